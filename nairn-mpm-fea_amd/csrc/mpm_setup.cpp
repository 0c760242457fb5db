// mpm_setup.cpp -- grid and particle generation, material constants, particle masses and the
// CFL time step: what NairnMPM does between reading the input and the first MPMStep.
#include <cmath>
#include <cstring>

#include "mpm_problem.h"

namespace mpmsetup {

// MPMReadHandler::grid() for a regular grid (Read_MPM/Generators.cpp:1675-1833): cell counts
// from cell sizes, then one extra element per side for GIMP/CPDI.
// lim receives the limits before the extra elements (the mesh limits mxmin..mzmax that
// ReadGridPoint's min and max keywords refer to, Generators.cpp:1760-1766).
static bool axis_cells(double &vmin, double &vmax, double cell, int &n, double *lim) {
    if (cell > 0.) {
        n = (int)((vmax - vmin) / cell);
        double newMax = vmin + n * cell;
        if (newMax < vmax) { vmax = newMax + cell; n++; }
    }
    if (n < 1) return false;
    lim[0] = vmin;
    lim[1] = vmax;
    double c = (vmax - vmin) / (double)n;
    n += 2;
    vmin -= c;
    vmax += c;
    return true;
}

bool build_grid(MpmProblem &p) {
    if (!p.haveGrid) { p.err = "no <Grid> in the input"; return false; }
    double xmn = p.gxmin, xmx = p.gxmax, ymn = p.gymin, ymx = p.gymax, zmn = p.gzmin, zmx = p.gzmax;
    int nx = p.nhoriz, ny = p.nvert, nz = p.ndepth;
    if (!axis_cells(xmn, xmx, p.cellx, nx, &p.meshLim[0]) || !axis_cells(ymn, ymx, p.celly, ny, &p.meshLim[2]) ||
        (p.dim == 3 && !axis_cells(zmn, zmx, p.cellz, nz, &p.meshLim[4]))) {
        p.err = "Number of grid elements in all direction must be >= 1.";
        return false;
    }
    MpmGridDesc &g = p.grid;
    memset(&g, 0, sizeof(g));
    g.analysis = p.analysis;
    g.nx = nx; g.ny = ny; g.nz = p.dim == 3 ? nz : 0;
    g.xmin = xmn; g.ymin = ymn; g.zmin = p.dim == 3 ? zmn : 0.;
    // MeshInfo::SetCartesian grid sizes and the xpts spacing (Generators.cpp:1804-1821)
    g.dx = (xmx - xmn) / (double)nx;
    g.dy = (ymx - ymn) / (double)ny;
    g.dz = p.dim == 3 ? (zmx - zmn) / (double)nz : 0.;
    g.thickness = p.thickness;
    long nelems = (long)nx * ny * (p.dim == 3 ? nz : 1);
    p.filled.assign(nelems, 0u);
    // MPMReadHandler::SetGIMPBorderAsHoles (Generators.cpp:1607-1619, called when </Grid> ends):
    // with GIMP or CPDI shape functions the elements of the outer layer take no particles
    // (MeshInfo::EdgeElement2D/3D, MeshInfo.cpp:171-199)
    if (p.run.shape != MPMGPU_SHAPE_LINEAR) {
        const long nzc = p.dim == 3 ? nz : 1;
        for (long e = 0; e < nelems; e++) {
            long i = e % nx, j = (e / nx) % ny, k = e / ((long)nx * ny);
            bool edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || (p.dim == 3 && (k == 0 || k == nzc - 1));
            if (edge) p.filled[e] = 0xffffffffu;
        }
    }
    p.gridBuilt = true;
    return true;
}

// Elastic::SetAnalysisProps + FillUnrotatedElasticProperties for an isotropic material
// (Common/Materials/Elastic.cpp:43-150,194-330; IsotropicMat.cpp:93-167)
static void isotropic_props(int analysis, double E, double nu, double G, double rho, MpmMaterial &o) {
    double e1 = E, e2 = E, e3 = E, v12 = nu, v13 = nu, v23 = nu, g12 = G, g13 = G, g23 = G;
    double v32 = v23 * e3 / e2;
    double v31 = v13 * e3 / e1;
    double v21 = v12 * e2 / e1;
    double xx, C11, C12, C13, C22, C23, C33, C44, C55, C66;
    memset(o.C, 0, sizeof(o.C));
    double rrho = 1. / rho;
    if (analysis == MPMGPU_PLANE_STRAIN) {
        xx = 1. - v13 * v31 - v23 * v32 - v12 * v21 - 2. * v12 * v23 * v31;
        C11 = e1 * (1. - v23 * v32) / xx;
        C12 = e2 * (v12 + v13 * v32) / xx;
        C22 = e2 * (1. - v13 * v31) / xx;
        C66 = g12;
        C13 = e3 * (v13 + v12 * v23) / xx;
        C23 = e3 * (v23 + v21 * v13) / xx;
        C33 = e3 * (1. - v21 * v12) / xx;
        o.C[1][1] = C11 * rrho;
        o.C[1][2] = C12 * rrho;
        o.C[2][2] = C22 * rrho;
        o.C[3][3] = C66 * rrho;
        o.C[4][1] = C13 * rrho;
        o.C[4][2] = C23 * rrho;
        o.C[4][4] = C33 * rrho;
        double S13 = -v13 / e1, S23 = -v23 / e2, S33 = 1. / e3;
        o.C[5][1] = S13 / S33;
        o.C[5][2] = S23 / S33;
        o.C[2][1] = o.C[1][2];
    } else {
        xx = 1. - v13 * v31 - v23 * v32 - v12 * v21 - 2. * v13 * v32 * v21;
        C11 = e1 * (1. - v23 * v32) / xx;
        C12 = e2 * (v12 + v13 * v32) / xx;
        C13 = e3 * (v13 + v12 * v23) / xx;
        C22 = e2 * (1. - v13 * v31) / xx;
        C23 = e3 * (v23 + v21 * v13) / xx;
        C33 = e3 * (1. - v21 * v12) / xx;
        C66 = g12; C44 = g23; C55 = g13;
        o.C[0][0] = C11 * rrho; o.C[0][1] = C12 * rrho; o.C[0][2] = C13 * rrho;
        o.C[1][0] = o.C[0][1]; o.C[1][1] = C22 * rrho; o.C[1][2] = C23 * rrho;
        o.C[2][0] = o.C[0][2]; o.C[2][1] = o.C[1][2]; o.C[2][2] = C33 * rrho;
        o.C[3][3] = C44 * rrho; o.C[4][4] = C55 * rrho; o.C[5][5] = C66 * rrho;
    }
    // the unspecific values kept for IsoPlasticity (IsoPlasticity.cpp:60-66 uses C66, C33)
    o.Gred = C66 / rho;
    o.Kred = C33 / rho - 4. * o.Gred / 3.;
}

bool material_verify(int analysis, const MaterialInput &in, MpmMaterial &o, std::string &err) {
    memset(&o, 0, sizeof(o));
    o.type = in.type;
    o.rho = in.rho * 0.001;                      // MaterialBaseMPM.cpp:56-58 legacy scaling
    o.ujOption = in.ujOption;
    if (in.pdamping > -1.e12) {                  // MaterialBase::SetDamping (MaterialBaseMPM.cpp:731-738)
        o.usePDamping = 1;
        o.pdamping = in.pdamping;
    }
    switch (in.type) {
        case MPMGPU_MAT_ISOTROPIC:
        case MPMGPU_MAT_ISOPLASTICITY: {
            // IsotropicMat::VerifyAndLoadProperties (Common/Materials/IsotropicMat.cpp:93-113)
            double E = in.E * 1.e6, G = in.G * 1.e6, nu = in.nu;
            if (!in.readG) G = E / (2. * (1. + nu));
            else if (!in.readE) E = 2. * G * (1. + nu);
            else if (!in.readNu) nu = E / (2. * G) - 1.;
            else { err = "E, nu, and G all specified. Only two allowed"; return false; }
            if (!(in.readE || in.readG) || !(in.readNu || (in.readE && in.readG))) {
                err = "A required material property is missing"; return false;
            }
            // legality of the elastic constants (Elastic::FillElasticProperties3D, Elastic.cpp:206-217);
            // isotropic: every v_ij v_ji = nu^2, and v21 v32 v13 = nu^3
            if (E < 0. || G < 0.) { err = "Modulus less than zero"; return false; }
            if (nu * nu > 1.) { err = "Illegal Poisson's ratio (v12*v21>1)"; return false; }
            if (nu * nu * nu > (1. - 3. * nu * nu) / 2.) {
                err = "Illegal Poisson's ratio (v21*v32*v13>(1-v12*v21-v31*v13-v23*v32)/2)"; return false;
            }
            if (o.rho <= 0.) { err = "Density less than or equal to zero"; return false; }
            o.E = E; o.nu = nu; o.G = G; o.K = E / (3. * (1 - 2 * nu));
            isotropic_props(analysis, E, nu, G, o.rho, o);
            // IsotropicMat::WaveSpeed (MoreIsotropicMat.cpp:388-396)
            o.waveSpeed = sqrt(2. * G * (1. - nu) / (o.rho * (1. - 2. * nu)));
            if (in.type == MPMGPU_MAT_ISOPLASTICITY) {
                // HardeningLawBase/LinearHardening::VerifyAndLoadProperties
                // (HardeningLawBase.cpp:58-65, LinearHardening.cpp:56-76)
                o.yield = in.yield * 1.e6;
                double yieldMin = in.yieldMin * 1.e6;
                o.yldred = o.yield / o.rho;
                o.yldredMin = yieldMin / o.rho;
                double beta = in.Khard, Ep = in.Ep >= 0. ? in.Ep * 1.e6 : -1.;
                if (Ep >= 0.) beta = Ep / o.yield;
                else Ep = beta * o.yield;
                o.Ep = Ep;
                o.Epred = o.yldred * beta;
                o.alphaMax = 1.e50;
                if (beta < 0.) o.alphaMax = ((o.yldredMin / o.yldred) - 1.) / beta;
            }
            return true;
        }
        case MPMGPU_MAT_NEOHOOKEAN: {
            // Neohookean::VerifyAndLoadProperties (Neohookean.cpp:88-143)
            double G = in.G >= 0. ? in.G * 1.e6 : -1., K = in.K >= 0. ? in.K * 1.e6 : -1.;
            double Lame = in.Lame >= 0. ? in.Lame * 1.e6 : -1.;
            double E = in.E >= 0. ? in.E * 1.e6 : -1., nu = in.nu;
            bool haveNu = in.readNu;
            if (G >= 0. && K >= 0.) {
                if (E >= 0. || haveNu || Lame >= 0.) { err = "Neohookean Hyperelastic material needs K and G, Lame and G, OR E and nu"; return false; }
                E = 9. * K * G / (3. * K + G);
                nu = (3. * K - 2. * G) / (6. * K + 2. * G);
                Lame = K - 2. * G / 3.;
            } else if (G >= 0. || Lame >= 0.) {
                if (E >= 0. || haveNu || K >= 0.) { err = "Neohookean Hyperelastic material needs K and G, Lame and G, OR E and nu"; return false; }
                K = Lame + 2. * G / 3.;
                E = 9. * K * G / (3. * K + G);
                nu = (3. * K - 2. * G) / (6. * K + 2. * G);
            } else if (E >= 0. && haveNu) {
                if (nu > 0.5 || nu < -1.) { err = "Neohookean Poisson's ratio out of range"; return false; }
                G = E / (2. * (1. + nu));
                if (nu < 0.5) { K = E / (3. * (1 - 2. * nu)); Lame = nu * E / (1. + nu) / (1. - 2. * nu); }
                else { err = "Neohookean Hyperelastic Poisson's ratio for solid material must be less than 1/2"; return false; }
            } else { err = "Neohookean Hyperelastic material needs K and G, Lame and G, OR E and nu"; return false; }
            if (o.rho <= 0.) { err = "Density less than or equal to zero"; return false; }
            o.E = E; o.nu = nu; o.G = G; o.K = K;
            o.Gsp = G / o.rho;
            o.Lamesp = Lame / o.rho;
            o.Ksp = o.Lamesp + 2. * o.Gsp / 3.;
            o.waveSpeed = sqrt((K + 4. * G / 3.) / o.rho);   // Neohookean::WaveSpeed (Neohookean.cpp:359)
            return true;
        }
        case MPMGPU_MAT_MOONEY: {
            // Mooney::VerifyAndLoadProperties (Mooney.cpp:99-150) and
            // HyperElastic::VerifyAndLoadProperties (HyperElastic.cpp:63-80)
            double G1 = in.readG ? in.G * 1.e6 : -1., G2 = in.G2 * 1.e6;
            double K = in.K >= 0. ? in.K * 1.e6 : -1.;
            double E = in.readE ? in.E * 1.e6 : -1., nu = in.readNu ? in.nu : -2.;
            if (G1 >= 0. && K >= 0.) {
                if (E >= 0. || nu >= -1.) { err = "Mooney-Rivlin Hyperelastic material needs K and G1 OR E and nu"; return false; }
                double Glow = G1 + G2;
                E = 9. * K * Glow / (3. * K + Glow);
                nu = (3. * K - 2. * Glow) / (6. * K + 2. * Glow);
            } else if (G1 >= 0. || K >= 0. || E < 0. || nu < -1.) {
                err = "Mooney-Rivlin Hyperelastic material needs K and G1 OR E and nu"; return false;
            } else {
                if (nu > 0.5) { err = "Mooney-Rivlin Hyperelastic Poisson's ratio must be between -1 and 1/2"; return false; }
                G1 = E / (2. * (1. + nu)) - G2;
                if (nu < 0.5) K = E / (3. * (1 - 2. * nu));
                else { err = "Mooney-Rivlin Hyperelastic Poisson's ratio for solid material must be less than 1/2"; return false; }
            }
            if (G2 < 0.) { err = "Mooney-Rivlin Hyperelastic material needs non-negative G2"; return false; }
            if (in.ujOption < 0 || in.ujOption > 2) { err = "Mooney-Rivlin dilational energy (UJOption) must be 0, 1, or 2"; return false; }
            if (o.rho <= 0.) { err = "Density less than or equal to zero"; return false; }
            if (K < 0.) { err = "Bulk modulus cannot be negative"; return false; }
            o.E = E; o.nu = nu; o.G = G1 + G2; o.K = K;
            o.G1sp = G1 / o.rho;
            o.G2sp = G2 / o.rho;
            o.Ksp = K / o.rho;
            o.waveSpeed = sqrt((K + 4. * (G1 + G2) / 3.) / o.rho);   // Mooney::WaveSpeed (Mooney.cpp:392-394)
            return true;
        }
        case MPMGPU_MAT_COULOMB:
            // CoulombFriction::VerifyAndLoadProperties (CoulombFriction.cpp:67-90)
            o.friction = in.friction;
            o.frictionStatic = in.friction < 0. ? -1. : in.frictionStatic;
            if (o.frictionStatic > 0. && o.frictionStatic < o.friction) {
                err = "The static coefficient of friction cannot be lower than the sliding coefficient of friction";
                return false;
            }
            return true;
        case MPMGPU_MAT_CONTACT:
            o.friction = 0.;
            o.frictionStatic = -1.;
            return true;
        default:
            err = "material type " + std::to_string(in.type) + " is not supported on this path";
            return false;
    }
}

bool add_material(MpmProblem &p, const MaterialInput &m) {
    p.matIn.push_back(m);
    return true;
}

// shape containment: BoxController (Common/Read_XML/BoxController.cpp:70-99),
// RectController (RectController.cpp:45-50), OvalController (OvalController.cpp:19-46),
// SphereController (Read_MPM/SphereController.cpp:21-41)
enum { SHAPE_BOX = 0, SHAPE_RECT = 1, SHAPE_OVAL = 2, SHAPE_SPHERE = 3 };

static bool contains(int kind, const double *e, int axis, const double *v) {
    double xmin = e[0], xmax = e[1], ymin = e[2], ymax = e[3], zmin = e[4], zmax = e[5];
    switch (kind) {
        case SHAPE_BOX: {
            if (axis == 0)
                return v[0] <= xmax && v[0] >= xmin && v[1] <= ymax && v[1] >= ymin && v[2] <= zmax && v[2] >= zmin;
            double xmid = (xmin + xmax) / 2., ymid = (ymin + ymax) / 2., zmid = (zmin + zmax) / 2.;
            double a2 = xmax - xmid; a2 *= a2;
            double b2 = ymax - ymid; b2 *= b2;
            double c2 = zmax - zmid; c2 *= c2;
            double R = 1.;          // coneRadius 1: a cylinder, (max-v + (v-min))/d
            if (axis == 1) {
                if (v[0] > xmax || v[0] < xmin) return false;
                double dx = xmax - xmin; R = (xmax - v[0] + 1. * (v[0] - xmin)) / dx;
                double dy = v[1] - ymid, dz = v[2] - zmid;
                return (dy * dy / b2 + dz * dz / c2) <= R * R;
            } else if (axis == 2) {
                if (v[1] > ymax || v[1] < ymin) return false;
                double dy = ymax - ymin; R = (ymax - v[1] + 1. * (v[1] - ymin)) / dy;
                double dx = v[0] - xmid, dz = v[2] - zmid;
                return (dx * dx / a2 + dz * dz / c2) <= R * R;
            } else {
                if (v[2] > zmax || v[2] < zmin) return false;
                double dz = zmax - zmin; R = (zmax - v[2] + 1. * (v[2] - zmin)) / dz;
                double dx = v[0] - xmid, dy = v[1] - ymid;
                return (dx * dx / a2 + dy * dy / b2) <= R * R;
            }
        }
        case SHAPE_RECT:
            return v[0] <= xmax && v[0] >= xmin && v[1] <= ymax && v[1] >= ymin;
        case SHAPE_OVAL: {
            double x0 = (xmin + xmax) / 2., y0 = (ymin + ymax) / 2.;
            double a = xmax - x0, b = ymax - y0;
            double x = v[0] - x0, y = v[1] - y0;
            return !((x * x / (a * a) + y * y / (b * b)) > 1.);
        }
        case SHAPE_SPHERE: {
            double x0 = (xmin + xmax) / 2., y0 = (ymin + ymax) / 2., z0 = (zmin + zmax) / 2.;
            double a = xmax - x0, b = ymax - y0, c = zmax - z0;
            return ((v[0] - x0) * (v[0] - x0) / a / a + (v[1] - y0) * (v[1] - y0) / b / b +
                    (v[2] - z0) * (v[2] - z0) / c / c) <= 1.;
        }
    }
    return false;
}

static void push_particle(MpmProblem &p, int mat1, int elem0, const double *pos, const double *vel,
                          const double *lp, double mass, double thick) {
    p.pthick.push_back(thick > 0. ? thick : p.thickness);
    for (int f = 0; f < MPF_NUM_FIELDS; f++) p.d[f].push_back(0.);
    for (int f = 0; f < MPI_NUM_FIELDS; f++) p.iv[f].push_back(0);
    size_t i = p.d[0].size() - 1;
    for (int c = 0; c < 3; c++) {
        p.d[MPF_POS_X + c][i] = pos[c];
        p.d[MPF_ORIG_X + c][i] = pos[c];
        p.d[MPF_VEL_X + c][i] = vel[c];
        p.d[MPF_LP_X + c][i] = lp[c];
    }
    p.d[MPF_MASS][i] = mass;               // <0: computed from rho and volume later
    p.iv[MPI_ELEM][i] = elem0;
    p.iv[MPI_MAT][i] = mat1 - 1;
    p.iv[MPI_ID][i] = (int32_t)i;
}

// MPMReadHandler::MPMPts (Read_MPM/Generators.cpp:1423-1545) with ElementBase::MPMPoints
// (Elements/MoreMPMElementBase.cpp:348-446): elements in order, points in element order, a point
// already taken by an earlier body is skipped (filled flags).
bool add_body(MpmProblem &p, int mat1, int kind, const double *ext, int axis, const double *vel,
              double thick) {
    if (!p.gridBuilt && !build_grid(p)) return false;
    const MpmGridDesc &g = p.grid;
    int n = p.ptsPerSide;
    int nz = p.dim == 3 ? n : 1;
    double xgap = 1. / ((double)n), ygap = xgap, zgap = 1. / ((double)nz);
    double lp[3] = {xgap, ygap, p.dim == 3 ? zgap : 1.};
    long nx = g.nx, ny = g.ny, nzc = p.dim == 3 ? g.nz : 1;
    // natural coordinates -1+(2j+1)/n, converted with the element's shape functions
    std::vector<double> mpos;
    for (int i = 0; i < nz; i++) {
        double zval = p.dim == 3 ? -1. + (2. * i + 1.) * zgap : 0.;
        for (int ii = 0; ii < n; ii++) {
            double yval = -1. + (2. * ii + 1.) * ygap;
            for (int j = 0; j < n; j++) {
                mpos.push_back(-1. + (2. * j + 1.) * xgap);
                mpos.push_back(yval);
                mpos.push_back(zval);
            }
        }
    }
    int numFound = (int)(mpos.size() / 3);
    for (long e = 0; e < nx * ny * nzc; e++) {
        long ei = e % nx, ej = (e / nx) % ny, ek = p.dim == 3 ? e / (nx * ny) : 0;
        double x0 = g.xmin + (double)ei * g.dx, x1 = g.xmin + (double)(ei + 1) * g.dx;
        double y0 = g.ymin + (double)ej * g.dy, y1 = g.ymin + (double)(ej + 1) * g.dy;
        double z0 = p.dim == 3 ? g.zmin + (double)ek * g.dz : 0., z1 = p.dim == 3 ? g.zmin + (double)(ek + 1) * g.dz : 0.;
        for (int k = 0; k < numFound; k++) {
            const double *xi = &mpos[3 * k];
            double pos[3] = {0., 0., 0.};
            if (p.dim == 3) {
                // EightNodeIsoparamBrick::ShapeFunction (EightNodeIsoparamBrick.cpp:87-108), nodes 0..7
                static const double xii[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
                static const double eti[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
                static const double zti[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};
                for (int q = 0; q < 8; q++) {
                    double f = 0.125 * (1. + xii[q] * xi[0]) * (1. + eti[q] * xi[1]) * (1. + zti[q] * xi[2]);
                    double nxp = xii[q] < 0 ? x0 : x1, nyp = eti[q] < 0 ? y0 : y1, nzp = zti[q] < 0 ? z0 : z1;
                    pos[0] += nxp * f; pos[1] += nyp * f; pos[2] += nzp * f;
                }
            } else {
                // FourNodeIsoparam::ShapeFunction (FourNodeIsoparam.cpp:189-214), nodes 0..3
                static const double xii[4] = {-1., 1., 1., -1.};
                static const double eti[4] = {-1., -1., 1., 1.};
                for (int q = 0; q < 4; q++) {
                    double f = 0.25 * (1. + xii[q] * xi[0]) * (1. + eti[q] * xi[1]);
                    double nxp = xii[q] < 0 ? x0 : x1, nyp = eti[q] < 0 ? y0 : y1;
                    pos[0] += nxp * f; pos[1] += nyp * f; pos[2] += 0. * f;
                }
            }
            uint32_t flag = 1u << k;
            if (p.filled[e] & flag) continue;
            if (!contains(kind, ext, axis, pos)) continue;
            push_particle(p, mat1, (int)e, pos, vel, lp, -1., thick);
            p.filled[e] |= flag;
        }
    }
    return true;
}

// explicit <mp> particle (Read_MPM/MPMReadHandler.cpp:1225-1290); its element is found from
// the position (MeshInfo::FindElementFromPoint, MeshInfo.cpp:593-640)
bool add_point(MpmProblem &p, int mat1, const double *pos, const double *vel, const double *lp,
               double thick, double mass) {
    if (!p.gridBuilt && !build_grid(p)) return false;
    const MpmGridDesc &g = p.grid;
    long col = (long)(int)((pos[0] - g.xmin) / g.dx);
    long row = (long)(int)((pos[1] - g.ymin) / g.dy);
    long zr = p.dim == 3 ? (long)(int)((pos[2] - g.zmin) / g.dz) : 0;
    if (col < 0 || col >= g.nx || row < 0 || row >= g.ny || (p.dim == 3 && (zr < 0 || zr >= g.nz))) {
        p.err = "A material point in an <mp> command is not within the grid.";
        return false;
    }
    long e = p.dim == 3 ? ((long)g.nx * (zr * g.ny + row) + col) : row * g.nx + col;
    double l[3] = {lp[0], lp[1], p.dim == 3 ? lp[2] : 1.};
    push_particle(p, mat1, (int)e, pos, vel, l, mass, thick);
    return true;
}

// NairnMPM::PreliminaryParticleCalcs (NairnMPM.cpp:589-730) and CFLTimeStep (:1207-1240)
bool finalize(MpmProblem &p) {
    if (p.finalized) return true;
    if (!p.gridBuilt && !build_grid(p)) return false;
    p.mats.clear();
    for (auto &mi : p.matIn) {
        MpmMaterial m;
        if (!material_verify(p.analysis, mi, m, p.err)) return false;
        p.mats.push_back(m);
    }
    // velocity fields in order of first use (MaterialBase::SetField, MaterialBaseMPM.cpp:615-680,
    // called from NairnMPM.cpp:617); one field turns multimaterial mode off (NairnMPM.cpp:739-744)
    {
        int64_t n = p.nparticles();
        std::vector<int> field(p.mats.size(), -1);
        int nf = 0;
        for (int64_t i = 0; i < n; i++) {
            int mat = p.iv[MPI_MAT][i];
            if (mat < 0 || mat >= (int)p.mats.size()) { p.err = "Material point with an undefined material type"; return false; }
            if (field[mat] < 0) field[mat] = nf++;
        }
        p.run.multiMaterial = p.mmRequested && nf >= 2 ? 1 : 0;
        for (size_t m = 0; m < p.mats.size(); m++)
            p.mats[m].field = field[m] < 0 ? -1 : (p.run.multiMaterial ? field[m] : 0);
        p.run.numFields = p.run.multiMaterial ? nf : 1;
        p.run.contactLaw = -1;
        if (p.run.multiMaterial) {
            // MeshInfo::MaterialOutput (MeshInfo.cpp:1126-1200): logistic regression is the default
            if (!p.normalsGiven) p.run.normalsMethod = 6;
            if (p.run.normalsMethod == 4) {
                // MeshInfo::SetContactNormal (MeshInfo.cpp:1211-1227)
                const double pi = 3.141592653589793;
                double sinp = 1.0;
                p.run.contactNormal[2] = 0.;
                if (p.dim == 3) {
                    double polar = pi * p.polarAngle / 180.;
                    p.run.contactNormal[2] = cos(polar);
                    sinp = sin(polar);
                }
                double az = pi * p.azimuthAngle / 180.;
                p.run.contactNormal[0] = cos(az) * sinp;
                p.run.contactNormal[1] = sin(az) * sinp;
            }
            int law = -1;
            if (!p.contactLawName.empty()) {
                for (size_t m = 0; m < p.matIn.size(); m++)
                    if (p.matIn[m].name == p.contactLawName) law = (int)m;
            } else if (p.contactLawNum > 0) {
                law = p.contactLawNum - 1;
            } else {
                // automatic law: old-style coefficient, or frictionless when no <Friction> was
                // given (MPMReadHandler.cpp:1780-1783)
                MaterialInput a;
                if (p.contactLawNum < 0) a = p.autoLaws[-p.contactLawNum - 1];
                else { a.type = MPMGPU_MAT_COULOMB; a.friction = 0.; a.name = "automatic contact law"; }
                MpmMaterial am;
                if (!material_verify(p.analysis, a, am, p.err)) return false;
                am.field = -1;
                p.mats.push_back(am);
                law = (int)p.mats.size() - 1;
            }
            if (law < 0 || law >= (int)p.mats.size() ||
                (p.mats[law].type != MPMGPU_MAT_COULOMB && p.mats[law].type != MPMGPU_MAT_CONTACT)) {
                p.err = "Multimaterial mode must select a default contact law";
                return false;
            }
            if ((int)p.mats.size() > MPMGPU_MAX_MATERIALS) { p.err = "too many materials"; return false; }
            p.run.contactLaw = law;
        }
    }
    const MpmGridDesc &g = p.grid;
    double dcell = g.dx;                          // MeshInfo::GetMinCellDimension
    if (g.dy < dcell) dcell = g.dy;
    if (p.dim == 3 && g.dz < dcell) dcell = g.dz;
    double timeStepMinMechanics = 1.e30;
    int64_t n = p.nparticles();
    for (int64_t i = 0; i < n; i++) {
        int mat = p.iv[MPI_MAT][i];
        if (mat < 0 || mat >= (int)p.mats.size()) { p.err = "Material point with an undefined material type"; return false; }
        const MpmMaterial &m = p.mats[mat];
        if (m.type == MPMGPU_MAT_COULOMB || m.type == MPMGPU_MAT_CONTACT) {
            p.err = "Material point with contact-law material"; return false;
        }
        // particle semi size and volume (MPMBase.cpp:519-523, MatPoint3D.cpp:227-233) from the
        // extents of the particle's element (ElementBase::GetDeltaBox, ElementBase.cpp:115)
        long e = p.iv[MPI_ELEM][i];
        long ei = e % g.nx, ej = (e / g.nx) % g.ny, ek = p.dim == 3 ? e / ((long)g.nx * g.ny) : 0;
        double ddx = (g.xmin + (double)(ei + 1) * g.dx) - (g.xmin + (double)ei * g.dx);
        double ddy = (g.ymin + (double)(ej + 1) * g.dy) - (g.ymin + (double)ej * g.dy);
        double psx = ddx * (0.5 * p.d[MPF_LP_X][i]);
        double psy = ddy * (0.5 * p.d[MPF_LP_Y][i]);
        double volume;
        if (p.dim == 3) {
            double ddz = (g.zmin + (double)(ek + 1) * g.dz) - (g.zmin + (double)ek * g.dz);
            double psz = ddz * (0.5 * p.d[MPF_LP_Z][i]);
            volume = 8. * psx * psy * psz;
        } else volume = 4. * psx * psy * p.pthick[i];
        if (p.d[MPF_MASS][i] < 0.) p.d[MPF_MASS][i] = m.rho * volume;     // InitializeMass
        // SetInitialParticleState / InitHistoryData
        if (m.type == MPMGPU_MAT_NEOHOOKEAN || m.type == MPMGPU_MAT_MOONEY) {
            p.d[MPF_ALT_XX][i] = p.d[MPF_ALT_YY][i] = p.d[MPF_ALT_ZZ][i] = 1.;
            p.d[MPF_HIST0][i] = 1.; p.d[MPF_HIST1][i] = 1.;
        }
        double tst = dcell / m.waveSpeed;
        if (tst < timeStepMinMechanics) timeStepMinMechanics = tst;
    }
    double timeStepMin = p.cfl * timeStepMinMechanics;
    double dt = p.timestepIn;
    if (timeStepMin < dt) dt = timeStepMin;
    p.run.timestep = dt;
    p.run.strainFirst = 0.5 * dt;               // fractionUSF = 0.5
    p.run.strainLast = dt - p.run.strainFirst;
    p.run.mpmMethod = 2;
    // NairnMPM::CreateWarnings (NairnMPM.cpp:814-848): LeaveLimit 0 means 1% of the particles;
    // then > 1 pushes leaving particles back and aborts once that many have left, 1 pushes back
    // and never aborts (a warning limit of -1), and <= 0 deletes them into the reservoir
    {
        int lim = p.leaveLimit == 0 ? (int)(n / 100) : p.leaveLimit;
        if (lim > 1) { p.run.deleteLeaving = 0; p.run.maxLeaving = lim; }
        else if (lim == 1) { p.run.deleteLeaving = 0; p.run.maxLeaving = 0; }
        else { p.run.deleteLeaving = 1; p.run.maxLeaving = -lim; }
        // DeleteLimit > 1 deletes particles whose position is NaN (into the reservoir); otherwise
        // the first NaN position stops the run
        if (p.deleteLimit > 1) {
            p.err = "<DeleteLimit> above 1 deletes particles into the reservoir, which is not on this path";
            return false;
        }
    }
    // PeriodicXPIC::Finalize and Initialize (PeriodicXPIC.cpp:160-252): the order, FMPM or XPIC,
    // and the mechanics schedule; steps override times, a CFL factor sets the time from the
    // mechanics time step
    if (p.xpicTask) {
        if (p.xpicOrderIn < 1 && p.fmpmOrderIn < 1) {
            p.err = "PeriodicXPIC custom task must set XPIC or FMPM order to 1 or higher";
            return false;
        }
        if (p.xpicOrderIn >= 1 && p.fmpmOrderIn >= 1) {
            p.err = "PeriodicXPIC custom task must set XPIC or FMPM order, but not both";
            return false;
        }
        p.run.xpicFMPM = p.fmpmOrderIn > 0 ? 1 : 0;
        p.run.xpicOrder = p.fmpmOrderIn > 0 ? p.fmpmOrderIn : p.xpicOrderIn;
        double ptime = p.xpicTimeIn, pcfl = p.xpicCFLIn;
        int psteps = p.xpicStepsIn;
        if (psteps > 0) ptime = pcfl = -1.;
        else if (!(ptime > 0. || pcfl > 0.)) {
            p.err = "PeriodicXPIC custom task did not activate any XPIC or FMPM calculations";
            return false;
        }
        if (pcfl > 0.) ptime = pcfl * timeStepMinMechanics;
        p.run.xpicPeriodSteps = psteps > 0 ? psteps : 0;
        p.run.xpicPeriodTime = psteps > 0 ? 0. : ptime;
        if (p.run.multiMaterial) {
            p.err = "PeriodicXPIC in multimaterial mode (incremental contact) is not on this path";
            return false;
        }
    }
    p.finalized = true;
    return true;
}

}  // namespace mpmsetup
