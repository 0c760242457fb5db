// mpm_kernels.hip -- gfx950 kernels of NairnMPM's explicit step (USAVG+, FLIP, uGIMP).
//
// Layout: particles stay in structure-of-arrays form; every step they are binned by cell with a
// stable counting sort into tiles of TileGeom<DIM> cells (tile-major cell keys), which gives a
// sorted slot -> particle map (order[]). Particle-to-grid sums never use floating-point atomics:
// a tile kernel stages its particles' per-axis GIMP factors in LDS; then one wave handles one
// cell at a time with each lane owning one of the 64 candidate node offsets, sums the cell's
// particles in sorted order in registers and adds the sum to the wave's own LDS copy of the
// tile's node footprint (wave_cells). The copies are added in wave order and the footprint sums
// of each tile go to a partial buffer that a node kernel adds in fixed tile order. Results are
// therefore deterministic; they differ from the reference's single-patch sums by summation
// order only.
//
// Reference tasks (paths relative to the reference root, NairnMPM/src/NairnMPM_Class/):
//   k_p2g_mass    MassAndMomentumTask.cpp:48-127      (+ InitializationTask xi, :43-85)
//   k_node_mass   PostExtrapolationTask.cpp:46-165 + GridValueCalculation (UpdateStrainsFirstTask.cpp:120)
//   k_usf_force   UpdateStrainsFirstTask.cpp:101-168 + GridForcesTask.cpp:40-153
//   k_node_force  PostForcesTask.cpp:43-111 + UpdateMomentaTask.cpp:44-64 + GridValueCalculation
//   k_g2p_p2g     UpdateParticlesTask.cpp:44-298 + UpdateStrainsLastContactTask.cpp:60-142
//   k_node_mom2   UpdateStrainsLastContactTask.cpp:139-149 + GridValueCalculation
//   k_usl_reset   UpdateStrainsFirstTask::FullStrainUpdate(last) + ResetElementsTask.cpp:44-265
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "mpm_device.h"
#include "mpm_kernels.h"

#define NT 256

// The kernels of one tile shape: this file builds shape 0, and mpm_kernels_s1.hip builds shape 1
// by defining MPM_TILE_SHAPE before including it. Each goes into its own namespace.
#ifndef MPM_TILE_SHAPE
#define MPM_TILE_SHAPE 0
#endif
#if MPM_TILE_SHAPE == 0
#define MPM_KNS mpmk_s0
#else
#define MPM_KNS mpmk_s1
#endif
namespace MPM_KNS {

template <int DIM> using TileGeom = TileShape<DIM, MPM_TILE_SHAPE>;
// waves per SIMD the uGIMP particle-update scatters are built for: shape 1's larger footprint
// leaves LDS for two workgroups per CU
#define G2P_WAVES (MPM_TILE_SHAPE == 0 ? 3 : 2)

template <int DIM>
__device__ __forceinline__ void tile_origin(const KGrid &g, int t, int &ox, int &oy, int &oz) {
    using G = TileGeom<DIM>;
    int tx = t % g.ntx;
    int r = t / g.ntx;
    int ty = r % g.nty;
    int tz = r / g.nty;
    ox = tx * G::TX; oy = ty * G::TY; oz = tz * G::TZ;
}

// bin key: field-major, then tile, then cell within the tile, so one (field, tile) pair owns a
// contiguous run of sorted particles
template <int DIM>
__device__ __forceinline__ int cell_key(const KGrid &g, int fld, int ci, int cj, int ck) {
    using G = TileGeom<DIM>;
    int tx = ci / G::TX, ty = cj / G::TY, tz = ck / G::TZ;
    int t = (tz * g.nty + ty) * g.ntx + tx;
    int l = ((ck - tz * G::TZ) * G::TY + (cj - ty * G::TY)) * G::TX + (ci - tx * G::TX);
    return (fld * g.tilesTotal + g.tileBase + t) * G::NC + l;
}

// node arrays of one velocity field
__device__ __forceinline__ KNodes field_view(const KGrid &g, const KNodes &N, int f) {
    KNodes v = N;
    long long nn = g.nstride, b = g.nodeBase;
    v.mass = N.mass + f * nn + b;
    v.pk = N.pk + 3 * f * nn + b;
    v.vk = N.vk + 3 * f * nn + b;
    v.ftot = N.ftot + 3 * f * nn + b;
    if (N.vprev) v.vprev = N.vprev + 3 * f * nn + b;
    v.cnt = N.cnt + f * nn + b;
    if (N.cvol) {
        v.cvol = N.cvol + f * nn + b;
        v.cdisp = N.cdisp + 3 * f * nn + b;
        v.cgrad = N.cgrad + 3 * f * nn + b;
    }
    return v;
}

// grid of the instance that owns context tile tg (the kernel's grid for one instance)
__device__ __forceinline__ KGrid tile_grid(const KGrid &g0, int tg) {
    if (!g0.inst) return g0;
    return g0.inst[g0.tileInst[tg]];
}

// grid of the instance that owns context node ng (node ranges are padded to NT per instance)
__device__ __forceinline__ KGrid node_grid(const KGrid &g0, long long ng) {
    if (!g0.inst) return g0;
    return g0.inst[g0.nodeBlockInst[ng / 256]];
}

__device__ __forceinline__ void elem_ijk(const KGrid &g, int e, int &i, int &j, int &k) {
    i = e % g.nx;
    int r = e / g.nx;
    j = r % g.ny;
    k = r / g.ny;
}

// natural coordinates in the particle's element (EightNodeIsoparamBrick.cpp:276-281,
// FourNodeIsoparam.cpp:351-356), node coordinates xmin + i*dx (Generators.cpp:1823-1833)
template <int DIM>
__device__ __forceinline__ void particle_xi(const KGrid &g, int ei, int ej, int ek, double x, double y,
                                            double z, double xi[3], double invd[3]) {
    double xmn = g.xmin + (double)ei * g.dx, xmx = g.xmin + (double)(ei + 1) * g.dx;
    double ymn = g.ymin + (double)ej * g.dy, ymx = g.ymin + (double)(ej + 1) * g.dy;
    xi[0] = (2. * x - xmn - xmx) / (xmx - xmn);
    xi[1] = (2. * y - ymn - ymx) / (ymx - ymn);
    invd[0] = 2.0 / (xmx - xmn);
    invd[1] = 2.0 / (ymx - ymn);
    if (DIM == 3) {
        double zmn = g.zmin + (double)ek * g.dz, zmx = g.zmin + (double)(ek + 1) * g.dz;
        xi[2] = (2. * z - zmn - zmx) / (zmx - zmn);
        invd[2] = 2.0 / (zmx - zmn);
    } else {
        xi[2] = 0.;
        invd[2] = 0.;
    }
}

// per-axis factors of one particle: uGIMP, or classic linear (the same kernels serve both; the
// choice is uniform over a launch)
template <int DIM>
__device__ __forceinline__ void particle_factors(const KGrid &g, const KPart &P, int p, const double xi[3],
                                                 double S[3][4], double dS[3][4]) {
    if (g.shape == MPMGPU_SHAPE_LINEAR) {
#pragma unroll
        for (int a = 0; a < DIM; a++) linear_axis(xi[a], S[a], dS[a]);
        return;
    }
    double lpx, lpy, lpz;
    if (g.uniformLp) { lpx = g.lp[0]; lpy = g.lp[1]; lpz = g.lp[2]; }
    else { lpx = P.f[MPF_LP_X][p]; lpy = P.f[MPF_LP_Y][p]; lpz = P.f[MPF_LP_Z][p]; }
    if (DIM == 3) {
        gimp_axis3(xi[0], lpx, 1. / (4. * lpx), S[0], dS[0]);
        gimp_axis3(xi[1], lpy, 1. / (4. * lpy), S[1], dS[1]);
        gimp_axis3(xi[2], lpz, 1. / (4. * lpy), S[2], dS[2]);   // EightNodeIsoparamBrick.cpp:303
    } else {
        gimp_axis2(xi[0], lpx, S[0], dS[0]);
        gimp_axis2(xi[1], lpy, S[1], dS[1]);
    }
}

// ---------------------------------------------------------------------------------------
// lCPDI (MatPoint3D::GetCPDINodesAndWeights, MatPoint3D.cpp:538-627; MatPoint2D.cpp:519-571;
// ElementBase::GetCPDIFunctions, MoreMPMElementBase.cpp:581-656). The semi-side vectors
// r1 = F(psz.x,0,0), r2, r3 are taken once per step, from the start-of-step F, by the first
// scatter and kept in P.cpr; every pass rebuilds the corners from them and the start position.
// A corner's element must be the particle's element or a neighbour of it (the 64-node candidate
// block); otherwise, or when a corner leaves the grid, the error counter is raised and the corner
// is clamped so that every access stays in bounds.
struct CpdiCorners {
    double xi[8][3];           // natural coordinates of each corner in its element
    int off[8][3];             // element of the corner relative to the particle's element (-1..1)
};

// MeshInfo::FindElementFromPoint on one axis of a regular grid (MeshInfo.cpp:597-625)
__device__ __forceinline__ int find_cell(double x, double xmin, double d, int n, bool &ok) {
    int c = (int)((x - xmin) / d);
    if (c < 0 || c >= n) {
        if (x == xmin + n * d) c = n - 1;
        else { ok = false; c = c < 0 ? 0 : n - 1; }
    }
    return c;
}

// semi-side vectors from the particle's F and size (GetSemiSideVectors, GetParticleSize)
template <int DIM>
__device__ __forceinline__ void cpdi_semi_sides(const KGrid &g, const KPart &P, int p, int ei, int ej, int ek,
                                                double r[9]) {
    double lpx, lpy, lpz;
    if (g.uniformLp) { lpx = g.lp[0]; lpy = g.lp[1]; lpz = g.lp[2]; }
    else { lpx = P.f[MPF_LP_X][p]; lpy = P.f[MPF_LP_Y][p]; lpz = P.f[MPF_LP_Z][p]; }
    double psx = ((g.xmin + (double)(ei + 1) * g.dx) - (g.xmin + (double)ei * g.dx)) * (0.5 * lpx);
    double psy = ((g.ymin + (double)(ej + 1) * g.dy) - (g.ymin + (double)ej * g.dy)) * (0.5 * lpy);
    double ep[6] = {P.f[MPF_EP_XX][p], P.f[MPF_EP_YY][p], P.f[MPF_EP_ZZ][p], 0., 0., P.f[MPF_EP_XY][p]};
    double wr[3] = {0., 0., P.f[MPF_WR_XY][p]};
    if (DIM == 3) {
        ep[3] = P.f[MPF_EP_YZ][p]; ep[4] = P.f[MPF_EP_XZ][p];
        wr[0] = P.f[MPF_WR_YZ][p]; wr[1] = P.f[MPF_WR_XZ][p];
    }
    double F[3][3];
    get_F3(ep, wr, F);
    if (DIM == 3) {
        double psz = ((g.zmin + (double)(ek + 1) * g.dz) - (g.zmin + (double)ek * g.dz)) * (0.5 * lpz);
        r[0] = F[0][0] * psx; r[1] = F[1][0] * psx; r[2] = F[2][0] * psx;
        r[3] = F[0][1] * psy; r[4] = F[1][1] * psy; r[5] = F[2][1] * psy;
        r[6] = F[0][2] * psz; r[7] = F[1][2] * psz; r[8] = F[2][2] * psz;
    } else {
        r[0] = F[0][0] * psx; r[1] = F[1][0] * psx; r[2] = 0.;
        r[3] = F[0][1] * psy; r[4] = F[1][1] * psy; r[5] = 0.;
        r[6] = r[7] = r[8] = 0.;
    }
}

// linear function of element node (jx,jy,jz) at natural coordinates x (EightNodeIsoparamBrick.cpp:87-108,
// FourNodeIsoparam.cpp:189-214): 0.125 (1 +- xi)(1 +- eta)(1 +- zeta)
template <int DIM>
__device__ __forceinline__ double cpdi_cfn(const double x[3], int jx, int jy, int jz) {
    double t1 = 1. + (jx ? 1. : -1.) * x[0];
    double t2 = 1. + (jy ? 1. : -1.) * x[1];
    if (DIM == 3) {
        double t3 = 1. + (jz ? 1. : -1.) * x[2];
        return 0.125 * t1 * t2 * t3;
    }
    return 0.25 * t1 * t2;
}


template <int DIM>
__device__ __forceinline__ int foot_slot(int lx, int ly, int lz) {
    using G = TileGeom<DIM>;
    if (DIM == 3) return ((lz + 1) * G::FY + (ly + 1)) * G::FX + (lx + 1);
    return (ly + 1) * G::FX + (lx + 1);
}

// global node index of footprint slot s of a tile with cell origin (ox,oy,oz); -1 outside grid
template <int DIM>
__device__ __forceinline__ long long foot_node(const KGrid &g, int s, int ox, int oy, int oz) {
    using G = TileGeom<DIM>;
    int fx = s % G::FX;
    int fy = (s / G::FX) % G::FY;
    int fz = DIM == 3 ? s / (G::FX * G::FY) : 1;
    int i = ox + fx - 1, j = oy + fy - 1, k = DIM == 3 ? oz + fz - 1 : 0;
    if (i < 0 || i >= g.nnx || j < 0 || j >= g.nny || k < 0 || k >= g.nnz) return -1;
    return ((long long)k * g.nny + j) * g.nnx + i;
}

// corner c of a particle (the body of cpdi_corners for one corner); false if the corner is unusable
template <int DIM>
__device__ __forceinline__ bool cpdi_corner(const KGrid &g, const double pos[3], const double r[9], int ei, int ej,
                                            int ek, int c, double xi[3], int off[3]) {
    bool ok = true;
    const int q = c & 3;
    const double s1 = (q == 1 || q == 2) ? 1. : -1.;
    const double s2 = q >= 2 ? 1. : -1.;
    const double s3 = c >= 4 ? 1. : -1.;
    double cx[3];
    if (DIM == 3) {
#pragma unroll
        for (int a = 0; a < 3; a++) cx[a] = pos[a] + s1 * r[a] + s2 * r[3 + a] + s3 * r[6 + a];
    } else {
        cx[0] = pos[0] + s1 * r[0] + s2 * r[3];
        cx[1] = pos[1] + s1 * r[1] + s2 * r[4];
        cx[2] = 0.;
    }
    const int ci = find_cell(cx[0], g.xmin, g.dx, g.nx, ok);
    const int cj = find_cell(cx[1], g.ymin, g.dy, g.ny, ok);
    const int ck = DIM == 3 ? find_cell(cx[2], g.zmin, g.dz, g.nz, ok) : 0;
    off[0] = ci - ei; off[1] = cj - ej; off[2] = DIM == 3 ? ck - ek : 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (off[a] < -1 || off[a] > 1) { ok = false; off[a] = off[a] < -1 ? -1 : 1; }
    }
    double invd[3];
    particle_xi<DIM>(g, ei + off[0], ej + off[1], ek + off[2], cx[0], cx[1], cx[2], xi, invd);
    return ok;
}

// corners of particle p for the step: from its saved semi-side vectors and start position
template <int DIM>
__device__ __forceinline__ bool cpdi_particle(const KGrid &g, const KPart &P, int p, int ei, int ej, int ek,
                                              CpdiCorners &cc) {
    double r[9];
#pragma unroll
    for (int a = 0; a < 9; a++) r[a] = P.cpr[a][p];
    double pos[3] = {P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p], DIM == 3 ? P.f[MPF_POS_Z][p] : 0.};
    bool ok = true;
    constexpr int NCOR = DIM == 3 ? 8 : 4;
#pragma unroll
    for (int c = 0; c < NCOR; c++) ok &= cpdi_corner<DIM>(g, pos, r, ei, ej, ek, c, cc.xi[c], cc.off[c]);
    return ok;
}

// gradient weight of corner c as the reference writes it out per corner (MatPoint3D.cpp:571-594,
// MatPoint2D.cpp:551-566): the same six products summed in the same order, each term's sign taken
// from the corner's signs s1, s2, s3 (the scatters use these; the gathers use cpdi_wg)
template <int DIM>
__device__ __forceinline__ void cpdi_wg_listed(const double r[9], int c, double w[3]) {
    const int q = c & 3;
    const double s1 = (q == 1 || q == 2) ? 1. : -1.;
    const double s2 = q >= 2 ? 1. : -1.;
    if (DIM == 3) {
        const double s3 = c >= 4 ? 1. : -1.;
        const double x1 = r[0], y1 = r[1], z1 = r[2], x2 = r[3], y2 = r[4], z2 = r[5];
        const double x3 = r[6], y3 = r[7], z3 = r[8];
        double Vp = 8. * (x1 * (y2 * z3 - z2 * y3) + y1 * (z2 * x3 - x2 * z3) + z1 * (x2 * y3 - y2 * x3));
        Vp = 1. / Vp;
        w[0] = ((((-s3 * (z1 * y2) + s3 * (y1 * z2)) + s2 * (z1 * y3)) - s1 * (z2 * y3)) - s2 * (y1 * z3) + s1 * (y2 * z3)) * Vp;
        w[1] = ((((s3 * (z1 * x2) - s3 * (x1 * z2)) - s2 * (z1 * x3)) + s1 * (z2 * x3)) + s2 * (x1 * z3) - s1 * (x2 * z3)) * Vp;
        w[2] = ((((-s3 * (y1 * x2) + s3 * (x1 * y2)) + s2 * (y1 * x3)) - s1 * (y2 * x3)) - s2 * (x1 * y3) + s1 * (x2 * y3)) * Vp;
    } else {
        double Ap = 4. * (r[0] * r[4] - r[1] * r[3]);
        Ap = 1. / Ap;
        w[0] = (-s2 * r[1] + s1 * r[4]) * Ap;
        w[1] = (s2 * r[0] - s1 * r[3]) * Ap;
        w[2] = 0.;
    }
}

// Corner tables of the lCPDI scatters. A scatter lane owns candidate node offset o1 - 1 per axis
// (o1 = 0..3); corner c's element holds the nodes at offsets off and off + 1, where the corner's
// linear factors are 1 - xi and 1 + xi. The table of an axis holds those two factors at o1 = off + 1
// and off + 2 and zeros elsewhere, so a lane reads the corner's function for its node with one load
// per axis, zero for nodes outside the element. Built by one thread per (particle, corner).
template <int DIM>
__device__ __forceinline__ void cpdi_table(const double xi[3], const int off[3], double (*T)[4]) {
#pragma unroll
    for (int a = 0; a < DIM; a++)
#pragma unroll
        for (int o = 0; o < 4; o++)
            T[a][o] = o == off[a] + 1 ? 1. - xi[a] : (o == off[a] + 2 ? 1. + xi[a] : 0.);
}

// Merged weight (and gradient) of the lane's node over the corners of one staged particle, summed
// in corner order; corners whose element lacks the node or whose function is below 1e-15 add
// nothing (GetCPDIFunctions, MoreMPMElementBase.cpp:581-656). The corner function 1/8 t1 t2 t3
// (1/4 t1 t2 in 2D) is formed without its power-of-two factor, which is exact to move: the
// threshold is scaled by 8 (4), the weight sum by 1/8 (1/4) once at the end, and the staged
// gradient weights carry it (cpdi_wg_staged). The weight is then the same to the last bit.
template <int DIM, bool GRAD>
__device__ __forceinline__ void cpdi_table_node(const double (*T)[DIM][4], const double (*wg)[3], int o1x, int o1y,
                                                int o1z, double &w, double gr[3]) {
    constexpr int NCOR = DIM == 3 ? 8 : 4;
    constexpr double thr = (DIM == 3 ? 8. : 4.) * 1e-15;
    constexpr double wscale = DIM == 3 ? 0.125 * 0.125 : 0.25 * 0.25;   // ws times the corner factor
    double sum = 0.;
    if (GRAD) gr[0] = gr[1] = gr[2] = 0.;
#pragma unroll
    for (int c = 0; c < NCOR; c++) {
        double f = DIM == 3 ? T[c][0][o1x] * T[c][1][o1y] * T[c][2][o1z] : T[c][0][o1x] * T[c][1][o1y];
        f = f < thr ? 0. : f;
        sum = c == 0 ? f : sum + f;
        if (GRAD) {
            gr[0] = fma(wg[c][0], f, gr[0]);
            gr[1] = fma(wg[c][1], f, gr[1]);
            if (DIM == 3) gr[2] = fma(wg[c][2], f, gr[2]);
        }
    }
    w = wscale * sum;
}

// gradient weight of corner c as the scatters stage it: times the corner function's factor
template <int DIM>
__device__ __forceinline__ void cpdi_wg_staged(const double r[9], int c, double w[3]) {
    cpdi_wg_listed<DIM>(r, c, w);
#pragma unroll
    for (int a = 0; a < 3; a++) w[a] *= DIM == 3 ? 0.125 : 0.25;
}

// ---------------------------------------------------------------------------------------
// binning: key + histogram, exclusive scan, scatter, per-cell sort (stable by particle index)

// count[key] += 1 for every active lane, with one atomic per distinct key of the wave
__device__ __forceinline__ void wave_count(int *count, int key) {
    const int lane = (int)__lane_id();
    unsigned long long active = __ballot(1);
    while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const int lkey = __shfl(key, leader);
        const unsigned long long same = __ballot(key == lkey) & active;
        if (lane == leader) atomicAdd(&count[lkey], __popcll(same));
        active &= ~same;
    }
}

template <int DIM>
__global__ void k_bin_keys(KGrid g0, KPart P, KBins B) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    if (B.stop[0]) { B.key[p] = 0; return; }   // input refused on the device: nothing is binned
    const KGrid g = g0.inst ? g0.inst[g0.matInst[P.mat[p]]] : g0;
    int i, j, k;
    elem_ijk(g, P.elem[p], i, j, k);
    // fieldOf covers the materials of one multimaterial context; batch instances are single-field
    // and their material numbers run past its end
    const int fld = g.nfields > 1 ? g.fieldOf[P.mat[p]] : 0;
    int key = cell_key<DIM>(g, fld, i, j, DIM == 3 ? k : 0);
    B.key[p] = key;
    wave_count(B.count, key);
}

#define SCAN_PER_THREAD 4
#define SCAN_BLOCK (NT * SCAN_PER_THREAD)

__device__ __forceinline__ int block_exclusive_scan(int v, int *s_tmp, int &total) {
    // Hillis-Steele over the block in LDS
    int tid = threadIdx.x;
    s_tmp[tid] = v;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        int add = tid >= off ? s_tmp[tid - off] : 0;
        __syncthreads();
        s_tmp[tid] += add;
        __syncthreads();
    }
    total = s_tmp[NT - 1];
    int incl = s_tmp[tid];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(NT) void k_scan_blocks(const int *in, int *out, int *blockSums, int n) {
    __shared__ int s_tmp[NT];
    int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER_THREAD;
    int v[SCAN_PER_THREAD];
    int sum = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        sum += v[i];
    }
    int total;
    int ex = block_exclusive_scan(sum, s_tmp, total);
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD; i++) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
    if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

__global__ __launch_bounds__(NT) void k_scan_sums(int *blockSums, int nb, int *ntiles) {
    __shared__ int s_tmp[NT];
    if (threadIdx.x == 0) ntiles[0] = 0;      // the tile list is rebuilt after the scan
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += NT) {
        int i = b0 + threadIdx.x;
        int v = i < nb ? blockSums[i] : 0;
        int total;
        int ex = block_exclusive_scan(v, s_tmp, total);
        if (i < nb) blockSums[i] = ex + carry;
        carry += total;
    }
}

__global__ void k_scan_add(int *out, const int *blockSums, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += blockSums[i / SCAN_BLOCK];
}

// Slots by key. The rows are kept in tile order, so the keys of a wave's 64 consecutive rows come
// in runs; the first lane of each run takes the run's slots with one atomic, all runs of the wave
// at once (one round trip per wave, not one per key). The order inside a cell is then fixed by
// k_bin_sort_cells, so the slots' first placement does not change any result.
__global__ void k_bin_scatter(KPart P, KBins B) {
    if (B.stop[0]) return;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < P.n;
    const int key = valid ? B.key[p] : -1;
    const int lane = (int)__lane_id();
    const int prev = __shfl_up(key, 1);
    const bool start = valid && (lane == 0 || prev != key);
    const unsigned long long starts = __ballot(start);
    const int nvalid = 64 - __clzll(__ballot(valid));                 // valid lanes are a prefix
    const unsigned long long upto = (2ull << lane) - 1ull;           // lanes 0..lane
    const int leader = 63 - __clzll(starts & upto);                   // this lane's run start
    int base = 0;
    if (start) {
        const unsigned long long later = starts & ~upto;
        const int end = later ? __ffsll((long long)later) - 1 : nvalid;
        base = B.cellStart[key] + atomicAdd(&B.cursor[key], end - lane);
    }
    base = __shfl(base, leader & 63);
    if (valid) {
        const int slot = base + (lane - leader);
        B.order[slot] = (int)p;
        B.shome[slot] = B.home[p];
    }
}

// the particles of a cell in upload order (their index at upload, which the memory order of the
// rows does not change), so every sum over a cell's particles has the same order in any layout.
// Most cells arrive in order already (their particles stayed put, in rows sorted by upload index):
// a cell of up to 16 is checked with independent loads first and only sorted when it is not.
__global__ void k_bin_sort_cells(KBins B) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= B.nkeys) return;
    int b = B.cellStart[c], e = B.cellStart[c + 1];
    if (B.stop[0]) e = b;                    // stopped: nothing was scattered
    bool sorted = false;
    if (e - b <= 16) {
        int h[16];
#pragma unroll
        for (int t = 0; t < 16; t++) h[t] = b + t < e ? B.shome[b + t] : 0x7fffffff;
        sorted = true;
#pragma unroll
        for (int t = 1; t < 16; t++) sorted = sorted && h[t - 1] <= h[t];
    }
    if (!sorted) {
        for (int i = b + 1; i < e; i++) {
            const int v = B.order[i];
            const int hv = B.shome[i];
            int j = i - 1;
            while (j >= b && B.shome[j] > hv) {
                B.order[j + 1] = B.order[j];
                B.shome[j + 1] = B.shome[j];
                j--;
            }
            B.order[j + 1] = v;
            B.shome[j + 1] = hv;
        }
    }
    B.cursor[c] = 0;
    B.count[c] = 0;
}

// list of the non-empty (field, tile) pairs; clears the lCPDI kernels' cursors
__global__ void k_tile_list(KBins B, int nc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < TILE_SLOTS * 8) B.work[threadIdx.x * TILE_SLOT_STRIDE] = 0;
    const bool stopped = B.stop[0] != 0;
    if (t == 0 && !stopped) B.stop[2] += 1;   // one more step completed
    const bool on = !stopped && t < B.ntall && B.cellStart[(size_t)(t + 1) * nc] > B.cellStart[(size_t)t * nc];
    if (t < B.ntall) B.tileOn[t] = on ? 1 : 0;
    const unsigned long long m = __ballot(on);
    if (m == 0ull) return;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(B.ntiles, __popcll(m));
    base = __shfl(base, leader);
    if (on) B.tiles[base + __popcll(m & ((1ull << lane) - 1ull))] = t;
}

// the (field, tile) pairs the next node passes visit (node_needed): needed = the pair or one of its
// neighbours is non-empty; needAny = needed now or at the previous binning. A stopped run needs none.
template <int DIM>
__global__ void k_tile_need(KGrid g0, KBins B) {
    const int tall = blockIdx.x * blockDim.x + threadIdx.x;
    if (tall >= B.ntall) return;
    const int fld = tall / g0.tilesTotal;
    const int tg = tall - fld * g0.tilesTotal;
    const KGrid g = tile_grid(g0, tg);
    const int t = tg - g.tileBase;
    const int tx = t % g.ntx, ty = (t / g.ntx) % g.nty, tz = DIM == 3 ? t / (g.ntx * g.nty) : 0;
    int need = 0;
    if (!B.stop[0]) {
        const int zl = DIM == 3 ? max(tz - 1, 0) : 0, zh = DIM == 3 ? min(tz + 1, g.ntz - 1) : 0;
        for (int z = zl; z <= zh && !need; z++)
            for (int y = max(ty - 1, 0); y <= min(ty + 1, g.nty - 1) && !need; y++)
                for (int x = max(tx - 1, 0); x <= min(tx + 1, g.ntx - 1); x++) {
                    const long long u = (long long)fld * g0.tilesTotal + g.tileBase + (z * g.nty + y) * g.ntx + x;
                    if (B.tileOn[u]) { need = 1; break; }
                }
    }
    B.needAny[tall] = need | B.need[tall];
    B.need[tall] = need;
}

// A tile kernel's blocks walk the list of non-empty tiles, so no workgroup is spent on an empty
// tile. Block b takes entry b first, then further entries claimed from one of eight cursors (the
// block's cursor is b mod 8, so the blocks sharing a cursor sit on one XCD and a claim never waits
// behind the whole device), entry gridDim.x + c + 8 k for the k-th claim on cursor c. The claim for
// the next tile is made while the current one runs. Results do not depend on which block takes a tile.
template <typename F>
__device__ __forceinline__ void for_each_tile(const KBins &B, F &tile) {
    const int n = B.ntiles[0];
    __shared__ int s_li;
    const int cur = blockIdx.x & 7;
    int *cursor = &B.work[(B.slot * 8 + cur) * TILE_SLOT_STRIDE];
    int li = blockIdx.x;
    while (li < n) {
        int next = 0;
        if (threadIdx.x == 0) next = (int)gridDim.x + cur + 8 * atomicAdd(cursor, 1);
        tile(B.tiles[li]);
        if (threadIdx.x == 0) s_li = next;
        __syncthreads();
        li = s_li;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Deterministic per-wave scatter. Wave w of the block takes the tile's cells w, w+NW, ... in
// order; for one cell, lane l owns the candidate node offset (l%4-1, (l/4)%4-1, l/16-1) (2D: the
// first 16 lanes, (l%4-1, l/4-1)), so the 64 lanes write 64 different nodes. The lane sums the
// cell's staged particles in sorted order in registers, then adds the sum to the wave's own
// accumulator copy in LDS. The copies are added in wave order at the end. Every node value is
// therefore the same fixed-order sum on every run.
#define NW (NT / 64)

// Footprint columns one wave's cells touch: wave w takes cells c = w, w+NW, ..., whose x index
// runs from w to w + TX - NW (TX is a multiple of NW), so with node offsets -1..2 its accumulator
// needs footprint columns w .. w + WX - 1 only.
template <int DIM> struct WaveFoot {
    static_assert(TileGeom<DIM>::TX % NW == 0, "tile width must be a multiple of the waves per block");
    static constexpr int WX = TileGeom<DIM>::TX - NW + 4;
    static constexpr int SIZE = WX * TileGeom<DIM>::FY * TileGeom<DIM>::FZ;
};

template <int DIM, int NQ, typename Load, typename Add>
__device__ __forceinline__ void wave_cells(const int *s_cs, int r0, int r1,
                                           double (*s_acc)[NQ][WaveFoot<DIM>::SIZE], Load load, Add add) {
    using G = TileGeom<DIM>;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int o1x = l & 3, o1y = (l >> 2) & 3, o1z = DIM == 3 ? (l >> 4) : 1;   // offset + 1
    if constexpr (DIM == 2) {
        // A 2D cell has 16 candidate nodes, so a wave takes four cells per pass: lane group
        // l / 16 one cell, lane l % 16 its candidate. The wave's cells sit at x = w + NW a
        // (a < A) on every row; the four cells of a pass are 4 cells apart in x or y, so they
        // touch different nodes and their sums go to different slots of the wave's copy. Each
        // node's slot takes its cells in a fixed pass order.
        constexpr int A = G::TX / NW;
        static_assert(A == 4 || (A == 2 && G::TY == 8), "2D tiles 4*NW cells wide, or 2*NW wide and 8 high");
        constexpr int PASSES = A * G::TY / 4;
        const int grp = l >> 4;
        const int lx = w + NW * (A == 4 ? grp : (grp & 1));
#pragma unroll 1
        for (int it = 0; it < PASSES; it++) {
            const int ly = A == 4 ? it : it + 4 * (grp >> 1);
            const int c = ly * G::TX + lx;
            const int b = max(s_cs[c], r0), e = min(s_cs[c + 1], r1);
            if (__ballot(b < e) == 0ull) continue;      // none of the four cells has particles here
            double acc[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[q] = 0.;
            int k = b;
            for (; k + 4 <= e; k += 4) {
                auto a0 = load(k - r0, o1x, o1y, o1z);
                auto a1 = load(k + 1 - r0, o1x, o1y, o1z);
                auto a2 = load(k + 2 - r0, o1x, o1y, o1z);
                auto a3 = load(k + 3 - r0, o1x, o1y, o1z);
                add(a0, acc, lx, ly, 0);
                add(a1, acc, lx, ly, 0);
                add(a2, acc, lx, ly, 0);
                add(a3, acc, lx, ly, 0);
            }
            for (; k < e; k++) {
                auto a0 = load(k - r0, o1x, o1y, o1z);
                add(a0, acc, lx, ly, 0);
            }
            if (b < e) {
                const int slot = (ly + o1y) * WaveFoot<DIM>::WX + (lx + o1x - w);
#pragma unroll
                for (int q = 0; q < NQ; q++) s_acc[w][q][slot] += acc[q];
            }
            __builtin_amdgcn_wave_barrier();
        }
        return;
    }
    const bool lane_on = true;
    for (int c = w; c < G::NC; c += NW) {
        int b = max(s_cs[c], r0), e = min(s_cs[c + 1], r1);
        if (b >= e) continue;
        int lx = c % G::TX, ly = (c / G::TX) % G::TY, lz = DIM == 3 ? c / (G::TX * G::TY) : 0;
        double acc[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) acc[q] = 0.;
        if (lane_on) {
            // records of four particles are read before they are added, in particle order
            int k = b;
            for (; k + 4 <= e; k += 4) {
                auto a0 = load(k - r0, o1x, o1y, o1z);
                auto a1 = load(k + 1 - r0, o1x, o1y, o1z);
                auto a2 = load(k + 2 - r0, o1x, o1y, o1z);
                auto a3 = load(k + 3 - r0, o1x, o1y, o1z);
                add(a0, acc, lx, ly, lz);
                add(a1, acc, lx, ly, lz);
                add(a2, acc, lx, ly, lz);
                add(a3, acc, lx, ly, lz);
            }
            for (; k < e; k++) {
                auto a0 = load(k - r0, o1x, o1y, o1z);
                add(a0, acc, lx, ly, lz);
            }
            const int row = DIM == 3 ? (lz + o1z) * G::FY + (ly + o1y) : ly + o1y;   // footprint row
            const int slot = row * WaveFoot<DIM>::WX + (lx + o1x - w);
#pragma unroll
            for (int q = 0; q < NQ; q++) s_acc[w][q][slot] += acc[q];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// particle records of the scatters: per-axis factors of the lane's candidate and the payload
struct RecMass { double wx, wy, wz, mp, vx, vy, vz; };
struct RecForce { double wx, wy, wz, dx, dy, dz, s0, s1, s2, s3, s4, s5; };
struct RecForce2 { double2 x, y, z; double a[6]; };
// with external particle loads or tractions: pFext and the particle's traction force at the node
struct RecForceX : RecForce { double e[3], t[3]; };
struct RecForce2X : RecForce2 { double e[3], t[3]; };

template <int DIM, int NQ>
__device__ __forceinline__ void zero_acc(double (*s_acc)[NQ][WaveFoot<DIM>::SIZE]) {
    double *a = &s_acc[0][0][0];
    for (int i = threadIdx.x; i < NW * NQ * WaveFoot<DIM>::SIZE; i += NT) a[i] = 0.;
}

// add the wave copies that cover each footprint node in wave order and store the tile's sums
template <int DIM, int NQ>
__device__ __forceinline__ void write_partials(const KNodes &N, int t, double (*s_acc)[NQ][WaveFoot<DIM>::SIZE]) {
    using G = TileGeom<DIM>;
    constexpr int WX = WaveFoot<DIM>::WX;
    for (int s = threadIdx.x; s < G::FOOT; s += NT) {
        const int fx = s % G::FX, row = s / G::FX;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            double v = 0.;
            bool first = true;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                const int c = fx - w;
                if (c < 0 || c >= WX) continue;
                const double a = s_acc[w][q][row * WX + c];
                v = first ? a : v + a;
                first = false;
            }
            N.part[((size_t)t * NQMAX + q) * G::FOOT + s] = v;
        }
    }
}

template <int DIM>
__device__ __forceinline__ void load_tile_starts(const KBins &B, int t, int *s_cs) {
    using G = TileGeom<DIM>;
    for (int i = threadIdx.x; i <= G::NC; i += NT) s_cs[i] = B.cellStart[t * G::NC + i];
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// P2G #1: mass and momentum (MassAndMomentumTask -> NodalPoint::AddMassMomentum,
// NodalPointMPM.cpp:419-453). The point count of a node is not summed: every particle has
// mp > 0 and every in-support weight is > 0, so numberPoints > 0 exactly when the mass is > 0.
template <int DIM, int SH>
__global__ __launch_bounds__(NT) void k_p2g_mass(KGrid g0, KPart P, KBins B, KNodes N, KStats st) {
    using G = TileGeom<DIM>;
    constexpr bool CP = SH == MPMGPU_SHAPE_LCPDI;
    constexpr int NCOR = DIM == 3 ? 8 : 4;
    // particles staged per chunk (three blocks per CU); lCPDI: one thread per particle corner
    constexpr int CH = CP ? NT / NCOR : NT / 2;
    __shared__ int s_cs[G::NC + 1];
    __shared__ double s_S[CP ? 1 : DIM * 4][CH + 1];   // [axis*4 + offset][particle]
    __shared__ double s_ct[CP ? CH : 1][NCOR][DIM][4];
    __shared__ double s_q[4][CH + 1];
    __shared__ double s_acc[NW][NQ_MASS][WaveFoot<DIM>::SIZE];
    auto tile = [&](const int tall) {      // (field, context tile) pair
        const KGrid g = tile_grid(g0, tall % g0.tilesTotal);
        load_tile_starts<DIM>(B, tall, s_cs);
        const int tb = s_cs[0], te = s_cs[G::NC];
        if (tb == te) return;
        zero_acc<DIM, NQ_MASS>(s_acc);
        auto scatter = [&](int c0, int c1) {
            wave_cells<DIM, NQ_MASS>(s_cs, c0, c1, s_acc,
                [&](int lq, int ix, int iy, int iz) {
                    RecMass r;
                    if (CP) {
                        double gr[3];
                        cpdi_table_node<DIM, false>(s_ct[lq], nullptr, ix, iy, iz, r.wx, gr);
                        r.wy = 1.; r.wz = 1.;
                    } else {
                        r.wx = s_S[ix][lq]; r.wy = s_S[4 + iy][lq]; r.wz = DIM == 3 ? s_S[(DIM - 1) * 4 + iz][lq] : 1.;
                    }
                    r.mp = s_q[0][lq]; r.vx = s_q[1][lq]; r.vy = s_q[2][lq]; r.vz = s_q[3][lq];
                    return r;
                },
                [&](const RecMass &r, double *acc, int, int, int) {
                    // lCPDI: merged weight (zero for non-member nodes) times mp; otherwise the
                    // product of the staged factors, mp included
                    const double fnmp = CP ? r.wx * r.mp : (DIM == 3 ? r.wx * (r.wy * r.wz) : r.wx * r.wy);
                    acc[0] += fnmp;
                    acc[1] = fma(fnmp, r.vx, acc[1]);
                    acc[2] = fma(fnmp, r.vy, acc[2]);
                    acc[3] = fma(fnmp, r.vz, acc[3]);
                });
        };
        if (!CP) {
            // Rounds of 2*NT slots: every thread first loads the particle words of its two slots of
            // the round (tb + t and tb + t + NT), so the round's global loads overlap; then the four
            // chunks of CH slots are staged from registers by the half of the block that holds them.
            for (int r0 = tb; r0 < te; r0 += 2 * NT) {
                int rp[2], re[2];
                double rx[2][3], rm[2], rv[2][3];
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const int k = r0 + threadIdx.x + j * NT;
                    rp[j] = k < te ? B.order[k] : -1;
                }
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const int p = rp[j];
                    if (p < 0) continue;
                    re[j] = P.elem[p];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        rx[j][a] = a < DIM ? P.f[MPF_POS_X + a][p] : 0.;
                        rv[j][a] = a < DIM ? P.f[MPF_VEL_X + a][p] : 0.;
                    }
                    rm[j] = P.f[MPF_MASS][p];
                }
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int c0 = r0 + c * CH;
                    if (c0 >= te) break;
                    const int j = c >> 1;
                    if ((int)(threadIdx.x / CH) == (c & 1) && rp[j] >= 0) {
                        const int i = threadIdx.x - (c & 1) * CH;
                        int ei, ej, ek;
                        elem_ijk(g, re[j], ei, ej, ek);
                        // factors with zeros outside the support; the x factor carries mp
                        double xi[3], invd[3], S[3][4], dS[3][4];
                        particle_xi<DIM>(g, ei, ej, ek, rx[j][0], rx[j][1], rx[j][2], xi, invd);
                        particle_factors<DIM>(g, P, rp[j], xi, S, dS);
#pragma unroll
                        for (int a = 0; a < DIM; a++)
#pragma unroll
                            for (int o = 0; o < 4; o++) {
                                const double v = S[a][o] >= 0. ? S[a][o] : 0.;
                                s_S[a * 4 + o][i] = a == 0 ? v * rm[j] : v;
                            }
                        s_q[0][i] = rm[j];
                        s_q[1][i] = rv[j][0]; s_q[2][i] = rv[j][1]; s_q[3][i] = rv[j][2];
                    }
                    __syncthreads();
                    scatter(c0, min(c0 + CH, te));
                    __syncthreads();
                }
            }
        } else {
            const int i = threadIdx.x / NCOR, c = threadIdx.x % NCOR;
            for (int c0 = tb; c0 < te; c0 += CH) {
                const int k = c0 + i;
                bool ok = true;
                if (k < te) {
                    const int p = B.order[k];
                    int ei, ej, ek;
                    elem_ijk(g, P.elem[p], ei, ej, ek);
                    double pos[3] = {P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p], DIM == 3 ? P.f[MPF_POS_Z][p] : 0.};
                    // InitializationTask -> GetCPDINodesAndWeights: semi-side vectors of the step (each
                    // corner's thread builds them; the first keeps them for the step's other passes)
                    double r[9];
                    cpdi_semi_sides<DIM>(g, P, p, ei, ej, ek, r);
                    double xi[3];
                    int off[3];
                    ok = cpdi_corner<DIM>(g, pos, r, ei, ej, ek, c, xi, off);
                    cpdi_table<DIM>(xi, off, s_ct[i][c]);
                    if (c == 0) {
#pragma unroll
                        for (int a = 0; a < 9; a++) P.cpr[a][p] = r[a];
                        s_q[0][i] = P.f[MPF_MASS][p];
                        s_q[1][i] = P.f[MPF_VEL_X][p];
                        s_q[2][i] = P.f[MPF_VEL_Y][p];
                        s_q[3][i] = DIM == 3 ? P.f[MPF_VEL_Z][p] : 0.;
                    }
                }
                // one count per particle with an unusable corner (its corner threads are adjacent lanes)
                const unsigned long long bad = __ballot(!ok);
                if (c == 0 && ((bad >> (threadIdx.x & 63)) & ((1ull << NCOR) - 1ull))) atomicAdd(st.cpdiErr, 1ull);
                __syncthreads();
                scatter(c0, min(c0 + CH, te));
                __syncthreads();
            }
        }
        write_partials<DIM, NQ_MASS>(N, tall, s_acc);
    };
    for_each_tile(B, tile);
}

// sum the partials of the tiles that cover node (i,j,k), in fixed tile order
template <int DIM, int NQ>
__device__ __forceinline__ void gather_partials(const KGrid &g, const KBins &B, const KNodes &N, int fld, int i,
                                                int j, int k, double out[NQ]) {
    using G = TileGeom<DIM>;
#pragma unroll
    for (int q = 0; q < NQ; q++) out[q] = 0.;
    int txh = min(g.ntx - 1, (i + 1) / G::TX), txl = i >= G::TX + 1 ? (i - 2) / G::TX : 0;
    int tyh = min(g.nty - 1, (j + 1) / G::TY), tyl = j >= G::TY + 1 ? (j - 2) / G::TY : 0;
    int tzh = 0, tzl = 0;
    if (DIM == 3) { tzh = min(g.ntz - 1, (k + 1) / G::TZ); tzl = k >= G::TZ + 1 ? (k - 2) / G::TZ : 0; }
    for (int tz = tzl; tz <= tzh; tz++)
        for (int ty = tyl; ty <= tyh; ty++)
            for (int tx = txl; tx <= txh; tx++) {
                int t = (tz * g.nty + ty) * g.ntx + tx + g.tileBase + fld * g.tilesTotal;
                if (!B.tileOn[t]) continue;
                int fx = i - (tx * G::TX - 1), fy = j - (ty * G::TY - 1);   // slot in the footprint
                int s = fy * G::FX + fx;
                if (DIM == 3) s += (k - (tz * G::TZ - 1)) * G::FX * G::FY;
#pragma unroll
                for (int q = 0; q < NQ; q++) out[q] += N.part[((size_t)t * NQMAX + q) * G::FOOT + s];
            }
}

// Node passes visit the nodes of the (field, tile) pairs that a non-empty pair's footprint reached at
// this binning or the one before (k_tile_need): a node of any other pair was cleared when its pair
// last was visited, or was never written, so it is skipped and stays zero. A node's own pair is the
// tile holding it (the grid's last node row belongs to the last tile); the pairs whose footprints
// can reach it are that tile and its neighbours.
template <int DIM>
__device__ __forceinline__ bool node_needed(const KGrid &g, const KBins &B, int fld, int i, int j, int k) {
    using G = TileGeom<DIM>;
    const int tx = min(i / G::TX, g.ntx - 1), ty = min(j / G::TY, g.nty - 1);
    const int tz = DIM == 3 ? min(k / G::TZ, g.ntz - 1) : 0;
    return B.needAny[fld * g.tilesTotal + g.tileBase + (tz * g.nty + ty) * g.ntx + tx] != 0;
}

// node pass after P2G #1: mass, momentum, count and v = p/m (MatVelocityField.cpp:158-164,239-251)
template <int DIM>
__global__ void k_node_mass(KGrid g0, KBins B, KNodes N) {
    long long nall = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nall >= g0.nstride * g0.nfields) return;
    const int fld = (int)(nall / g0.nstride);
    const long long ng = nall - (long long)fld * g0.nstride;
    const KGrid g = node_grid(g0, ng);
    const long long n = ng - g.nodeBase;
    if (n >= g.nnodes) return;                // padding between instances
    const KNodes Nf = field_view(g, N, fld);
    if (nall == 0 && N.clist) N.clist[0] = 0;   // contact-node list of the step (regression normals)
    int i = (int)(n % g.nnx);
    long long r = n / g.nnx;
    int j = (int)(r % g.nny);
    int k = (int)(r / g.nny);
    if (!node_needed<DIM>(g, B, fld, i, j, k)) return;
    double a[NQ_MASS];
    gather_partials<DIM, NQ_MASS>(g, B, N, fld, i, j, k, a);
    long long nn = g.nstride;
    Nf.mass[n] = a[0];
    Nf.pk[n] = a[1]; Nf.pk[nn + n] = a[2]; Nf.pk[2 * nn + n] = a[3];
    int cnt = a[0] > 0. ? 1 : 0;       // numberPoints > 0 <=> mass > 0 (all mp > 0)
    Nf.cnt[n] = cnt;
    double vx = 0., vy = 0., vz = 0.;
    if (cnt > 0 && a[0] != 0.) {
        double s = 1. / a[0];
        vx = s * a[1]; vy = s * a[2]; vz = s * a[3];
    }
    Nf.vk[n] = vx; Nf.vk[nn + n] = vy; Nf.vk[2 * nn + n] = vz;
    Nf.ftot[n] = 0.; Nf.ftot[nn + n] = 0.; Nf.ftot[2 * nn + n] = 0.;
}

// ---------------------------------------------------------------------------------------
// load the footprint node vectors of a tile into LDS (zero outside the grid)
template <int DIM, int NV>
__device__ __forceinline__ void load_foot(const KGrid &g, int ox, int oy, int oz, const double *src[NV],
                                          const long long stride[NV], double (*dst)[TileGeom<DIM>::FOOT]) {
    using G = TileGeom<DIM>;
    for (int s = threadIdx.x; s < G::FOOT; s += NT) {
        long long n = foot_node<DIM>(g, s, ox, oy, oz);
#pragma unroll
        for (int v = 0; v < NV; v++) dst[v][s] = n >= 0 ? src[v][n + stride[v]] : 0.;
    }
}

// lCPDI gathers corner by corner. A node's merged weight and gradient are sums over the corners
// whose element holds it (cpdi_table_node), so a gather over the merged node list is the same sum taken
// corner by corner: sum_n w(n) v(n) = ws sum_c sum_j cfn_c(j) v(n_cj) and
// sum_n v(n) grad(n) = sum_c (sum_j cfn_c(j) v(n_cj)) wg_c, where j runs over the nodes of corner c's
// element (cfn below 1e-15 left out, as in the reference's list). Each corner is built on the fly.

// gradient weight of corner c: the reference's per-corner expressions (MatPoint3D.cpp:571-594,
// MatPoint2D.cpp:551-566) collected as (s3 r1 x r2 - s2 r1 x r3 + s1 r2 x r3) / (8 V) in 3D and
// (s1 r2y - s2 r1y, s2 r1x - s1 r2x) / (4 A) in 2D; inv = 1/(8 V) or 1/(4 A)
template <int DIM>
__device__ __forceinline__ void cpdi_wg(const double r[9], double inv, int c, double w[3]) {
    const int q = c & 3;
    const double s1 = (q == 1 || q == 2) ? 1. : -1.;
    const double s2 = q >= 2 ? 1. : -1.;
    if (DIM == 3) {
        const double s3 = c >= 4 ? 1. : -1.;
        const double *r1 = r, *r2 = r + 3, *r3 = r + 6;
        const double c12[3] = {r1[1] * r2[2] - r1[2] * r2[1], r1[2] * r2[0] - r1[0] * r2[2], r1[0] * r2[1] - r1[1] * r2[0]};
        const double c13[3] = {r1[1] * r3[2] - r1[2] * r3[1], r1[2] * r3[0] - r1[0] * r3[2], r1[0] * r3[1] - r1[1] * r3[0]};
        const double c23[3] = {r2[1] * r3[2] - r2[2] * r3[1], r2[2] * r3[0] - r2[0] * r3[2], r2[0] * r3[1] - r2[1] * r3[0]};
#pragma unroll
        for (int a = 0; a < 3; a++) w[a] = (s3 * c12[a] - s2 * c13[a] + s1 * c23[a]) * inv;
    } else {
        w[0] = (s1 * r[4] - s2 * r[1]) * inv;
        w[1] = (s2 * r[0] - s1 * r[3]) * inv;
        w[2] = 0.;
    }
}

// 1/(8 V) (3D) or 1/(4 A) (2D) of the particle domain with semi-side vectors r
template <int DIM>
__device__ __forceinline__ double cpdi_inv_measure(const double r[9]) {
    if (DIM == 3)
        return 1. / (8. * (r[0] * (r[4] * r[8] - r[5] * r[7]) + r[1] * (r[5] * r[6] - r[3] * r[8]) +
                           r[2] * (r[3] * r[7] - r[4] * r[6])));
    return 1. / (4. * (r[0] * r[4] - r[1] * r[3]));
}

// sum over the nodes of corner c's element of cfn * rows[v][node], footprint slots relative to the
// particle's cell (lx, ly, lz) in its tile
template <int DIM, int NV>
__device__ __forceinline__ void cpdi_corner_sum(const double xi[3], const int off[3], int lx, int ly, int lz,
                                                const double (*rows)[TileGeom<DIM>::FOOT], double out[NV]) {
#pragma unroll
    for (int v = 0; v < NV; v++) out[v] = 0.;
    constexpr int NN = DIM == 3 ? 8 : 4;
#pragma unroll
    for (int jn = 0; jn < NN; jn++) {
        const int q = jn & 3;
        const int jx = (q == 1 || q == 2) ? 1 : 0, jy = q >= 2 ? 1 : 0, jz = jn >= 4 ? 1 : 0;
        double cfn = cpdi_cfn<DIM>(xi, jx, jy, jz);
        cfn = cfn < 1e-15 ? 0. : cfn;
        const int sl = foot_slot<DIM>(lx + off[0] + jx, ly + off[1] + jy, DIM == 3 ? lz + off[2] + jz : 0);
#pragma unroll
        for (int v = 0; v < NV; v++) out[v] = fma(cfn, rows[v][sl], out[v]);
    }
}

// Support window of one particle: with a half size lp <= 1 (natural units) a particle with xi < 0
// touches no node at offset 2 and one with xi >= 0 none at offset -1, so three offsets per axis
// starting at w = 0 (offsets -1..1) or w = 1 (0..2) hold every node of its support. The window's
// factors are stored with zeros for nodes outside the support, and derivatives times 2/dx.
template <int DIM>
__device__ __forceinline__ void support_window(const double xi[3], const double S[3][4], const double dS[3][4],
                                               const double invd[3], int w[3], double Sw[3][3], double Dw[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        w[a] = a < DIM && xi[a] >= 0. ? 1 : 0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if (a >= DIM) { Sw[a][j] = j == 1 ? 1. : 0.; Dw[a][j] = 0.; continue; }
            const double sv = w[a] ? S[a][j + 1] : S[a][j];
            const double dv = w[a] ? dS[a][j + 1] : dS[a][j];
            const bool in = sv >= 0.;
            Sw[a][j] = in ? sv : 0.;
            Dw[a][j] = in ? dv * invd[a] : 0.;
        }
    }
}

// velocity gradient of one particle from footprint velocities over its support window
// (MatPoint3D::ExtrapVelocityGradient, MatPoint3D.cpp:62-94; MatPoint2D.cpp:57-89)
template <int DIM>
__device__ __forceinline__ void velocity_gradient(const double xi[3], const double S[3][4], const double dS[3][4],
                                                  const double invd[3], int lx, int ly, int lz,
                                                  const double (*s_v)[TileGeom<DIM>::FOOT], double L[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) L[a][b] = 0.;
    int w[3];
    double Sw[3][3], Dw[3][3];
    support_window<DIM>(xi, S, dS, invd, w, Sw, Dw);
    constexpr int NZ = DIM == 3 ? 3 : 1;
#pragma unroll 1
    for (int jz = 0; jz < NZ; jz++)
#pragma unroll
        for (int jy = 0; jy < 3; jy++)
#pragma unroll
            for (int jx = 0; jx < 3; jx++) {
                const int zz = DIM == 3 ? jz : 1;
                const double syz = Sw[1][jy] * Sw[2][zz], sxz = Sw[0][jx] * Sw[2][zz], sxy = Sw[0][jx] * Sw[1][jy];
                const double gx = Dw[0][jx] * syz, gy = Dw[1][jy] * sxz, gz = Dw[2][zz] * sxy;
                const int s = foot_slot<DIM>(lx + w[0] + jx - 1, ly + w[1] + jy - 1, DIM == 3 ? lz + w[2] + jz - 1 : 0);
                const double vx = s_v[0][s], vy = s_v[1][s];
                L[0][0] = fma(vx, gx, L[0][0]); L[0][1] = fma(vx, gy, L[0][1]);
                L[1][0] = fma(vy, gx, L[1][0]); L[1][1] = fma(vy, gy, L[1][1]);
                if (DIM == 3) {
                    const double vz = s_v[2][s];
                    L[0][2] = fma(vx, gz, L[0][2]); L[1][2] = fma(vy, gz, L[1][2]);
                    L[2][0] = fma(vz, gx, L[2][0]); L[2][1] = fma(vz, gy, L[2][1]); L[2][2] = fma(vz, gz, L[2][2]);
                }
            }
}

// Matrix3::Scale (Matrix3.cpp:410-424)
template <int DIM>
__device__ __forceinline__ void scale_gradient(double L[3][3], double f) {
    L[0][0] *= f; L[0][1] *= f; L[1][0] *= f; L[1][1] *= f; L[2][2] *= f;
    if (DIM == 3) { L[0][2] *= f; L[1][2] *= f; L[2][0] *= f; L[2][1] *= f; }
}

// velocity gradient of particle p from footprint velocities (lx,ly,lz = its cell in the tile):
// the support window of its shape functions, or lCPDI nodes in list order
template <int DIM, int SH>
__device__ __forceinline__ void particle_gradient(const KGrid &g, const KPart &P, int p, int ei, int ej, int ek,
                                                  int lx, int ly, int lz, const double (*s_v)[TileGeom<DIM>::FOOT],
                                                  double L[3][3]) {
    if (SH == MPMGPU_SHAPE_LCPDI) {
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) L[a][b] = 0.;
        double r[9];
#pragma unroll
        for (int a = 0; a < 9; a++) r[a] = P.cpr[a][p];
        const double pos[3] = {P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p], DIM == 3 ? P.f[MPF_POS_Z][p] : 0.};
        const double inv = cpdi_inv_measure<DIM>(r);
        constexpr int NCOR = DIM == 3 ? 8 : 4;
#pragma unroll 1
        for (int c = 0; c < NCOR; c++) {
            double xi[3], vc[3], w[3];
            int off[3];
            cpdi_corner<DIM>(g, pos, r, ei, ej, ek, c, xi, off);
            cpdi_corner_sum<DIM, 3>(xi, off, lx, ly, lz, s_v, vc);
            cpdi_wg<DIM>(r, inv, c, w);
#pragma unroll
            for (int a = 0; a < DIM; a++)
#pragma unroll
                for (int b = 0; b < DIM; b++) L[a][b] = fma(vc[a], w[b], L[a][b]);
        }
    } else {
        double xi[3], invd[3], S[3][4], dS[3][4];
        particle_xi<DIM>(g, ei, ej, ek, P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p],
                         DIM == 3 ? P.f[MPF_POS_Z][p] : 0., xi, invd);
        particle_factors<DIM>(g, P, p, xi, S, dS);
        velocity_gradient<DIM>(xi, S, dS, invd, lx, ly, lz, s_v, L);
    }
}

// IsoPlasticity: the plastic strain (alt rows) and the cumulative plastic strain alpha (hist0) grow
// together, and rotating a zero plastic strain keeps it zero, so a particle with alpha = 0 has
// zero plastic strain: its alt rows (and alpha) are neither read nor written until it yields.
// Upload checks that the state keeps to this.
__device__ __forceinline__ bool alt_rows(int type, double h0) {
    return type != MPMGPU_MAT_ISOPLASTICITY || h0 != 0.;
}

template <int DIM>
__device__ __forceinline__ void load_state(const KPart &P, int p, int type, PState &s) {
    s.ep[0] = P.f[MPF_EP_XX][p]; s.ep[1] = P.f[MPF_EP_YY][p]; s.ep[2] = P.f[MPF_EP_ZZ][p];
    s.ep[5] = P.f[MPF_EP_XY][p]; s.wr[2] = P.f[MPF_WR_XY][p];
    s.sp[0] = P.f[MPF_SP_XX][p]; s.sp[1] = P.f[MPF_SP_YY][p]; s.sp[2] = P.f[MPF_SP_ZZ][p];
    s.sp[5] = P.f[MPF_SP_XY][p];
    if (DIM == 3) {
        s.ep[3] = P.f[MPF_EP_YZ][p]; s.ep[4] = P.f[MPF_EP_XZ][p];
        s.wr[0] = P.f[MPF_WR_YZ][p]; s.wr[1] = P.f[MPF_WR_XZ][p];
        s.sp[3] = P.f[MPF_SP_YZ][p]; s.sp[4] = P.f[MPF_SP_XZ][p];
    } else {
        s.ep[3] = s.ep[4] = 0.; s.wr[0] = s.wr[1] = 0.; s.sp[3] = s.sp[4] = 0.;
    }
    s.P = 0.; s.h0 = 0.; s.h1 = 0.;
#pragma unroll
    for (int c = 0; c < 6; c++) s.alt[c] = 0.;
    if (type != MPMGPU_MAT_ISOTROPIC) {
        s.P = P.f[MPF_PRESSURE][p];
        s.h0 = P.f[MPF_HIST0][p];
        if (alt_rows(type, s.h0)) {
            s.alt[0] = P.f[MPF_ALT_XX][p]; s.alt[1] = P.f[MPF_ALT_YY][p]; s.alt[2] = P.f[MPF_ALT_ZZ][p];
            s.alt[5] = P.f[MPF_ALT_XY][p];
            if (DIM == 3) { s.alt[3] = P.f[MPF_ALT_YZ][p]; s.alt[4] = P.f[MPF_ALT_XZ][p]; }
        }
        if (type == MPMGPU_MAT_NEOHOOKEAN || type == MPMGPU_MAT_MOONEY) s.h1 = P.f[MPF_HIST1][p];
    }
}

template <int DIM>
__device__ __forceinline__ void store_state(const KPart &P, int p, int type, const PState &s) {
    P.f[MPF_EP_XX][p] = s.ep[0]; P.f[MPF_EP_YY][p] = s.ep[1]; P.f[MPF_EP_ZZ][p] = s.ep[2];
    P.f[MPF_EP_XY][p] = s.ep[5]; P.f[MPF_WR_XY][p] = s.wr[2];
    P.f[MPF_SP_XX][p] = s.sp[0]; P.f[MPF_SP_YY][p] = s.sp[1]; P.f[MPF_SP_ZZ][p] = s.sp[2];
    P.f[MPF_SP_XY][p] = s.sp[5];
    if (DIM == 3) {
        P.f[MPF_EP_YZ][p] = s.ep[3]; P.f[MPF_EP_XZ][p] = s.ep[4];
        P.f[MPF_WR_YZ][p] = s.wr[0]; P.f[MPF_WR_XZ][p] = s.wr[1];
        P.f[MPF_SP_YZ][p] = s.sp[3]; P.f[MPF_SP_XZ][p] = s.sp[4];
    }
    if (type != MPMGPU_MAT_ISOTROPIC) {
        P.f[MPF_PRESSURE][p] = s.P;
        if (alt_rows(type, s.h0)) {
            P.f[MPF_ALT_XX][p] = s.alt[0]; P.f[MPF_ALT_YY][p] = s.alt[1]; P.f[MPF_ALT_ZZ][p] = s.alt[2];
            P.f[MPF_ALT_XY][p] = s.alt[5];
            if (DIM == 3) { P.f[MPF_ALT_YZ][p] = s.alt[3]; P.f[MPF_ALT_XZ][p] = s.alt[4]; }
            P.f[MPF_HIST0][p] = s.h0;
        }
        if (type == MPMGPU_MAT_NEOHOOKEAN || type == MPMGPU_MAT_MOONEY) P.f[MPF_HIST1][p] = s.h1;
    }
}

// ---------------------------------------------------------------------------------------
// USF strain update, first half step (UpdateStrainsFirstTask::FullStrainUpdate,
// UpdateStrainsFirstTask.cpp:101-168 -> MatPoint3D::UpdateStrain, MatPoint3D.cpp:45-60)
template <int DIM, int MAT, int SH>
__global__ __launch_bounds__(NT, 2) void k_usf(KGrid g0, KPart P, KBins B, KNodes N, const MpmMaterial *mats,
                                               int only) {
    using G = TileGeom<DIM>;
    __shared__ int s_cs[G::NC + 1];
    __shared__ double s_v[3][G::FOOT];
    auto tile = [&](const int tall) {      // (field, context tile) pair
        if (only >= 0 && g0.tileType[tall % g0.tilesTotal] != only) return;   // another launch's type
        const int fld = tall / g0.tilesTotal;
        const KGrid g = tile_grid(g0, tall - fld * g0.tilesTotal);
        const int t = tall - fld * g0.tilesTotal - g.tileBase;
        const KNodes Nf = field_view(g, N, fld);
        load_tile_starts<DIM>(B, tall, s_cs);
        const int tb = s_cs[0], te = s_cs[G::NC];
        if (tb == te) return;
        int ox, oy, oz;
        tile_origin<DIM>(g, t, ox, oy, oz);
        {
            const double *src[3] = {Nf.vk, Nf.vk, Nf.vk};
            const long long st[3] = {0, g.nstride, 2 * g.nstride};
            load_foot<DIM, 3>(g, ox, oy, oz, src, st, s_v);
        }
        __syncthreads();
        for (int k = tb + threadIdx.x; k < te; k += NT) {
            int p = B.order[k];
            int ei, ej, ek;
            elem_ijk(g, P.elem[p], ei, ej, ek);
            double L[3][3];
            particle_gradient<DIM, SH>(g, P, p, ei, ej, ek, ei - ox, ej - oy, DIM == 3 ? ek - oz : 0, s_v, L);
            scale_gradient<DIM>(L, g.strainFirst);
            const MpmMaterial &m = mats[P.mat[p]];
            const int mtype = MAT ? MAT : m.type;
            PState s;
            load_state<DIM>(P, p, mtype, s);
            constitutive<DIM, MAT>(L, m, s);
            store_state<DIM>(P, p, mtype, s);
        }
    };
    for_each_tile(B, tile);
}

// internal-force scatter (GridForcesTask::Execute, GridForcesTask.cpp:40-153 ->
// MatPoint3D::GetFintPlusFext, MatPoint3D.cpp:248-270; MatPoint2D.cpp:242-263) with the stress
// just updated by k_usf; FX adds N * pFext of the loaded particles (pFext is zero otherwise)
template <int DIM, int SH, bool FX>
__global__ __launch_bounds__(NT) void k_p2g_force(KGrid g0, KPart P, KBins B, KNodes N, KLoads L,
                                                  const int *loaded) {
    using G = TileGeom<DIM>;
    constexpr bool CP = SH == MPMGPU_SHAPE_LCPDI;
    constexpr int NCOR = DIM == 3 ? 8 : 4;
    constexpr int CH = CP ? NT / NCOR : NT / 2;   // particles staged per chunk (lCPDI: a thread per corner)
    constexpr int NA = DIM == 3 ? 6 : 3;      // -mp times the stress terms (shape-function path)
    __shared__ int s_cs[G::NC + 1];
    // shape-function path: per axis and candidate offset, the factor and its derivative times
    // 2/dx (zero outside the support), read as one pair
    __shared__ double2 s_F[CP ? 1 : CH][DIM][4];
    __shared__ double s_a[CP ? 1 : CH][NA];
    __shared__ double s_ct[CP ? CH : 1][NCOR][DIM][4];
    __shared__ double s_cwg[CP ? CH : 1][NCOR][3];
    __shared__ double s_st[CP ? CH : 1][NA];
    __shared__ double s_e[FX ? CH : 1][3];    // pFext of the staged particles
    __shared__ int s_tt[FX ? CH : 1];         // their traction tables (-1: none)
    __shared__ double s_acc[NW][NQ_FORCE][WaveFoot<DIM>::SIZE];
    using R1 = std::conditional_t<FX, RecForceX, RecForce>;
    using R2 = std::conditional_t<FX, RecForce2X, RecForce2>;
    // MatPtLoadBC's pFext of particle p (0 when it carries no load) and its traction table
    auto stage_fext = [&](int slot, int p) {
        const int q = L.map[B.home[p]];
#pragma unroll
        for (int c = 0; c < 3; c++) s_e[slot][c] = q >= 0 ? L.fext[3 * q + c] : 0.;
        s_tt[slot] = q >= 0 && L.nt > 0 ? L.tslot[q] : -1;
    };
    // the staged particle's traction force at candidate node (ix, iy, iz) (zero without tractions)
    auto trac = [&](int lq, int ix, int iy, int iz, double t[3]) {
        const int tt = s_tt[lq];
        const int o = DIM == 3 ? (iz * 4 + iy) * 4 + ix : iy * 4 + ix;
        const double *tab = L.ttab + (size_t)3 * (tt < 0 ? 0 : tt) * TRAC_NODES + o;
#pragma unroll
        for (int c = 0; c < 3; c++) t[c] = tt >= 0 && c < DIM ? tab[c * TRAC_NODES] : 0.;
    };
    auto tile = [&](const int tall) {      // (field, context tile) pair
        // with loads: the launch with the load term takes the marked tiles, the other the rest
        if (loaded && (loaded[tall] != 0) != FX) return;
        const int fld = tall / g0.tilesTotal;
        const KGrid g = tile_grid(g0, tall - fld * g0.tilesTotal);
        load_tile_starts<DIM>(B, tall, s_cs);
        const int tb = s_cs[0], te = s_cs[G::NC];
        if (tb == te) return;
        zero_acc<DIM, NQ_FORCE>(s_acc);
        for (int c0 = tb; c0 < te; c0 += CH) {
            if (CP) {
                // thread (i, c): corner c of staged particle i; the first corner's thread stages the stress
                const int i = threadIdx.x / NCOR, c = threadIdx.x % NCOR;
                const int k = c0 + i;
                if (k < te) {
                    const int p = B.order[k];
                    int ei, ej, ek;
                    elem_ijk(g, P.elem[p], ei, ej, ek);
                    double r[9];
#pragma unroll
                    for (int a = 0; a < 9; a++) r[a] = P.cpr[a][p];
                    double pos[3] = {P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p], DIM == 3 ? P.f[MPF_POS_Z][p] : 0.};
                    double xi[3], w[3];
                    int off[3];
                    cpdi_corner<DIM>(g, pos, r, ei, ej, ek, c, xi, off);
                    cpdi_table<DIM>(xi, off, s_ct[i][c]);
                    cpdi_wg_staged<DIM>(r, c, w);
                    s_cwg[i][c][0] = w[0]; s_cwg[i][c][1] = w[1]; s_cwg[i][c][2] = w[2];
                    if (c == 0) {
                        if constexpr (FX) stage_fext(i, p);
                        const double Pr = P.f[MPF_PRESSURE][p];
                        const double negmp = -P.f[MPF_MASS][p];
                        s_st[i][0] = negmp * (P.f[MPF_SP_XX][p] - Pr);
                        s_st[i][1] = negmp * (P.f[MPF_SP_YY][p] - Pr);
                        s_st[i][2] = negmp * P.f[MPF_SP_XY][p];
                        if constexpr (DIM == 3) {
                            s_st[i][3] = negmp * (P.f[MPF_SP_ZZ][p] - Pr);
                            s_st[i][4] = negmp * P.f[MPF_SP_XZ][p];
                            s_st[i][5] = negmp * P.f[MPF_SP_YZ][p];
                        }
                    }
                }
            }
            int k = c0 + threadIdx.x;
            if (!CP && threadIdx.x < CH && k < te) {
                int p = B.order[k];
                int ei, ej, ek;
                elem_ijk(g, P.elem[p], ei, ej, ek);
                double Pr = P.f[MPF_PRESSURE][p];      // stays 0 for small-strain elastic particles
                const double negmp = -P.f[MPF_MASS][p];
                const double sxx = P.f[MPF_SP_XX][p] - Pr, syy = P.f[MPF_SP_YY][p] - Pr, sxy = P.f[MPF_SP_XY][p];
                {
                    double xi[3], invd[3], S[3][4], dS[3][4];
                    particle_xi<DIM>(g, ei, ej, ek, P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p],
                                     DIM == 3 ? P.f[MPF_POS_Z][p] : 0., xi, invd);
                    particle_factors<DIM>(g, P, p, xi, S, dS);
#pragma unroll
                    for (int a = 0; a < DIM; a++)
#pragma unroll
                        for (int o = 0; o < 4; o++) {
                            const bool in = S[a][o] >= 0.;
                            s_F[threadIdx.x][a][o] = make_double2(in ? S[a][o] : 0., in ? dS[a][o] * invd[a] : 0.);
                        }
                    if constexpr (FX) stage_fext(threadIdx.x, p);
                    s_a[threadIdx.x][0] = negmp * sxx;
                    s_a[threadIdx.x][1] = negmp * syy;
                    s_a[threadIdx.x][2] = negmp * sxy;
                    if (DIM == 3) {
                        s_a[threadIdx.x][3] = negmp * (P.f[MPF_SP_ZZ][p] - Pr);
                        s_a[threadIdx.x][4] = negmp * P.f[MPF_SP_XZ][p];
                        s_a[threadIdx.x][5] = negmp * P.f[MPF_SP_YZ][p];
                    }
                }
            }
            __syncthreads();
            if (!CP) {
                // f = -mp (sigma - P) . grad N; grad N from the staged factor pairs, accumulated with FMA
                wave_cells<DIM, NQ_FORCE>(s_cs, c0, min(c0 + CH, te), s_acc,
                    [&](int lq, int ix, int iy, int iz) {
                        R2 r;
                        r.x = s_F[lq][0][ix];
                        r.y = s_F[lq][1][iy];
                        r.z = DIM == 3 ? s_F[lq][DIM - 1][iz] : make_double2(1., 0.);
#pragma unroll
                        for (int q = 0; q < NA; q++) r.a[q] = s_a[lq][q];
                        if constexpr (FX) {
                            r.e[0] = s_e[lq][0]; r.e[1] = s_e[lq][1]; r.e[2] = s_e[lq][2];
                            trac(lq, ix, iy, iz, r.t);
                        }
                        return r;
                    },
                    [&](const R2 &r, double *acc, int, int, int) {
                        if constexpr (FX) {
                            // + N pFext, N the product of the axis factors, + the traction force
                            const double w = DIM == 3 ? r.x.x * (r.y.x * r.z.x) : r.x.x * r.y.x;
                            acc[0] = fma(r.e[0], w, acc[0]) + r.t[0];
                            acc[1] = fma(r.e[1], w, acc[1]) + r.t[1];
                            if (DIM == 3) acc[2] = fma(r.e[2], w, acc[2]) + r.t[2];
                        }
                        if (DIM == 3) {
                            const double gx = r.x.y * (r.y.x * r.z.x);
                            const double gy = r.y.y * (r.x.x * r.z.x);
                            const double gz = r.z.y * (r.x.x * r.y.x);
                            // a: xx, yy, xy, zz, xz, yz
                            acc[0] = fma(r.a[0], gx, fma(r.a[2], gy, fma(r.a[4], gz, acc[0])));
                            acc[1] = fma(r.a[2], gx, fma(r.a[1], gy, fma(r.a[5], gz, acc[1])));
                            acc[2] = fma(r.a[4], gx, fma(r.a[5], gy, fma(r.a[3], gz, acc[2])));
                        } else {
                            const double gx = r.x.y * r.y.x;
                            const double gy = r.y.y * r.x.x;
                            acc[0] = fma(r.a[0], gx, fma(r.a[2], gy, acc[0]));
                            acc[1] = fma(r.a[2], gx, fma(r.a[1], gy, acc[1]));
                        }
                    });
                __syncthreads();
                continue;
            }
            wave_cells<DIM, NQ_FORCE>(s_cs, c0, min(c0 + CH, te), s_acc,
                [&](int lq, int ix, int iy, int iz) {
                    R1 r;
                    {
                        double gr[3];
                        cpdi_table_node<DIM, true>(s_ct[lq], s_cwg[lq], ix, iy, iz, r.wx, gr);
                        r.wy = r.wz = 1.;
                        r.dx = gr[0]; r.dy = gr[1]; r.dz = gr[2];
                    }
                    r.s0 = s_st[lq][0]; r.s1 = s_st[lq][1]; r.s2 = s_st[lq][2];
                    if constexpr (DIM == 3) { r.s3 = s_st[lq][3]; r.s4 = s_st[lq][4]; r.s5 = s_st[lq][5]; }
                    else { r.s3 = r.s4 = r.s5 = 0.; }
                    if constexpr (FX) {
                        r.e[0] = s_e[lq][0]; r.e[1] = s_e[lq][1]; r.e[2] = s_e[lq][2];
                        trac(lq, ix, iy, iz, r.t);
                    }
                    return r;
                },
                [&](const R1 &r, double *acc, int, int, int) {
                    if constexpr (FX) {              // + N pFext, N the merged corner weight, + traction
                        acc[0] = fma(r.e[0], r.wx, acc[0]) + r.t[0];
                        acc[1] = fma(r.e[1], r.wx, acc[1]) + r.t[1];
                        if (DIM == 3) acc[2] = fma(r.e[2], r.wx, acc[2]) + r.t[2];
                    }
                    // s: -mp times xx, yy, xy, zz, xz, yz of sigma - P; non-member nodes have zero gradient
                    if (DIM == 3) {
                        acc[0] = fma(r.s0, r.dx, fma(r.s2, r.dy, fma(r.s4, r.dz, acc[0])));
                        acc[1] = fma(r.s2, r.dx, fma(r.s1, r.dy, fma(r.s5, r.dz, acc[1])));
                        acc[2] = fma(r.s4, r.dx, fma(r.s5, r.dy, fma(r.s3, r.dz, acc[2])));
                    } else {
                        acc[0] = fma(r.s0, r.dx, fma(r.s2, r.dy, acc[0]));
                        acc[1] = fma(r.s2, r.dx, fma(r.s1, r.dy, acc[1]));
                    }
                });
            __syncthreads();
        }
        write_partials<DIM, NQ_FORCE>(N, tall, s_acc);
    };
    for_each_tile(B, tile);
}

// node pass after the forces: ftot, restore p (identity without BCs), p += f dt, v = p/m
// (PostForcesTask.cpp:46-48, MatVelocityField.cpp:301-314, :239-251)
template <int DIM>
__global__ void k_node_force(KGrid g0, KBins B, KNodes N) {
    long long nall = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nall >= g0.nstride * g0.nfields) return;
    const int fld = (int)(nall / g0.nstride);
    const long long ng = nall - (long long)fld * g0.nstride;
    const KGrid g = node_grid(g0, ng);
    const long long n = ng - g.nodeBase;
    if (n >= g.nnodes) return;                // padding between instances
    const KNodes Nf = field_view(g, N, fld);
    int i = (int)(n % g.nnx);
    long long r = n / g.nnx;
    int j = (int)(r % g.nny);
    int k = (int)(r / g.nny);
    if (!node_needed<DIM>(g, B, fld, i, j, k)) return;
    double f[NQ_FORCE];
    gather_partials<DIM, NQ_FORCE>(g, B, N, fld, i, j, k, f);
    long long nn = g.nstride;
    if (DIM == 2) f[2] = 0.;
    int cnt = Nf.cnt[n];
    if (g.hasGravity && cnt > 0) {         // PostForcesTask.cpp:62-91 -> AddGravityAndBodyForceTask3
        const double m = Nf.mass[n];       // (MatVelocityField.cpp:254-258): ftot += mass g
        for (int c = 0; c < 3; c++) f[c] += m * g.gravity[c];
    }
    Nf.ftot[n] = f[0]; Nf.ftot[nn + n] = f[1]; Nf.ftot[2 * nn + n] = f[2];
    if (cnt <= 0) return;
    double dt = g.timestep;
    double px = Nf.pk[n] + f[0] * dt;
    double py = Nf.pk[nn + n] + f[1] * dt;
    double pz = Nf.pk[2 * nn + n] + f[2] * dt;
    Nf.pk[n] = px; Nf.pk[nn + n] = py; Nf.pk[2 * nn + n] = pz;
    double m = Nf.mass[n];
    if (m != 0.) {
        double s = 1. / m;
        Nf.vk[n] = s * px; Nf.vk[nn + n] = s * py; Nf.vk[2 * nn + n] = s * pz;
    }
}

// ---------------------------------------------------------------------------------------
// Particle update (FLIP, no damping) fused with the momentum re-extrapolation of the USL pass
// with start-of-step shape functions (MatVelocityField::IncrementNodalVelAcc,
// MatVelocityField.cpp:410-426; MatPoint3D::MoveParticle, MatPoint3D.cpp:104-194;
// NodalPoint::AddMassMomentumLast, NodalPointMPM.cpp:470-486)
// MOVE: G2P_FLIP (no damping), G2P_FLIP_DAMPED, or G2P_SERIES, which moves by FMPM(mode) for
// mode > 0 and by XPIC(-mode) for mode < 0 with the damping of the grid (zero damping gives the
// same numbers as none). The FLIP instances keep their registers for the common update.
enum { G2P_FLIP = 0, G2P_FLIP_DAMPED = 1, G2P_SERIES = 2 };
template <int DIM, int SH, int MOVE>
__global__ __launch_bounds__(NT, SH == MPMGPU_SHAPE_LCPDI ? 1 : G2P_WAVES) void k_g2p_p2g(KGrid g0, KPart P, KBins B, KNodes N,
                                                                                  const MpmMaterial *mats, int mode) {
    constexpr bool DAMP = MOVE != G2P_FLIP;
    using G = TileGeom<DIM>;
    constexpr bool CP = SH == MPMGPU_SHAPE_LCPDI;
    constexpr int NCOR = DIM == 3 ? 8 : 4;
    // particles per scatter round: half the block, or (lCPDI) as many as the corner tables hold
    constexpr int SUB = CP ? NT / NCOR : NT / 2;
    __shared__ int s_cs[G::NC + 1];
    __shared__ double s_n[6][G::FOOT];         // v(3), f/m(3)
    __shared__ double s_S[CP ? 1 : DIM * 4][SUB + 1];   // [axis*4 + offset][particle]
    __shared__ double s_ct[CP ? SUB : 1][NCOR][DIM][4];
    __shared__ double s_q[4][SUB + 1];
    __shared__ double s_acc[NW][NQ_MOM][WaveFoot<DIM>::SIZE];
    auto tile = [&](const int tall) {      // (field, context tile) pair
        const int fld = tall / g0.tilesTotal;
        const KGrid g = tile_grid(g0, tall - fld * g0.tilesTotal);
        const int t = tall - fld * g0.tilesTotal - g.tileBase;
        const KNodes Nf = field_view(g, N, fld);
        load_tile_starts<DIM>(B, tall, s_cs);
        const int tb = s_cs[0], te = s_cs[G::NC];
        if (tb == te) return;
        int ox, oy, oz;
        tile_origin<DIM>(g, t, ox, oy, oz);
        {
            // footprint velocities and accelerations f/m (zero outside the grid and on nodes without
            // mass, which no particle of the tile touches)
            const long long nn = g.nstride;
            for (int s = threadIdx.x; s < G::FOOT; s += NT) {
                const long long n = foot_node<DIM>(g, s, ox, oy, oz);
                const double m = n >= 0 ? Nf.mass[n] : 0.;
                const double r = m > 0. ? 1. / m : 0.;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    s_n[c][s] = n >= 0 ? Nf.vk[c * nn + n] : 0.;
                    s_n[3 + c][s] = n >= 0 ? Nf.ftot[c * nn + n] * r : 0.;
                }
            }
        }
        zero_acc<DIM, NQ_MOM>(s_acc);
        __syncthreads();
        const double dt = g.timestep;
        for (int c0 = tb; c0 < te; c0 += NT) {
            const int c1 = min(c0 + NT, te);
            const int k = c0 + threadIdx.x;
            double S[3][4], q4[4];
            CpdiCorners cc;
            if (k < c1) {
                int p = B.order[k];
                // every word of the particle is loaded before anything is stored, so the loads are
                // issued together (a later load could not move above a store to another row)
                const int e0 = P.elem[p];
                double pos[3] = {P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p], DIM == 3 ? P.f[MPF_POS_Z][p] : 0.};
                double vold[3] = {P.f[MPF_VEL_X][p], P.f[MPF_VEL_Y][p], DIM == 3 ? P.f[MPF_VEL_Z][p] : 0.};
                const double mp = P.f[MPF_MASS][p];
                int ei, ej, ek;
                elem_ijk(g, e0, ei, ej, ek);
                int lx = ei - ox, ly = ej - oy, lz = DIM == 3 ? ek - oz : 0;
                double Svk[3] = {0., 0., 0.}, Sacc[3] = {0., 0., 0.};
                if (CP) {
                    cpdi_particle<DIM>(g, P, p, ei, ej, ek, cc);
                    constexpr int NCOR = DIM == 3 ? 8 : 4;
                    constexpr double ws = DIM == 3 ? 0.125 : 0.25;
                    double sum[6] = {0., 0., 0., 0., 0., 0.};
#pragma unroll
                    for (int c = 0; c < NCOR; c++) {
                        double vc[6];
                        cpdi_corner_sum<DIM, 6>(cc.xi[c], cc.off[c], lx, ly, lz, s_n, vc);
#pragma unroll
                        for (int v = 0; v < 6; v++) sum[v] += vc[v];
                    }
#pragma unroll
                    for (int a = 0; a < 3; a++) { Svk[a] = ws * sum[a]; Sacc[a] = ws * sum[3 + a]; }
                } else {
                double xi[3], invd[3], dS[3][4];
                particle_xi<DIM>(g, ei, ej, ek, pos[0], pos[1], pos[2], xi, invd);
                particle_factors<DIM>(g, P, p, xi, S, dS);
                int w[3];
                double Sw[3][3], Dw[3][3];
                support_window<DIM>(xi, S, dS, invd, w, Sw, Dw);
                constexpr int NZ = DIM == 3 ? 3 : 1;
#pragma unroll
                for (int jz = 0; jz < NZ; jz++)
#pragma unroll
                    for (int jy = 0; jy < 3; jy++)
#pragma unroll
                        for (int jx = 0; jx < 3; jx++) {
                            const double fi = DIM == 3 ? Sw[0][jx] * (Sw[1][jy] * Sw[2][jz]) : Sw[0][jx] * Sw[1][jy];
                            const int s = foot_slot<DIM>(lx + w[0] + jx - 1, ly + w[1] + jy - 1,
                                                         DIM == 3 ? lz + w[2] + jz - 1 : 0);
                            Svk[0] = fma(s_n[0][s], fi, Svk[0]);
                            Svk[1] = fma(s_n[1][s], fi, Svk[1]);
                            Svk[2] = fma(s_n[2][s], fi, Svk[2]);
                            Sacc[0] = fma(s_n[3][s], fi, Sacc[0]);
                            Sacc[1] = fma(s_n[4][s], fi, Sacc[1]);
                            Sacc[2] = fma(s_n[5][s], fi, Sacc[2]);
                        }
                }
                double vnew[3] = {0., 0., 0.};
                // particle damping: the material's own <PDamping> in place of the run's
                // (MaterialBase::GetMaterialDamping, MaterialBaseMPM.cpp:742-744)
                double alphaP = 0.;
                if (DAMP) {
                    const MpmMaterial &mt = mats[P.mat[p]];
                    alphaP = mt.usePDamping ? mt.pdamping : g.particleDamping;
                }
                // MatPoint3D::MoveParticle (MatPoint3D.cpp:104-194): Adamp0 = alpha_p v + alpha_g vm
                // (the DAMP instance; zero damping gives delV = Sacc dt for FLIP)
                if (MOVE != G2P_SERIES) {
                    // FLIP: vm = S(v - a dt), v += (Sa - Adamp0) dt, x += (vm + delV/2) dt (:104-160)
#pragma unroll
                    for (int c = 0; c < DIM; c++) {
                        double vm = Svk[c] - Sacc[c] * dt;
                        double delV = DAMP ? (Sacc[c] - (vold[c] * alphaP + vm * g.gridDamping)) * dt
                                           : Sacc[c] * dt;
                        vnew[c] = vold[c] + delV;
                        P.f[MPF_VEL_X + c][p] = vnew[c];
                        double delXRate = vm + 0.5 * delV;
                        P.posn[c][p] = pos[c] + delXRate * dt;
                    }
                } else {
                    // FMPM(k) (mode k): vm = S v+(k), v = vm - Adamp0 dt, x += (v(n) + v(n+1)) dt/2
                    // (:122-137); XPIC(k) (mode -k): vm = S(v(k) + a dt), v = vm - Adamp0 dt,
                    // x += (vm + delV/2) dt (:152-170)
#pragma unroll
                    for (int c = 0; c < DIM; c++) {
                        const double vm = Svk[c];
                        const double adamp = DAMP ? vold[c] * alphaP + vm * g.gridDamping : 0.;
                        if (mode > 0) {
                            vnew[c] = vm + adamp * -dt;
                            const double delXRate = vold[c] + vnew[c];
                            P.posn[c][p] = pos[c] + delXRate * (0.5 * dt);
                        } else {
                            vnew[c] = vm - adamp * dt;
                            const double delV = vnew[c] - vold[c];
                            const double delXRate = vm + 0.5 * delV;
                            P.posn[c][p] = pos[c] + delXRate * dt;
                        }
                        P.f[MPF_VEL_X + c][p] = vnew[c];
                    }
                }
                q4[0] = mp;
                q4[1] = vnew[0]; q4[2] = vnew[1]; q4[3] = vnew[2];
            }
            for (int h = 0; h < NT / SUB; h++) {
                const int r0 = c0 + h * SUB, r1 = min(r0 + SUB, c1);
                if (r0 >= r1) break;
                if ((int)(threadIdx.x / SUB) == h && k < c1) {
                    const int i = threadIdx.x - h * SUB;
                    if (CP) {
                        // the tables are built here, at the thread's turn (an empty asm makes the corner
                        // coordinates opaque, so the compiler cannot build them once before the rounds
                        // and hold all of them in registers)
#pragma unroll
                        for (int c = 0; c < NCOR; c++) {
                            double xo[3];
#pragma unroll
                            for (int a = 0; a < 3; a++) {
                                xo[a] = cc.xi[c][a];
                                asm volatile("" : "+v"(xo[a]));
                            }
                            cpdi_table<DIM>(xo, cc.off[c], s_ct[i][c]);
                        }
                    } else {
                        // factors with zeros outside the support; the x factor carries mp
#pragma unroll
                        for (int a = 0; a < DIM; a++)
#pragma unroll
                            for (int o = 0; o < 4; o++) {
                                const double v = S[a][o] >= 0. ? S[a][o] : 0.;
                                s_S[a * 4 + o][i] = a == 0 ? v * q4[0] : v;
                            }
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) s_q[q][i] = q4[q];
                }
                __syncthreads();
                wave_cells<DIM, NQ_MOM>(s_cs, r0, r1, s_acc,
                    [&](int lq, int ix, int iy, int iz) {
                        RecMass r;
                        if (CP) {
                            double gr[3];
                            cpdi_table_node<DIM, false>(s_ct[lq], nullptr, ix, iy, iz, r.wx, gr);
                            r.wy = 1.; r.wz = 1.;
                        } else {
                            r.wx = s_S[ix][lq]; r.wy = s_S[4 + iy][lq]; r.wz = DIM == 3 ? s_S[(DIM - 1) * 4 + iz][lq] : 1.;
                        }
                        r.mp = s_q[0][lq]; r.vx = s_q[1][lq]; r.vy = s_q[2][lq]; r.vz = s_q[3][lq];
                        return r;
                    },
                    [&](const RecMass &r, double *acc, int, int, int) {
                        const double fnmp = CP ? r.wx * r.mp : (DIM == 3 ? r.wx * (r.wy * r.wz) : r.wx * r.wy);
                        acc[0] = fma(r.vx, fnmp, acc[0]);
                        acc[1] = fma(r.vy, fnmp, acc[1]);
                        acc[2] = fma(r.vz, fnmp, acc[2]);
                    });
                __syncthreads();
            }
        }
        write_partials<DIM, NQ_MOM>(N, tall, s_acc);
    };
    for_each_tile(B, tile);
}

// node pass of the USL re-extrapolation: p (rezeroed, RezeroNodeTask6) and v = p/m
template <int DIM>
__global__ void k_node_mom2(KGrid g0, KBins B, KNodes N) {
    long long nall = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nall >= g0.nstride * g0.nfields) return;
    const int fld = (int)(nall / g0.nstride);
    const long long ng = nall - (long long)fld * g0.nstride;
    const KGrid g = node_grid(g0, ng);
    const long long n = ng - g.nodeBase;
    if (n >= g.nnodes) return;                // padding between instances
    const KNodes Nf = field_view(g, N, fld);
    int cnt = Nf.cnt[n];
    if (cnt <= 0) return;
    int i = (int)(n % g.nnx);
    long long r = n / g.nnx;
    int j = (int)(r % g.nny);
    int k = (int)(r / g.nny);
    if (!node_needed<DIM>(g, B, fld, i, j, k)) return;
    double a[NQ_MOM];
    gather_partials<DIM, NQ_MOM>(g, B, N, fld, i, j, k, a);
    long long nn = g.nstride;
    Nf.pk[n] = a[0]; Nf.pk[nn + n] = a[1]; Nf.pk[2 * nn + n] = a[2];
    double m = Nf.mass[n];
    if (m != 0.) {
        double s = 1. / m;
        Nf.vk[n] = s * a[0]; Nf.vk[nn + n] = s * a[1]; Nf.vk[2 * nn + n] = s * a[2];
    }
}

// ---------------------------------------------------------------------------------------
// XPIC(k) and FMPM(k) grid velocities (XPICExtrapolationTask::Execute, XPICExtrapolationTask.cpp:50-164)
// for one material velocity field per material: v(k) = sum_{r<k} (P(I - A))^r v0 with
// A = diag(1/m) S^T Mp S, built as vk = v0 + Delta v(2) + ... + Delta v(k), each increment
// Delta v(r) = (I - A) Delta v(r-1) with the held components of the conditions zeroed.

// InitializeXPICData -> MatVelocityField::XPICSupport(INITIALIZE_XPIC) (MatVelocityField.cpp:321-345):
// FMPM starts from v0 = pk/m (the momenta the caller's pass left), XPIC from the velocities
// before the momentum update, v0 = (pk - ftot dt)/m, with ftot dt/m added to vk so the result
// is v(k) + a dt. vprev = v0 on every node with particles.
template <int DIM>
__global__ void k_xpic_init(KGrid g0, KBins B, KNodes N, int fmpm) {
    long long nall = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nall >= g0.nstride * g0.nfields) return;
    const int fld = (int)(nall / g0.nstride);
    const long long ng = nall - (long long)fld * g0.nstride;
    const KGrid g = node_grid(g0, ng);
    const long long n = ng - g.nodeBase;
    if (n >= g.nnodes) return;
    const KNodes Nf = field_view(g, N, fld);
    if (Nf.cnt[n] <= 0) return;
    const long long nn = g.nstride;
    const double m = Nf.mass[n], dt = g.timestep, sc = 1. / m;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double p = Nf.pk[c * nn + n];
        double vp, vs;
        if (fmpm) {
            vp = sc * p;
            vs = vp;
        } else {
            const double f = Nf.ftot[c * nn + n];
            vp = (p + f * -dt) * sc;
            vs = vp + f * (dt / m);
        }
        Nf.vprev[c * nn + n] = vp;
        Nf.vk[c * nn + n] = vs;
    }
}

// Particle loop of one increment (XPICExtrapolationTask::XPICDoubleLoop, XPICExtrapolationTask.cpp:182-211):
// node i gains sum_p mp S_ip (sum_j S_jp vprev_j), the particle's gather of the last increment
// scattered as a momentum; the node pass divides by m_i. Shape functions are those of the step's
// start, as in every pass of the step.
template <int DIM, int SH>
__global__ __launch_bounds__(NT, SH == MPMGPU_SHAPE_LCPDI ? 1 : G2P_WAVES) void k_xpic_p2g(KGrid g0, KPart P, KBins B, KNodes N) {
    using G = TileGeom<DIM>;
    constexpr bool CP = SH == MPMGPU_SHAPE_LCPDI;
    constexpr int NCOR = DIM == 3 ? 8 : 4;
    constexpr int SUB = CP ? NT / NCOR : NT / 2;
    __shared__ int s_cs[G::NC + 1];
    __shared__ double s_n[3][G::FOOT];         // vprev
    __shared__ double s_S[CP ? 1 : DIM * 4][SUB + 1];
    __shared__ double s_ct[CP ? SUB : 1][NCOR][DIM][4];
    __shared__ double s_q[4][SUB + 1];
    __shared__ double s_acc[NW][NQ_MOM][WaveFoot<DIM>::SIZE];
    auto tile = [&](const int tall) {
        const int fld = tall / g0.tilesTotal;
        const KGrid g = tile_grid(g0, tall - fld * g0.tilesTotal);
        const int t = tall - fld * g0.tilesTotal - g.tileBase;
        const KNodes Nf = field_view(g, N, fld);
        load_tile_starts<DIM>(B, tall, s_cs);
        const int tb = s_cs[0], te = s_cs[G::NC];
        if (tb == te) return;
        int ox, oy, oz;
        tile_origin<DIM>(g, t, ox, oy, oz);
        {
            // footprint increments (zero outside the grid and on nodes without particles)
            const long long nn = g.nstride;
            for (int s = threadIdx.x; s < G::FOOT; s += NT) {
                const long long n = foot_node<DIM>(g, s, ox, oy, oz);
                const bool on = n >= 0 && Nf.cnt[n] > 0;
#pragma unroll
                for (int c = 0; c < 3; c++) s_n[c][s] = on ? Nf.vprev[c * nn + n] : 0.;
            }
        }
        zero_acc<DIM, NQ_MOM>(s_acc);
        __syncthreads();
        for (int c0 = tb; c0 < te; c0 += NT) {
            const int c1 = min(c0 + NT, te);
            const int k = c0 + threadIdx.x;
            double S[3][4], q4[4];
            CpdiCorners cc;
            if (k < c1) {
                const int p = B.order[k];
                const int e0 = P.elem[p];
                const double pos[3] = {P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p], DIM == 3 ? P.f[MPF_POS_Z][p] : 0.};
                const double mp = P.f[MPF_MASS][p];
                int ei, ej, ek;
                elem_ijk(g, e0, ei, ej, ek);
                const int lx = ei - ox, ly = ej - oy, lz = DIM == 3 ? ek - oz : 0;
                double u[3] = {0., 0., 0.};
                if (CP) {
                    cpdi_particle<DIM>(g, P, p, ei, ej, ek, cc);
                    constexpr double ws = DIM == 3 ? 0.125 : 0.25;
                    double sum[3] = {0., 0., 0.};
#pragma unroll
                    for (int c = 0; c < NCOR; c++) {
                        double vc[3];
                        cpdi_corner_sum<DIM, 3>(cc.xi[c], cc.off[c], lx, ly, lz, s_n, vc);
#pragma unroll
                        for (int v = 0; v < 3; v++) sum[v] += vc[v];
                    }
#pragma unroll
                    for (int a = 0; a < 3; a++) u[a] = ws * sum[a];
                } else {
                    double xi[3], invd[3], dS[3][4];
                    particle_xi<DIM>(g, ei, ej, ek, pos[0], pos[1], pos[2], xi, invd);
                    particle_factors<DIM>(g, P, p, xi, S, dS);
                    int w[3];
                    double Sw[3][3], Dw[3][3];
                    support_window<DIM>(xi, S, dS, invd, w, Sw, Dw);
                    constexpr int NZ = DIM == 3 ? 3 : 1;
#pragma unroll
                    for (int jz = 0; jz < NZ; jz++)
#pragma unroll
                        for (int jy = 0; jy < 3; jy++)
#pragma unroll
                            for (int jx = 0; jx < 3; jx++) {
                                const double fi = DIM == 3 ? Sw[0][jx] * (Sw[1][jy] * Sw[2][jz]) : Sw[0][jx] * Sw[1][jy];
                                const int s = foot_slot<DIM>(lx + w[0] + jx - 1, ly + w[1] + jy - 1,
                                                             DIM == 3 ? lz + w[2] + jz - 1 : 0);
                                u[0] = fma(s_n[0][s], fi, u[0]);
                                u[1] = fma(s_n[1][s], fi, u[1]);
                                u[2] = fma(s_n[2][s], fi, u[2]);
                            }
                }
                q4[0] = mp;
                q4[1] = u[0]; q4[2] = u[1]; q4[3] = u[2];
            }
            for (int h = 0; h < NT / SUB; h++) {
                const int r0 = c0 + h * SUB, r1 = min(r0 + SUB, c1);
                if (r0 >= r1) break;
                if ((int)(threadIdx.x / SUB) == h && k < c1) {
                    const int i = threadIdx.x - h * SUB;
                    if (CP) {
#pragma unroll
                        for (int c = 0; c < NCOR; c++) {
                            double xo[3];
#pragma unroll
                            for (int a = 0; a < 3; a++) {
                                xo[a] = cc.xi[c][a];
                                asm volatile("" : "+v"(xo[a]));
                            }
                            cpdi_table<DIM>(xo, cc.off[c], s_ct[i][c]);
                        }
                    } else {
#pragma unroll
                        for (int a = 0; a < DIM; a++)
#pragma unroll
                            for (int o = 0; o < 4; o++) {
                                const double v = S[a][o] >= 0. ? S[a][o] : 0.;
                                s_S[a * 4 + o][i] = a == 0 ? v * q4[0] : v;
                            }
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) s_q[q][i] = q4[q];
                }
                __syncthreads();
                wave_cells<DIM, NQ_MOM>(s_cs, r0, r1, s_acc,
                    [&](int lq, int ix, int iy, int iz) {
                        RecMass r;
                        if (CP) {
                            double gr[3];
                            cpdi_table_node<DIM, false>(s_ct[lq], nullptr, ix, iy, iz, r.wx, gr);
                            r.wy = 1.; r.wz = 1.;
                        } else {
                            r.wx = s_S[ix][lq]; r.wy = s_S[4 + iy][lq]; r.wz = DIM == 3 ? s_S[(DIM - 1) * 4 + iz][lq] : 1.;
                        }
                        r.mp = s_q[0][lq]; r.vx = s_q[1][lq]; r.vy = s_q[2][lq]; r.vz = s_q[3][lq];
                        return r;
                    },
                    [&](const RecMass &r, double *acc, int, int, int) {
                        const double fnmp = CP ? r.wx * r.mp : (DIM == 3 ? r.wx * (r.wy * r.wz) : r.wx * r.wy);
                        acc[0] = fma(r.vx, fnmp, acc[0]);
                        acc[1] = fma(r.vy, fnmp, acc[1]);
                        acc[2] = fma(r.vz, fnmp, acc[2]);
                    });
                __syncthreads();
            }
        }
        write_partials<DIM, NQ_MOM>(N, tall, s_acc);
    };
    for_each_tile(B, tile);
}

// Node pass of one increment (GetDeltaV and UpdateXStar, XPICExtrapolationTask.cpp:222-240;
// MatVelocityField::XPICSupport GET_DELTAV and UPDATE_VSTAR, MatVelocityField.cpp:347-378):
// vprev -= (the scatter's sum)/m, vk += vprev. The XPIC pass of the velocity conditions then zeroes
// the held components of vprev and takes them back out of vk. Block 0 clears the cursors of the
// XPIC tile slot for the next increment's scatter (the scatter before this pass has finished).
template <int DIM>
__global__ void k_xpic_node(KGrid g0, KBins B, KNodes N) {
    if (blockIdx.x == 0 && threadIdx.x < 8) B.work[(TILE_SLOT_XPIC * 8 + threadIdx.x) * TILE_SLOT_STRIDE] = 0;
    long long nall = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nall >= g0.nstride * g0.nfields) return;
    const int fld = (int)(nall / g0.nstride);
    const long long ng = nall - (long long)fld * g0.nstride;
    const KGrid g = node_grid(g0, ng);
    const long long n = ng - g.nodeBase;
    if (n >= g.nnodes) return;
    const KNodes Nf = field_view(g, N, fld);
    if (Nf.cnt[n] <= 0) return;
    int i = (int)(n % g.nnx);
    long long r = n / g.nnx;
    int j = (int)(r % g.nny);
    int k = (int)(r / g.nny);
    double a[NQ_MOM];
    gather_partials<DIM, NQ_MOM>(g, B, N, fld, i, j, k, a);
    const long long nn = g.nstride;
    const double m = Nf.mass[n];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double d = Nf.vprev[c * nn + n] - a[c] / m;
        Nf.vprev[c * nn + n] = d;
        Nf.vk[c * nn + n] += d;
    }
}

// ---------------------------------------------------------------------------------------
// element test of a rectangle (EightNodeIsoparamBrick.cpp:207-213, FourNodeIsoparam.cpp:413-423)
template <int DIM>
__device__ __forceinline__ bool pt_in_elem(const KGrid &g, int ei, int ej, int ek, const double p[3]) {
    double xmn = g.xmin + (double)ei * g.dx, xmx = g.xmin + (double)(ei + 1) * g.dx;
    if (p[0] < xmn || p[0] >= xmx) return false;
    double ymn = g.ymin + (double)ej * g.dy, ymx = g.ymin + (double)(ej + 1) * g.dy;
    if (p[1] < ymn || p[1] >= ymx) return false;
    if (DIM == 3) {
        double zmn = g.zmin + (double)ek * g.dz, zmx = g.zmin + (double)(ek + 1) * g.dz;
        if (p[2] < zmn || p[2] >= zmx) return false;
    }
    return true;
}

// USL strain update (second half step) and ResetElementsTask; also the cell key and histogram
// of the next binning
template <int DIM, int MAT, int SH>
__global__ __launch_bounds__(NT, 2) void k_usl_reset(KGrid g0, KPart P, KBins B, KNodes N, const MpmMaterial *mats,
                                                  KStats st, int only) {
    using G = TileGeom<DIM>;
    __shared__ int s_cs[G::NC + 1];
    __shared__ double s_v[3][G::FOOT];
    auto tile = [&](const int tall) {      // (field, context tile) pair
        if (only >= 0 && g0.tileType[tall % g0.tilesTotal] != only) return;   // another launch's type
        const int fld = tall / g0.tilesTotal;
        const KGrid g = tile_grid(g0, tall - fld * g0.tilesTotal);
        const int t = tall - fld * g0.tilesTotal - g.tileBase;
        const KNodes Nf = field_view(g, N, fld);
        load_tile_starts<DIM>(B, tall, s_cs);
        const int tb = s_cs[0], te = s_cs[G::NC];
        if (tb == te) return;
        int ox, oy, oz;
        tile_origin<DIM>(g, t, ox, oy, oz);
        {
            const double *src[3] = {Nf.vk, Nf.vk, Nf.vk};
            const long long stt[3] = {0, g.nstride, 2 * g.nstride};
            load_foot<DIM, 3>(g, ox, oy, oz, src, stt, s_v);
        }
        __syncthreads();
        for (int k = tb + threadIdx.x; k < te; k += NT) {
            int p = B.order[k];
            int e = P.elem[p];
            // the end-of-step position is loaded here, before the state stores below, so its load is
            // issued with the others instead of after them
            double pos[3] = {P.posn[0][p], P.posn[1][p], DIM == 3 ? P.posn[2][p] : 0.};
            int ei, ej, ek;
            elem_ijk(g, e, ei, ej, ek);
            double L[3][3];
            particle_gradient<DIM, SH>(g, P, p, ei, ej, ek, ei - ox, ej - oy, DIM == 3 ? ek - oz : 0, s_v, L);
            scale_gradient<DIM>(L, g.strainLast);
            const MpmMaterial &m = mats[P.mat[p]];
            const int mtype = MAT ? MAT : m.type;
            PState s;
            load_state<DIM>(P, p, mtype, s);
            constitutive<DIM, MAT>(L, m, s);
            store_state<DIM>(P, p, mtype, s);

            // ResetElementsTask::ResetElement (ResetElementsTask.cpp:196-228)
            if (pos[0] != pos[0] || pos[1] != pos[1] || pos[2] != pos[2]) {
                atomicAdd(st.nanCount, 1ull);
                if (g.stopOnNan) stop_run(B, STOP_NAN, -1);
            } else if (!pt_in_elem<DIM>(g, ei, ej, ek, pos)) {
                // MeshInfo::FindElementFromPoint (MeshInfo.cpp:593-640) and EdgeElement2D/3D (:171-199)
                bool left = false;
                long long horiz = g.nx, vert = g.ny, depth = DIM == 3 ? g.nz : 1;
                long long total = horiz * vert * depth;
                long long col = (long long)(int)((pos[0] - g.xmin) / g.dx);
                if (col < 0 || col >= horiz) {
                    if (pos[0] == g.xmin + horiz * g.dx) col = horiz - 1; else left = true;
                }
                long long row = 0, zrow = 0;
                if (!left) {
                    row = (long long)(int)((pos[1] - g.ymin) / g.dy);
                    if (row < 0 || row >= vert) {
                        if (pos[1] == g.ymin + vert * g.dy) row = vert - 1; else left = true;
                    }
                }
                if (!left && DIM == 3) {
                    zrow = (long long)(int)((pos[2] - g.zmin) / g.dz);
                    if (zrow < 0 || zrow >= depth) {
                        if (pos[2] == g.zmin + depth * g.dz) zrow = depth - 1; else left = true;
                    }
                }
                long long num = 0;
                if (!left) {
                    num = (DIM == 3 ? horiz * (zrow * vert + row) + col : row * horiz + col) + 1;
                    if (DIM == 3) {
                        if (num <= horiz * vert || num > total - horiz * vert) left = true;
                        else if (num % horiz <= 1) left = true;
                        else {
                            long long xz = num % (horiz * vert);
                            if (xz <= horiz || xz > horiz * (vert - 1)) left = true;
                        }
                    } else {
                        if (num <= horiz || num > total - horiz) left = true;
                        else if (num % horiz <= 1) left = true;
                    }
                }
                if (!left) {
                    e = (int)(num - 1);
                    P.elem[p] = e;
                    elem_ijk(g, e, ei, ej, ek);
                } else {
                    // the warning count comes first, then the deletion (ResetElementsTask.cpp:71-104)
                    if (!P.left[p]) {
                        P.left[p] = 1;
                        const unsigned long long before = atomicAdd(st.leftGrid, 1ull);
                        if (g.maxLeaving > 0 && before + 1ull >= (unsigned long long)g.maxLeaving)
                            stop_run(B, STOP_TOO_MANY_LEFT, (long long)before);
                        else if (g.deleteLeaving)
                            stop_run(B, STOP_DELETE, (long long)before);
                    } else if (g.deleteLeaving) {
                        stop_run(B, STOP_DELETE, (long long)(*st.leftGrid));
                    }
                    // ReturnToElement (ResetElementsTask.cpp:232-265)
                    double dtt = g.timestep;
                    double outside[3] = {pos[0], pos[1], pos[2]};
                    double inside[3];
                    inside[0] = outside[0] - dtt * P.f[MPF_VEL_X][p];
                    inside[1] = outside[1] - dtt * P.f[MPF_VEL_Y][p];
                    inside[2] = DIM == 3 ? outside[2] - dtt * P.f[MPF_VEL_Z][p] : 0.;
                    if (!pt_in_elem<DIM>(g, ei, ej, ek, inside)) {
                        inside[0] = ((g.xmin + (double)ei * g.dx) + (g.xmin + (double)(ei + 1) * g.dx)) / 2.;
                        inside[1] = ((g.ymin + (double)ej * g.dy) + (g.ymin + (double)(ej + 1) * g.dy)) / 2.;
                        inside[2] = DIM == 3 ? ((g.zmin + (double)ek * g.dz) + (g.zmin + (double)(ek + 1) * g.dz)) / 2. : 0.;
                    }
                    for (int pass = 1; pass <= 10; pass++) {
                        double mid[3] = {(outside[0] + inside[0]) / 2., (outside[1] + inside[1]) / 2.,
                                         (outside[2] + inside[2]) / 2.};
                        if (pt_in_elem<DIM>(g, ei, ej, ek, mid)) { inside[0] = mid[0]; inside[1] = mid[1]; inside[2] = mid[2]; }
                        else { outside[0] = mid[0]; outside[1] = mid[1]; outside[2] = mid[2]; }
                    }
                    P.posn[0][p] = inside[0];
                    P.posn[1][p] = inside[1];
                    if (DIM == 3) P.posn[2][p] = inside[2];
                }
            }
            int key = cell_key<DIM>(g, fld, ei, ej, DIM == 3 ? ek : 0);
            B.key[p] = key;
            wave_count(B.count, key);
        }
    };
    for_each_tile(B, tile);
}

// ---------------------------------------------------------------------------------------
// Multimaterial contact extrapolations (NodalPoint::AddMassMomentum / AddMassMomentumLast extras,
// NodalPointMPM.cpp:428-437,477-485 -> CrackVelocityField::AddVolumeDisplacement,
// CrackVelocityField.cpp:395-402, CrackVelocityFieldMulti::AddVolumeGradient, :108-112): per field,
// volume sum N Vp, displacement sum (x - x0) N mp and volume gradient sum grad N Vp, with
// Vp = det(F) mp / rho (MatPoint3D::GetVolume, MatPoint2D::GetVolume) and the start-of-step
// shape functions. The displacement uses the new positions in the USL pass.
struct RecContact { double wx, wy, wz, dx, dy, dz, vp, mp, ux, uy, uz; };

template <int DIM, int SH>
__global__ __launch_bounds__(NT) void k_p2g_contact(KGrid g0, KPart P, KBins B, KNodes N, const MpmMaterial *mats,
                                                    int newPositions) {
    using G = TileGeom<DIM>;
    constexpr bool CP = SH == MPMGPU_SHAPE_LCPDI;
    constexpr int NCOR = DIM == 3 ? 8 : 4;
    constexpr int CH = CP ? NT / NCOR : NT / 2;   // lCPDI: a thread per particle corner
    __shared__ int s_cs[G::NC + 1];
    __shared__ double s_S[CP ? 1 : CH][DIM][4];
    __shared__ double s_dS[CP ? 1 : CH][DIM][4];
    __shared__ double s_ct[CP ? CH : 1][NCOR][DIM][4];
    __shared__ double s_cwg[CP ? CH : 1][NCOR][3];
    __shared__ double s_q[CH][5];              // Vp, mp, displacement
    __shared__ double s_acc[NW][NQ_CONTACT][WaveFoot<DIM>::SIZE];
    __shared__ double s_id[DIM][G::TX];
    auto tile = [&](const int tall) {      // (field, context tile) pair
        const int fld = tall / g0.tilesTotal;
        const KGrid g = tile_grid(g0, tall - fld * g0.tilesTotal);
        const int t = tall - fld * g0.tilesTotal - g.tileBase;
        load_tile_starts<DIM>(B, tall, s_cs);
        const int tb = s_cs[0], te = s_cs[G::NC];
        if (tb == te) return;
        int ox, oy, oz;
        tile_origin<DIM>(g, t, ox, oy, oz);
        if (threadIdx.x < G::TX) {
            int i = threadIdx.x;
            s_id[0][i] = 2.0 / ((g.xmin + (double)(ox + i + 1) * g.dx) - (g.xmin + (double)(ox + i) * g.dx));
            if (i < G::TY) s_id[1][i] = 2.0 / ((g.ymin + (double)(oy + i + 1) * g.dy) - (g.ymin + (double)(oy + i) * g.dy));
            if (DIM == 3 && i < G::TZ)
                s_id[DIM - 1][i] = 2.0 / ((g.zmin + (double)(oz + i + 1) * g.dz) - (g.zmin + (double)(oz + i) * g.dz));
        }
        zero_acc<DIM, NQ_CONTACT>(s_acc);
        // staging thread: particle si of the chunk, and (lCPDI) its corner sc; the particle's own
        // quantities are staged by its first thread
        const int si = CP ? threadIdx.x / NCOR : threadIdx.x, sc = CP ? threadIdx.x % NCOR : 0;
        for (int c0 = tb; c0 < te; c0 += CH) {
            const int k = c0 + si;
            if (si < CH && k < te) {
                const int p = B.order[k];
                int ei, ej, ek;
                elem_ijk(g, P.elem[p], ei, ej, ek);
                double pos[3] = {P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p], DIM == 3 ? P.f[MPF_POS_Z][p] : 0.};
                if (CP) {
                    double r[9];
#pragma unroll
                    for (int a = 0; a < 9; a++) r[a] = P.cpr[a][p];
                    double xi[3], w[3];
                    int off[3];
                    cpdi_corner<DIM>(g, pos, r, ei, ej, ek, sc, xi, off);
                    cpdi_table<DIM>(xi, off, s_ct[si][sc]);
                    cpdi_wg_staged<DIM>(r, sc, w);
                    s_cwg[si][sc][0] = w[0]; s_cwg[si][sc][1] = w[1]; s_cwg[si][sc][2] = w[2];
                } else {
                    double xi[3], invd[3], S[3][4], dS[3][4];
                    particle_xi<DIM>(g, ei, ej, ek, pos[0], pos[1], pos[2], xi, invd);
                    particle_factors<DIM>(g, P, p, xi, S, dS);
#pragma unroll
                    for (int a = 0; a < DIM; a++)
#pragma unroll
                        for (int o = 0; o < 4; o++) {
                            s_S[threadIdx.x][a][o] = S[a][o];
                            s_dS[threadIdx.x][a][o] = dS[a][o];
                        }
                }
                if (sc == 0) {
                    double mp = P.f[MPF_MASS][p];
                    double rho = mats[P.mat[p]].rho;
                    double Vp;
                    if (DIM == 3) {
                        double Fm[3][3];
                        double ep[6] = {P.f[MPF_EP_XX][p], P.f[MPF_EP_YY][p], P.f[MPF_EP_ZZ][p],
                                        P.f[MPF_EP_YZ][p], P.f[MPF_EP_XZ][p], P.f[MPF_EP_XY][p]};
                        double wr[3] = {P.f[MPF_WR_YZ][p], P.f[MPF_WR_XZ][p], P.f[MPF_WR_XY][p]};
                        get_F3(ep, wr, Fm);
                        Vp = (Fm[0][0] * (Fm[1][1] * Fm[2][2] - Fm[2][1] * Fm[1][2])
                            - Fm[0][1] * (Fm[1][0] * Fm[2][2] - Fm[2][0] * Fm[1][2])
                            + Fm[0][2] * (Fm[1][0] * Fm[2][1] - Fm[2][0] * Fm[1][1])) * mp / rho;
                    } else {
                        double exy = P.f[MPF_EP_XY][p], wxy = P.f[MPF_WR_XY][p];
                        double f00 = 1. + P.f[MPF_EP_XX][p], f11 = 1. + P.f[MPF_EP_YY][p];
                        double f01 = 0.5 * (exy - wxy), f10 = 0.5 * (exy + wxy);
                        Vp = (f00 * f11 - f10 * f01) * mp / rho;
                    }
                    s_q[si][0] = Vp;
                    s_q[si][1] = mp;
#pragma unroll
                    for (int a = 0; a < 3; a++)
                        s_q[si][2 + a] = a < DIM ? (newPositions ? P.posn[a][p] : pos[a]) - P.f[MPF_ORIG_X + a][p] : 0.;
                }
            }
            __syncthreads();
            wave_cells<DIM, NQ_CONTACT>(s_cs, c0, min(c0 + CH, te), s_acc,
                [&](int lq, int ix, int iy, int iz) {
                    RecContact r;
                    if (CP) {
                        double gr[3];
                        cpdi_table_node<DIM, true>(s_ct[lq], s_cwg[lq], ix, iy, iz, r.wx, gr);
                        r.wy = r.wz = 1.;
                        r.dx = gr[0]; r.dy = gr[1]; r.dz = gr[2];
                    } else {
                        r.wx = s_S[lq][0][ix]; r.wy = s_S[lq][1][iy];
                        r.dx = s_dS[lq][0][ix]; r.dy = s_dS[lq][1][iy];
                        if (DIM == 3) { r.wz = s_S[lq][DIM - 1][iz]; r.dz = s_dS[lq][DIM - 1][iz]; }
                        else { r.wz = 1.; r.dz = 0.; }
                    }
                    r.vp = s_q[lq][0]; r.mp = s_q[lq][1];
                    r.ux = s_q[lq][2]; r.uy = s_q[lq][3]; r.uz = s_q[lq][4];
                    return r;
                },
                [&](const RecContact &r, double *acc, int lx, int ly, int lz) {
                    bool on = r.wx >= 0. && r.wy >= 0. && r.wz >= 0.;
                    double vp = on ? r.vp : 0., mp = on ? r.mp : 0.;
                    double fn, gx, gy, gz;
                    if (CP) {
                        fn = r.wx; gx = r.dx; gy = r.dy; gz = DIM == 3 ? r.dz : 0.;
                    } else if (DIM == 3) {
                        fn = r.wx * r.wy * r.wz;
                        gx = r.dx * r.wy * r.wz * s_id[0][lx];
                        gy = r.wx * r.dy * r.wz * s_id[1][ly];
                        gz = r.wx * r.wy * r.dz * s_id[DIM - 1][lz];
                    } else {
                        fn = r.wx * r.wy;
                        gx = r.dx * r.wy * s_id[0][lx];
                        gy = r.wx * r.dy * s_id[1][ly];
                        gz = 0.;
                    }
                    acc[0] = fma(fn, vp, acc[0]);
                    const double fnmp = fn * mp;
                    acc[1] = fma(r.ux, fnmp, acc[1]);
                    acc[2] = fma(r.uy, fnmp, acc[2]);
                    acc[3] = fma(r.uz, fnmp, acc[3]);
                    acc[4] = fma(gx, vp, acc[4]);
                    acc[5] = fma(gy, vp, acc[5]);
                    acc[6] = fma(gz, vp, acc[6]);
                });
            __syncthreads();
        }
        write_partials<DIM, NQ_CONTACT>(N, tall, s_acc);
    };
    for_each_tile(B, tile);
}

template <int DIM>
__global__ void k_node_contact(KGrid g0, KBins B, KNodes N) {
    long long nall = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nall >= g0.nstride * g0.nfields) return;
    const int fld = (int)(nall / g0.nstride);
    const long long ng = nall - (long long)fld * g0.nstride;
    const KGrid g = node_grid(g0, ng);
    const long long n = ng - g.nodeBase;
    if (n >= g.nnodes) return;                // padding between instances
    const KNodes Nf = field_view(g, N, fld);
    int i = (int)(n % g.nnx);
    long long r = n / g.nnx;
    int j = (int)(r % g.nny);
    int k = (int)(r / g.nny);
    if (!node_needed<DIM>(g, B, fld, i, j, k)) return;
    double a[NQ_CONTACT];
    gather_partials<DIM, NQ_CONTACT>(g, B, N, fld, i, j, k, a);
    long long nn = g.nstride;
    Nf.cvol[n] = a[0];
    Nf.cdisp[n] = a[1]; Nf.cdisp[nn + n] = a[2]; Nf.cdisp[2 * nn + n] = a[3];
    Nf.cgrad[n] = a[4]; Nf.cgrad[nn + n] = a[5]; Nf.cgrad[2 * nn + n] = a[6];
}

__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// CoulombFriction::GetFrictionalDeltaMomentum and GetSslideAcDt (CoulombFriction.cpp:146-282);
// false leaves the momenta unchanged
__device__ __forceinline__ bool coulomb_delta(const KGrid &g, double *delPi, const double *norm, double dotn,
                                              double deltaDotn, double mred, const double *delFi) {
    double coeff = g.friction, coeffStatic = g.frictionStatic;
    int style = coeff < 0. ? 0 : (coeff == 0. ? 1 : 2);      // STICK, FRICTIONLESS, FRICTIONAL
    if (coeffStatic > 0. && style == 1) style = 2;
    if (style == 0) return true;
    double deltime = g.timestep;
    if (delFi) {
        double fn = dot3(delFi, norm);
        deltaDotn += deltime * (dotn - 0.5 * fn * deltime) / mred;
    }
    bool inContact = deltaDotn < 0. && dotn < 0.;
    if (style == 1) {
        if (inContact) { for (int c = 0; c < 3; c++) delPi[c] = norm[c] * dotn; return true; }
        return false;
    }
    if (!inContact) return false;
    double dott = 0., tang[3];
    for (int c = 0; c < 3; c++) tang[c] = delPi[c];
    for (int c = 0; c < 3; c++) tang[c] += norm[c] * (-dotn);
    double tangMag = sqrt(dot3(tang, tang));
    if (tangMag > 0.) {
        double sc = 1. / tangMag;
        for (int c = 0; c < 3; c++) tang[c] *= sc;
        dott = dot3(delPi, tang);
        if (dott < 0.) { for (int c = 0; c < 3; c++) tang[c] *= -1.; dott = -dott; }
    }
    double NAcDt = -dotn, SslideAcDt;
    if (coeffStatic > 0. && dott <= coeffStatic * NAcDt) SslideAcDt = coeffStatic * NAcDt;
    else SslideAcDt = coeff * NAcDt;
    if (SslideAcDt <= 0.) {
        for (int c = 0; c < 3; c++) delPi[c] = norm[c] * dotn;
    } else if (dott > SslideAcDt) {
        for (int c = 0; c < 3; c++) delPi[c] = norm[c] * dotn;
        for (int c = 0; c < 3; c++) delPi[c] += tang[c] * SslideAcDt;
    }
    return true;
}

// MatVelocityField::ChangeMatMomentum (MatVelocityField.cpp:197-224) on field f of node n, then
// the field's velocity for the next gather (GridValueCalculation, :239-251)
__device__ __forceinline__ void change_momentum(const KGrid &g, const KNodes &N, int f, long long n,
                                                const double *delP, int callType) {
    long long nn = g.nstride, b = 3 * f * nn + n;
    double p[3];
    for (int c = 0; c < 3; c++) { p[c] = N.pk[b + c * nn] + delP[c]; N.pk[b + c * nn] = p[c]; }
    if (callType == CONTACT_UPDATE_MOMENTUM) {
        double sc = 1. / g.timestep;
        for (int c = 0; c < 3; c++) N.ftot[b + c * nn] += delP[c] * sc;
    }
    double m = N.mass[f * nn + n];
    if (N.cnt[f * nn + n] > 0 && m != 0.) {
        double s = 1. / m;
        for (int c = 0; c < 3; c++) N.vk[b + c * nn] = s * p[c];
    }
}

// MaterialContactNode::ContactOnKnownNodes -> CrackVelocityFieldMulti::MaterialContactOnCVFLumped
// (CrackVelocityFieldMulti.cpp:302-671) for non-rigid materials, contact by displacements, normals
// by GetNormalVector (:960-1069) methods 0-3, no symmetry planes. One thread per node; nodes with
// fewer than two materials return at once.
template <int DIM>
__global__ void k_contact(KGrid g, KNodes N, int callType) {
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nnodes) return;
    const long long nn = g.nstride;
    int act[MPMGPU_MAX_MATERIALS];
    int numMats = 0;
    for (int f = 0; f < g.nfields; f++)
        if (N.cnt[f * nn + n] > 0) act[numMats++] = f;
    if (numMats <= 1) return;
    double Pc[3] = {0., 0., 0.}, Mc = 0.;
    for (int a = 0; a < numMats; a++) {
        long long b = 3 * act[a] * nn + n;
        for (int c = 0; c < 3; c++) Pc[c] += N.pk[b + c * nn];
        Mc += N.mass[act[a] * nn + n];
    }
    const int normals = g.normals;
    const bool doingPairs = numMats == 2 && normals != 3;
    const int miMax = doingPairs ? numMats - 1 : numMats;
    const bool postUpdate = callType == CONTACT_UPDATE_MOMENTUM;
    double dispc[3] = {0., 0., 0.};
    for (int a = 0; a < numMats; a++)
        for (int c = 0; c < 3; c++) dispc[c] += N.cdisp[(3 * act[a] + c) * nn + n];
    double sMc = 1. / Mc;
    for (int c = 0; c < 3; c++) dispc[c] *= sMc;
    double volAll = 0.;
    for (int a = 0; a < numMats; a++) volAll += N.cvol[act[a] * nn + n];
    for (int mi = 0; mi < miMax; mi++) {
        const int i = act[mi];
        double massi = N.mass[i * nn + n];
        double voli = N.cvol[i * nn + n];
        double volj = volAll - voli;
        int j = -1;
        double gradj[3] = {0., 0., 0.};
        double maxOther = 0.;
        for (int kj = 0; kj < numMats; kj++) {
            if (kj == mi) continue;
            int jj = act[kj];
            double mv = N.cvol[jj * nn + n];
            if (mv > maxOther) { maxOther = mv; j = jj; }
            for (int c = 0; c < 3; c++) gradj[c] += N.cgrad[(3 * jj + c) * nn + n] * -1.;
        }
        if (j < 0) continue;
        double massRatio = massi / Mc;
        double mred = 1. - massRatio;
        double delPi[3];
        for (int c = 0; c < 3; c++) delPi[c] = N.pk[(3 * i + c) * nn + n] * -1.;
        for (int c = 0; c < 3; c++) delPi[c] += Pc[c] * massRatio;
        if (!g.ignoreContact) {
            double gradi[3], norm[3];
            for (int c = 0; c < 3; c++) gradi[c] = N.cgrad[(3 * i + c) * nn + n] * 1.;
            if (normals == 0) {
                double magi = sqrt(dot3(gradi, gradi)), magj = sqrt(dot3(gradj, gradj));
                if (magi >= 1. * magj) { double sc = 1. / magi; for (int c = 0; c < 3; c++) norm[c] = gradi[c] * sc; }
                else { double sc = 1. / magj; for (int c = 0; c < 3; c++) norm[c] = gradj[c] * sc; }
            } else if (normals == 1) {
                if (voli >= volj) for (int c = 0; c < 3; c++) norm[c] = gradi[c];
                else for (int c = 0; c < 3; c++) norm[c] = gradj[c];
                double sc = 1. / sqrt(dot3(norm, norm));
                for (int c = 0; c < 3; c++) norm[c] = norm[c] * sc;
            } else if (normals == 2) {
                for (int c = 0; c < 3; c++) norm[c] = gradi[c];
                for (int c = 0; c < 3; c++) norm[c] += gradj[c];
                double sc = 1. / sqrt(dot3(norm, norm));
                for (int c = 0; c < 3; c++) norm[c] *= sc;
            } else if (normals == 3) {
                for (int c = 0; c < 3; c++) norm[c] = gradi[c];
                double sc = 1. / sqrt(dot3(norm, norm));
                for (int c = 0; c < 3; c++) norm[c] *= sc;
            } else if (normals == 4) {            // SPECIFIED_NORMAL (no symmetry planes)
                for (int c = 0; c < 3; c++) norm[c] = g.contactNormal[c];
            } else {                              // regression: k_contact_normals found it
                for (int c = 0; c < 3; c++) norm[c] = N.cnorm[(4 * i + c) * nn + n];
            }
            if (norm[0] != norm[0] || norm[1] != norm[1] || norm[2] != norm[2]) continue;
            double dotn = dot3(delPi, norm);
            double dispa[3], sa = 1. / massi;
            for (int c = 0; c < 3; c++) dispa[c] = N.cdisp[(3 * i + c) * nn + n] * sa;
            double delta[3], sr = 1. / mred;
            for (int c = 0; c < 3; c++) delta[c] = dispc[c];
            for (int c = 0; c < 3; c++) delta[c] -= dispa[c];
            for (int c = 0; c < 3; c++) delta[c] *= sr;
            double dispb[3];
            for (int c = 0; c < 3; c++) dispb[c] = delta[c];
            for (int c = 0; c < 3; c++) dispb[c] += dispa[c];
            // the regression normals come with the separation of the point cloud
            double deln = normals >= 5 ? N.cnorm[(4 * i + 3) * nn + n] : dot3(dispb, norm) - dot3(dispa, norm);
            mred *= massi;
            double delFi[3];
            if (postUpdate) {
                double fs[3] = {0., 0., 0.};
                for (int a = 0; a < numMats; a++)
                    for (int c = 0; c < 3; c++) fs[c] += N.ftot[(3 * act[a] + c) * nn + n];
                double sc = massi / Mc;
                for (int c = 0; c < 3; c++) delFi[c] = fs[c] * sc;
                for (int c = 0; c < 3; c++) delFi[c] += N.ftot[(3 * i + c) * nn + n] * -1.;
            }
            // the separation's momentum-update term is skipped for regression normals (CoulombFriction.cpp:162)
            if (!coulomb_delta(g, delPi, norm, dotn, deln, mred, postUpdate && normals < 5 ? delFi : nullptr)) continue;
        }
        change_momentum(g, N, i, n, delPi, callType);
        if (doingPairs) {
            for (int c = 0; c < 3; c++) delPi[c] *= -1.;
            change_momentum(g, N, j, n, delPi, callType);
        }
    }
}

// ---------------------------------------------------------------------------------------
// Regression contact normals (normals methods 5 and 6): CrackVelocityFieldMulti::
// LinearRegressionNormal and LogisticRegressionNormal (Nodes/MachineLearning.cpp:58-633) over the
// particles that see each contact node (MaterialContactNode::GetMaterialContactData,
// MaterialContactNode.cpp:253-350). A wave takes one contact node at a time: it lists the node's
// particles in LDS (cells and fields in a fixed order, particles of a cell in upload order), then
// the lanes split the list for every sum of the fit and a fixed reduction tree adds their parts,
// so each node's result is the same on every run.

// contact nodes of the step: two or more fields with particles (CalcTotalMassAndCount,
// PostExtrapolationTask.cpp:62-80); the count is cleared by the mass node pass before
template <int DIM>
__global__ void k_contact_list(KGrid g, KNodes N) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool on = false;
    if (n < g.nnodes) {
        int numMats = 0;
        for (int f = 0; f < g.nfields; f++) numMats += N.cnt[f * g.nstride + n] > 0 ? 1 : 0;
        on = numMats > 1;
    }
    const unsigned long long m = __ballot(on);
    if (m == 0ull) return;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&N.clist[0], __popcll(m));
    base = __shfl(base, leader);
    if (on) N.clist[1 + base + __popcll(m & ((1ull << lane) - 1ull))] = (int)n;
}

// node (i,j,k) in particle p's shape-function node list (start-of-step shape functions):
// uGIMP/classic: inside the support on every axis; lCPDI: a corner's function at the node is
// at least 1e-15 (GetCPDIFunctions, MoreMPMElementBase.cpp:581-656)
template <int DIM, int SH>
__device__ __forceinline__ bool lr_member(const KGrid &g, const KPart &P, int p, int i, int j, int k) {
    int ei, ej, ek;
    elem_ijk(g, P.elem[p], ei, ej, ek);
    const int o[3] = {i - ei + 1, j - ej + 1, DIM == 3 ? k - ek + 1 : 1};
#pragma unroll
    for (int a = 0; a < DIM; a++)
        if (o[a] < 0 || o[a] > 3) return false;
    if (SH == MPMGPU_SHAPE_LCPDI) {
        CpdiCorners cc;
        cpdi_particle<DIM>(g, P, p, ei, ej, ek, cc);
        constexpr int NCOR = DIM == 3 ? 8 : 4;
        constexpr double thr = (DIM == 3 ? 8. : 4.) * 1e-15;
        bool in = false;
#pragma unroll
        for (int c = 0; c < NCOR; c++) {
            double T[DIM][4];
            cpdi_table<DIM>(cc.xi[c], cc.off[c], T);
            const double f = DIM == 3 ? T[0][o[0]] * T[1][o[1]] * T[2][o[2]] : T[0][o[0]] * T[1][o[1]];
            in |= f >= thr;
        }
        return in;
    }
    double xi[3], invd[3], S[3][4], dS[3][4];
    particle_xi<DIM>(g, ei, ej, ek, P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p], DIM == 3 ? P.f[MPF_POS_Z][p] : 0., xi, invd);
    particle_factors<DIM>(g, P, p, xi, S, dS);
    bool in = true;
#pragma unroll
    for (int a = 0; a < DIM; a++) in &= S[a][o[a]] >= 0.;
    return in;
}

// sum over the wave in a fixed tree, the total in every lane
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off);
    return __shfl(v, 0);
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_down(v, off));
    return __shfl(v, 0);
}

// Matrix4::cofactor, determinant, Inverse and Times (Matrix4.cpp:96-103,133-140,206-231): the
// inverse as signed cofactors scaled by 1/det, then times rhs
__device__ __forceinline__ double lr_cofactor4(const double m[4][4], int row, int col) {
    double c[3][3];
    int cr = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (i == row) continue;
        int cc = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j == col) continue;
            c[cr][cc++] = m[i][j];
        }
        cr++;
    }
    return c[0][0] * (c[1][1] * c[2][2] - c[2][1] * c[1][2]) - c[1][0] * (c[0][1] * c[2][2] - c[2][1] * c[0][2]) +
           c[2][0] * (c[0][1] * c[1][2] - c[1][1] * c[0][2]);
}
__device__ __forceinline__ void lr_solve4(const double m[4][4], const double rhs[4], double out[4]) {
    const double det = m[0][0] * lr_cofactor4(m, 0, 0) - m[1][0] * lr_cofactor4(m, 1, 0) +
                       m[2][0] * lr_cofactor4(m, 2, 0) - m[3][0] * lr_cofactor4(m, 3, 0);
    const double f = 1. / det;
    double inv[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) inv[r][c] = ((r + c) & 1 ? -lr_cofactor4(m, c, r) : lr_cofactor4(m, c, r)) * f;
#pragma unroll
    for (int r = 0; r < 4; r++) out[r] = inv[r][0] * rhs[0] + inv[r][1] * rhs[1] + inv[r][2] * rhs[2] + inv[r][3] * rhs[3];
}
// symmetric Matrix3 (Matrix3.cpp:107-121) Inverse (3D form, :446-460) times rhs (:394-408)
__device__ __forceinline__ void lr_solve3(const double s[6], const double rhs[3], double out[3]) {
    const double m[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], s[5]}};
    double inv[3][3] = {{m[1][1] * m[2][2] - m[2][1] * m[1][2], -(m[0][1] * m[2][2] - m[2][1] * m[0][2]),
                         m[0][1] * m[1][2] - m[1][1] * m[0][2]},
                        {-(m[1][0] * m[2][2] - m[2][0] * m[1][2]), m[0][0] * m[2][2] - m[2][0] * m[0][2],
                         -(m[0][0] * m[1][2] - m[1][0] * m[0][2])},
                        {m[1][0] * m[2][1] - m[2][0] * m[1][1], -(m[0][0] * m[2][1] - m[2][0] * m[0][1]),
                         m[0][0] * m[1][1] - m[1][0] * m[0][1]}};
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                       m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) + m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    const double f = 1. / det;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) inv[r][c] *= f;
#pragma unroll
    for (int r = 0; r < 3; r++) out[r] = inv[r][0] * rhs[0] + inv[r][1] * rhs[1] + inv[r][2] * rhs[2];
}

// MatPoint3D/MatPoint2D::GetDeformedRadius, ellipse methods (MatPoint3D.cpp:497-510,
// MatPoint2D.cpp:482-497): the distance from the particle's center to its deformed domain's
// surface along the unit vector n (2D applies F^-T where 3D applies F^-1, as the reference does)
template <int DIM>
__device__ __forceinline__ double lr_radius(const KGrid &g, const KPart &P, int p, const double n[3]) {
    double lpx, lpy, lpz;
    if (g.uniformLp) { lpx = g.lp[0]; lpy = g.lp[1]; lpz = g.lp[2]; }
    else { lpx = P.f[MPF_LP_X][p]; lpy = P.f[MPF_LP_Y][p]; lpz = P.f[MPF_LP_Z][p]; }
    const double psx = g.dx * (0.5 * lpx), psy = g.dy * (0.5 * lpy);      // GetParticleSize
    if (DIM == 2) {
        const double exx = P.f[MPF_EP_XX][p], eyy = P.f[MPF_EP_YY][p], exy = P.f[MPF_EP_XY][p], wxy = P.f[MPF_WR_XY][p];
        const double F00 = 1. + exx, F01 = 0.5 * (exy - wxy), F10 = 0.5 * (exy + wxy), F11 = 1. + eyy;
        const double rdet2D = 1. / (F00 * F11 - F01 * F10);
        const double n0x = (F11 * n[0] - F10 * n[1]) * rdet2D;
        const double n0y = (-F01 * n[0] + F00 * n[1]) * rdet2D;
        return 1. / sqrt(n0x * n0x / (psx * psx) + n0y * n0y / (psy * psy));
    }
    const double psz = g.dz * (0.5 * lpz);
    double ep[6], wr[3], m[3][3];
#pragma unroll
    for (int c = 0; c < 6; c++) ep[c] = P.f[MPF_EP_XX + c][p];
#pragma unroll
    for (int c = 0; c < 3; c++) wr[c] = P.f[MPF_WR_YZ + c][p];
    get_F3(ep, wr, m);
    double inv[3][3] = {{m[1][1] * m[2][2] - m[2][1] * m[1][2], -(m[0][1] * m[2][2] - m[2][1] * m[0][2]),
                         m[0][1] * m[1][2] - m[1][1] * m[0][2]},
                        {-(m[1][0] * m[2][2] - m[2][0] * m[1][2]), m[0][0] * m[2][2] - m[2][0] * m[0][2],
                         -(m[0][0] * m[1][2] - m[1][0] * m[0][2])},
                        {m[1][0] * m[2][1] - m[2][0] * m[1][1], -(m[0][0] * m[2][1] - m[2][0] * m[0][1]),
                         m[0][0] * m[1][1] - m[1][0] * m[0][1]}};
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                       m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) + m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    const double f = 1. / det;
    double n0[3];
#pragma unroll
    for (int r = 0; r < 3; r++) n0[r] = (inv[r][0] * f) * n[0] + (inv[r][1] * f) * n[1] + (inv[r][2] * f) * n[2];
    return 1. / sqrt(n0[0] * n0[0] / (psx * psx) + n0[1] * n0[1] / (psy * psy) + n0[2] * n0[2] / (psz * psz));
}

#define LR_MAX 1024        // particles listed at one contact node (a wave's LDS list)

template <int DIM, int SH>
__global__ __launch_bounds__(NT) void k_contact_normals(KGrid g, KPart P, KBins B, KNodes N, const MpmMaterial *mats,
                                                        KStats st, int newPos) {
    __shared__ int s_list[NW][LR_MAX];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int *list = s_list[w];
    const int count = N.clist[0];
    const long long nn = g.nstride;
    const double *px = newPos ? P.posn[0] : P.f[MPF_POS_X];
    const double *py = newPos ? P.posn[1] : P.f[MPF_POS_Y];
    const double *pz = DIM == 3 ? (newPos ? P.posn[2] : P.f[MPF_POS_Z]) : nullptr;
    const bool logistic = g.normals == 6;
    for (int e = blockIdx.x * NW + w; e < count; e += gridDim.x * NW) {
        const long long n = N.clist[1 + e];
        const int i = (int)(n % g.nnx), j = (int)((n / g.nnx) % g.nny), k = DIM == 3 ? (int)(n / ((long long)g.nnx * g.nny)) : 0;
        const double nd[3] = {g.xmin + (double)i * g.dx, g.ymin + (double)j * g.dy, DIM == 3 ? g.zmin + (double)k * g.dz : 0.};
        int act[MPMGPU_MAX_MATERIALS];
        int numMats = 0;
        for (int f = 0; f < g.nfields; f++)
            if (N.cnt[f * nn + n] > 0) act[numMats++] = f;
        // the particles seeing the node: cells i-2..i+1 (per axis) of every field, in (field, z, y, x)
        // cell order and sorted order inside a cell. Lane q < 4^DIM holds cell q's slot range; a wave
        // prefix sum of the counts lays the field's particles out one after another, so every pass
        // tests 64 of them whatever the cells hold (the list comes out in the same order).
        constexpr int NCELL = DIM == 3 ? 64 : 16;
        int nl = 0;
        for (int f = 0; f < g.nfields; f++) {
            int cb = 0, cn = 0;
            if (lane < NCELL) {
                const int cx = i - 2 + (lane & 3), cy = j - 2 + ((lane >> 2) & 3);
                const int cz = DIM == 3 ? k - 2 + (lane >> 4) : 0;
                if (cx >= 0 && cx < g.nx && cy >= 0 && cy < g.ny && cz >= 0 && cz < g.nz) {
                    const int key = cell_key<DIM>(g, f, cx, cy, cz);
                    cb = B.cellStart[key];
                    cn = B.cellStart[key + 1] - cb;
                }
            }
            int incl = cn;                                   // inclusive prefix over the lanes
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int u = __shfl_up(incl, d);
                if (lane >= d) incl += u;
            }
            const int total = __shfl(incl, 63);
            for (int t0 = 0; t0 < total; t0 += 64) {
                const int t = t0 + lane;
                // the cell of flattened index t: the first lane whose inclusive prefix exceeds t
                int lo = 0, hi = NCELL - 1;
#pragma unroll
                for (int it = 0; it < (DIM == 3 ? 6 : 4); it++) {
                    const int mid = (lo + hi) >> 1;
                    if (__shfl(incl, mid) > t) hi = mid;
                    else lo = mid + 1;
                }
                const int qb = __shfl(cb, lo), qe = __shfl(incl - cn, lo);
                const int p = t < total ? B.order[qb + (t - qe)] : -1;
                const bool in = p >= 0 && lr_member<DIM, SH>(g, P, p, i, j, DIM == 3 ? k : 0);
                const unsigned long long m = __ballot(in);
                if (in) {
                    const int idx = nl + __popcll(m & ((1ull << lane) - 1ull));
                    if (idx < LR_MAX) list[idx] = p;
                }
                nl += __popcll(m);
            }
        }
        if (nl > LR_MAX) {
            if (lane == 0) atomicAdd(st.lrOverflow, 1ull);   // a fatal error, reported after the step
            nl = LR_MAX;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int miMax = numMats == 2 ? 1 : numMats;
        for (int mi = 0; mi < miMax; mi++) {
            const int a = act[mi];
            double norm[3] = {0., 0., 0.}, linNorm[3] = {0., 0., 0.};
            const double cell = fmin(g.dx, fmin(g.dy, DIM == 3 ? g.dz : g.dy));     // GetMinCellDimension
            const double normPenalty = 1.e-7 * (cell * cell);
            if (DIM == 3) {
                double mt[10], rhs[4], lambda[4];
                if (logistic) { lambda[0] = lambda[1] = lambda[2] = normPenalty; lambda[3] = 0.; }
                else { lambda[0] = lambda[1] = lambda[2] = 0.; lambda[3] = 0.1; }
                for (int q = 0; q < 10; q++) mt[q] = 0.;
                for (int q = 0; q < 4; q++) rhs[q] = 0.;
                for (int t = lane; t < nl; t += 64) {
                    const int p = list[t];
                    const double mi_ = mats[P.mat[p]].field == a ? -1. : 1.;
                    const double x = px[p] - nd[0], y = py[p] - nd[1], z = pz[p] - nd[2];
                    mt[0] += x * x; mt[1] += x * y; mt[2] += x * z; mt[3] += x;
                    mt[4] += y * y; mt[5] += y * z; mt[6] += y;
                    mt[7] += z * z; mt[8] += z; mt[9] += 1.;
                    rhs[0] += mi_ * x; rhs[1] += mi_ * y; rhs[2] += mi_ * z; rhs[3] += mi_;
                }
                for (int q = 0; q < 10; q++) mt[q] = wave_sum(mt[q]);
                for (int q = 0; q < 4; q++) rhs[q] = wave_sum(rhs[q]);
                auto mat4 = [&](double M[4][4]) {
                    M[0][0] = mt[0] + lambda[0]; M[0][1] = mt[1]; M[0][2] = mt[2]; M[0][3] = mt[3];
                    M[1][0] = mt[1]; M[1][1] = mt[4] + lambda[1]; M[1][2] = mt[5]; M[1][3] = mt[6];
                    M[2][0] = mt[2]; M[2][1] = mt[5]; M[2][2] = mt[7] + lambda[2]; M[2][3] = mt[8];
                    M[3][0] = mt[3]; M[3][1] = mt[6]; M[3][2] = mt[8]; M[3][3] = mt[9] + lambda[3];
                };
                double M[4][4], beta[4];
                mat4(M);
                lr_solve4(M, rhs, beta);
                if (!logistic) {
                    const double magi = sqrt(beta[0] * beta[0] + beta[1] * beta[1] + beta[2] * beta[2]);
                    for (int c = 0; c < 3; c++) norm[c] = beta[c] * (1. / magi);
                } else {
                    for (int q = 0; q < 4; q++) beta[q] *= 2.;
                    double prevNorm[3] = {beta[0], beta[1], beta[2]};
                    double nmag = sqrt(dot3(prevNorm, prevNorm));
                    for (int c = 0; c < 3; c++) prevNorm[c] *= 1. / nmag;
                    for (int c = 0; c < 3; c++) linNorm[c] = prevNorm[c];
                    double delta = 1.;
                    int iter = 0;
                    while (delta > 1.e-5 && iter < 15) {
                        for (int q = 0; q < 10; q++) mt[q] = 0.;
                        for (int q = 0; q < 4; q++) rhs[q] = 0.;
                        for (int t = lane; t < nl; t += 64) {
                            const int p = list[t];
                            const double mi_ = mats[P.mat[p]].field == a ? -1. : 1.;
                            const double x = px[p] - nd[0], y = py[p] - nd[1], z = pz[p] - nd[2];
                            const double dotp = x * beta[0] + y * beta[1] + z * beta[2] + beta[3];
                            const double arg = 1. + exp(-dotp);
                            const double phi = 2. * (arg - 1.) / (arg * arg);
                            const double phi2 = phi * phi;
                            mt[0] += phi2 * x * x; mt[1] += phi2 * x * y; mt[2] += phi2 * x * z; mt[3] += phi2 * x;
                            mt[4] += phi2 * y * y; mt[5] += phi2 * y * z; mt[6] += phi2 * y;
                            mt[7] += phi2 * z * z; mt[8] += phi2 * z; mt[9] += phi2;
                            const double mifi = (mi_ - 2. / arg + 1.) * phi;
                            rhs[0] += mifi * x; rhs[1] += mifi * y; rhs[2] += mifi * z; rhs[3] += mifi;
                        }
                        for (int q = 0; q < 10; q++) mt[q] = wave_sum(mt[q]);
                        // the penalty term starts the right-hand side (MachineLearning.cpp:327-328)
                        for (int q = 0; q < 4; q++) rhs[q] = -lambda[q] * beta[q] + wave_sum(rhs[q]);
                        mat4(M);
                        double dbeta[4];
                        lr_solve4(M, rhs, dbeta);
                        for (int q = 0; q < 4; q++) beta[q] += dbeta[q];
                        for (int c = 0; c < 3; c++) norm[c] = beta[c];
                        nmag = sqrt(dot3(norm, norm));
                        for (int c = 0; c < 3; c++) norm[c] *= 1. / nmag;
                        delta = 1. - dot3(norm, prevNorm);
                        iter++;
                        for (int c = 0; c < 3; c++) prevNorm[c] = norm[c];
                    }
                }
            } else {
                double mt[6], rhs[3];
                double lambda[3] = {0., 0., 0.1};
                if (logistic) { lambda[0] = normPenalty; lambda[1] = normPenalty; lambda[2] = 0.; }
                for (int q = 0; q < 6; q++) mt[q] = 0.;
                for (int q = 0; q < 3; q++) rhs[q] = 0.;
                for (int t = lane; t < nl; t += 64) {
                    const int p = list[t];
                    const double mi_ = mats[P.mat[p]].field == a ? -1. : 1.;
                    const double x = px[p] - nd[0], y = py[p] - nd[1];
                    mt[0] += x * x; mt[1] += x * y; mt[2] += x;
                    mt[3] += y * y; mt[4] += y; mt[5] += 1.;
                    rhs[0] += mi_ * x; rhs[1] += mi_ * y; rhs[2] += mi_;
                }
                for (int q = 0; q < 6; q++) mt[q] = wave_sum(mt[q]);
                for (int q = 0; q < 3; q++) rhs[q] = wave_sum(rhs[q]);
                double beta[3];
                double S[6] = {mt[0] + lambda[0], mt[1], mt[2], mt[3] + lambda[1], mt[4], mt[5] + lambda[2]};
                lr_solve3(S, rhs, beta);
                if (!logistic) {
                    norm[0] = beta[0]; norm[1] = beta[1];
                    const double magi = sqrt(dot3(norm, norm));
                    for (int c = 0; c < 3; c++) norm[c] *= 1. / magi;
                } else {
                    for (int q = 0; q < 3; q++) beta[q] *= 2.;
                    double prevNorm[3] = {beta[0], beta[1], beta[2]};
                    double nmag = sqrt(beta[0] * beta[0] + beta[1] * beta[1]);
                    prevNorm[0] /= nmag; prevNorm[1] /= nmag;
                    linNorm[0] = prevNorm[0]; linNorm[1] = prevNorm[1];
                    double delta = 1., nv[3] = {0., 0., 0.};
                    int iter = 0;
                    while (delta > 1.e-5 && iter < 15) {
                        for (int q = 0; q < 6; q++) mt[q] = 0.;
                        for (int q = 0; q < 3; q++) rhs[q] = 0.;
                        for (int t = lane; t < nl; t += 64) {
                            const int p = list[t];
                            const double mi_ = mats[P.mat[p]].field == a ? -1. : 1.;
                            const double x = px[p] - nd[0], y = py[p] - nd[1], z = 1.;
                            const double arg = 1. + exp(-(x * beta[0] + y * beta[1] + z * beta[2]));
                            const double phi = 2. * (arg - 1.) / (arg * arg);
                            const double phi2 = phi * phi;
                            mt[0] += phi2 * x * x; mt[1] += phi2 * x * y; mt[2] += phi2 * x;
                            mt[3] += phi2 * y * y; mt[4] += phi2 * y; mt[5] += phi2;
                            const double mifi = (mi_ - 2. / arg + 1.) * phi;
                            rhs[0] += mifi * x; rhs[1] += mifi * y; rhs[2] += mifi;
                        }
                        for (int q = 0; q < 6; q++) mt[q] = wave_sum(mt[q]);
                        for (int q = 0; q < 3; q++) rhs[q] = -lambda[q] * beta[q] + wave_sum(rhs[q]);
                        double S2[6] = {mt[0] + lambda[0], mt[1], mt[2], mt[3] + lambda[1], mt[4], mt[5] + lambda[2]};
                        double dbeta[3];
                        lr_solve3(S2, rhs, dbeta);
                        for (int q = 0; q < 3; q++) beta[q] += dbeta[q];
                        for (int q = 0; q < 3; q++) nv[q] = beta[q];
                        nmag = sqrt(nv[0] * nv[0] + nv[1] * nv[1]);
                        nv[0] /= nmag; nv[1] /= nmag;
                        delta = 1. - nv[0] * prevNorm[0] - nv[1] * prevNorm[1];
                        iter++;
                        for (int q = 0; q < 3; q++) prevNorm[q] = nv[q];
                    }
                    norm[0] = nv[0]; norm[1] = nv[1]; norm[2] = 0.;
                }
            }
            if (logistic && norm[0] != norm[0])
                for (int c = 0; c < 3; c++) norm[c] = linNorm[c];
            // separation from the point cloud (MachineLearning.cpp:195-226, 614-633)
            double supA = 1.e30, supB = 1.e30;
            for (int t = lane; t < nl; t += 64) {
                const int p = list[t];
                const double xp[3] = {px[p], py[p], DIM == 3 ? pz[p] : 0.};
                const double r = lr_radius<DIM>(g, P, p, norm);
                if (mats[P.mat[p]].field == a) supA = fmin(supA, -dot3(norm, xp) - r);
                else supB = fmin(supB, dot3(norm, xp) - r);
            }
            supA = wave_min(supA);
            supB = wave_min(supB);
            if (lane == 0) {
#pragma unroll
                for (int c = 0; c < 3; c++) N.cnorm[(4 * a + c) * nn + n] = norm[c];
                N.cnorm[(4 * a + 3) * nn + n] = supA + supB;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------------------
// launchers
static inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_bin_keys(int dim, const KGrid &g, const KPart &p, const KBins &b, hipStream_t s) {
    if (p.n == 0) return;
    if (dim == 3) hipLaunchKernelGGL(k_bin_keys<3>, dim3(blocks_for(p.n, NT)), dim3(NT), 0, s, g, p, b);
    else hipLaunchKernelGGL(k_bin_keys<2>, dim3(blocks_for(p.n, NT)), dim3(NT), 0, s, g, p, b);
}

void launch_bin_finish(int dim, const KGrid &g, const KPart &p, const KBins &b, hipStream_t s) {
    int n = b.nkeys + 1;          // count[nkeys] stays zero, so cellStart[nkeys] = total
    int nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(NT), 0, s, b.count, b.cellStart, b.blockSums, n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(NT), 0, s, b.blockSums, nb, b.ntiles);
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n, NT)), dim3(NT), 0, s, b.cellStart, b.blockSums, n);
    if (p.n > 0) hipLaunchKernelGGL(k_bin_scatter, dim3(blocks_for(p.n, NT)), dim3(NT), 0, s, p, b);
    hipLaunchKernelGGL(k_bin_sort_cells, dim3(blocks_for(b.nkeys, NT)), dim3(NT), 0, s, b);
    hipLaunchKernelGGL(k_tile_list, dim3(blocks_for(b.ntall, NT)), dim3(NT), 0, s, b,
                       dim == 3 ? TileGeom<3>::NC : TileGeom<2>::NC);
    if (dim == 3) hipLaunchKernelGGL(k_tile_need<3>, dim3(blocks_for(b.ntall, NT)), dim3(NT), 0, s, g, b);
    else hipLaunchKernelGGL(k_tile_need<2>, dim3(blocks_for(b.ntall, NT)), dim3(NT), 0, s, g, b);
}

static int g_cus = 256;
void set_tile_cus(int cus) { g_cus = cus > 0 ? cus : 256; }

// workgroups of one tile-kernel launch: eight per CU (at least as many as the device holds at once
// for any of the tile kernels), at most one per (field, tile) pair
static dim3 tile_grid_dim(const KBins &b) {
    const long long want = 8LL * g_cus;
    return dim3((unsigned)(b.ntall < want ? b.ntall : want));
}
static KBins tile_bins(const KBins &b, int slot) {
    KBins r = b;
    r.slot = slot;
    return r;
}
#define TILE_LAUNCH(KERNEL, B, S, ...) hipLaunchKernelGGL(KERNEL, tile_grid_dim(B), dim3(NT), 0, S, __VA_ARGS__)

void launch_p2g_mass(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                     const MpmMaterial *mats, const KStats &st, hipStream_t s) {
    const KBins tb = tile_bins(b, TILE_SLOT_MASS);
    (void)mats;
    const bool cp = g.shape == MPMGPU_SHAPE_LCPDI;
    if (dim == 3) {
        if (cp) TILE_LAUNCH((k_p2g_mass<3, MPMGPU_SHAPE_LCPDI>), tb, s, g, p, tb, nd, st);
        else TILE_LAUNCH((k_p2g_mass<3, MPMGPU_SHAPE_UGIMP>), tb, s, g, p, tb, nd, st);
    } else {
        if (cp) TILE_LAUNCH((k_p2g_mass<2, MPMGPU_SHAPE_LCPDI>), tb, s, g, p, tb, nd, st);
        else TILE_LAUNCH((k_p2g_mass<2, MPMGPU_SHAPE_UGIMP>), tb, s, g, p, tb, nd, st);
    }
}

void launch_node_mass(int dim, const KGrid &g, const KBins &b, const KNodes &nd, hipStream_t s) {
    if (dim == 3) hipLaunchKernelGGL(k_node_mass<3>, dim3(blocks_for(g.nstride * g.nfields, NT)), dim3(NT), 0, s, g, b, nd);
    else hipLaunchKernelGGL(k_node_mass<2>, dim3(blocks_for(g.nstride * g.nfields, NT)), dim3(NT), 0, s, g, b, nd);
}

template <int DIM, int SH>
static void usf_mat(int mat, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                    const MpmMaterial *mats, int only, int slot, hipStream_t s) {
    const KBins tb = tile_bins(b, slot);
    switch (mat) {
        case MPMGPU_MAT_ISOTROPIC:
            TILE_LAUNCH((k_usf<DIM, MPMGPU_MAT_ISOTROPIC, SH>), tb, s, g, p, tb, nd, mats, only); break;
        case MPMGPU_MAT_ISOPLASTICITY:
            TILE_LAUNCH((k_usf<DIM, MPMGPU_MAT_ISOPLASTICITY, SH>), tb, s, g, p, tb, nd, mats, only); break;
        case MPMGPU_MAT_NEOHOOKEAN:
            TILE_LAUNCH((k_usf<DIM, MPMGPU_MAT_NEOHOOKEAN, SH>), tb, s, g, p, tb, nd, mats, only); break;
        case MPMGPU_MAT_MOONEY:
            TILE_LAUNCH((k_usf<DIM, MPMGPU_MAT_MOONEY, SH>), tb, s, g, p, tb, nd, mats, only); break;
        default:
            TILE_LAUNCH((k_usf<DIM, 0, SH>), tb, s, g, p, tb, nd, mats, only); break;
    }
}

// the strain updates of a launch: one launch with the law of every particle when they share a
// type (or the grid has no per-tile types), else one launch per type present, each taking only the
// tiles of its type, and one with every law for tiles whose instance mixes types
template <typename L>
static void by_type(const KGrid &g, int firstTypedSlot, int plainSlot, L launch) {
    if (g.matType != 0 || !g.tileType) {
        launch(g.matType, -1, plainSlot);
        return;
    }
    for (int j = 0; j < g.ntypes; j++) launch(g.types[j], g.types[j], firstTypedSlot + j);
    if (g.anyMixed) launch(0, 0, plainSlot);
}

void launch_usf(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                const MpmMaterial *mats, hipStream_t s) {
    const bool cp = g.shape == MPMGPU_SHAPE_LCPDI;
    by_type(g, TILE_SLOT_TYPED, TILE_SLOT_USF, [&](int mat, int only, int slot) {
        if (dim == 3) {
            if (cp) usf_mat<3, MPMGPU_SHAPE_LCPDI>(mat, g, p, b, nd, mats, only, slot, s);
            else usf_mat<3, MPMGPU_SHAPE_UGIMP>(mat, g, p, b, nd, mats, only, slot, s);
        } else {
            if (cp) usf_mat<2, MPMGPU_SHAPE_LCPDI>(mat, g, p, b, nd, mats, only, slot, s);
            else usf_mat<2, MPMGPU_SHAPE_UGIMP>(mat, g, p, b, nd, mats, only, slot, s);
        }
    });
}

// BoundaryCondition::GetNodeNum and BCValue (BoundaryCondition.cpp:72-140): 1 if the condition
// acts at time t, its value in v
__device__ __forceinline__ int style_value(int style, double val, double ft, double t, double &v) {
    v = 0.;
    switch (style) {
        case 1: if (t >= ft) { v = val; return 1; } return 0;
        case 2: if (t >= ft) { v = 0. + val * (t - ft); return 1; } return 0;
        case 3: v = val * sin(t * ft); return 1;
        case 4: v = val * cos(t * ft); return 1;
        default: return 0;
    }
}

// ---------------------------------------------------------------------------------------
// Traction conditions (MatPtTractionBC::AddMPFluxBC, MatPtTractionBC.cpp:64-227: planar and 3D,
// no cracks, no ExactTractions). GetSurfaceInfo (MatPoint2D.cpp:1138-1300, MatPoint3D.cpp:628-851)
// gives the corners of the loaded face and hat n times its weight: uGIMP and classic particles
// take the corners of the undeformed particle and the size of the deformed face, lCPDI particles
// the deformed corners and their own face. Each corner's element adds its linear shape functions
// at the corner (GetShapeFunctionsForTractions, MoreMPMElementBase.cpp:91-107) times
// tmag * weight to the particle's material field at the element's nodes where that field has
// points (NodalPoint::AddTractionTask3, NodalPointMPM.cpp:644-650).

// signs of r1, r2 (and r3) at the corners of each face, in the reference's corner order
__constant__ signed char k_face_corners2[4][2][2] = {
    {{-1, -1}, {1, -1}}, {{1, -1}, {1, 1}}, {{1, 1}, {-1, 1}}, {{-1, 1}, {-1, -1}}};
__constant__ signed char k_face_corners3[6][4][3] = {
    {{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1}},    // 1: -y, nodes 1, 2, 6, 5
    {{1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {1, -1, 1}},        // 2: +x, nodes 2, 3, 7, 6
    {{1, 1, -1}, {-1, 1, -1}, {-1, 1, 1}, {1, 1, 1}},        // 3: +y, nodes 3, 4, 8, 7
    {{-1, 1, -1}, {-1, -1, -1}, {-1, -1, 1}, {-1, 1, 1}},    // 4: -x, nodes 4, 1, 5, 8
    {{1, -1, -1}, {-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}},    // 5: -z, nodes 2, 1, 4, 3
    {{-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}}};       // 6: +z, nodes 5, 6, 7, 8

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - b[1] * a[2];          // CrossProduct (CommonUtilities.cpp:215-221)
    c[1] = b[0] * a[2] - a[0] * b[2];
    c[2] = a[0] * b[1] - b[0] * a[1];
}

// corners (cx[k]) and weighted normal (wn) of one face; returns the corner count
template <int DIM>
__device__ int surface_info(const KGrid &g, const KPart &P, int p, int ei, int ej, int ek, int face, int dof,
                            double thick, double cx[4][3], double wn[3]) {
    double r[9];
    cpdi_semi_sides<DIM>(g, P, p, ei, ej, ek, r);            // deformed semi sides F (psz)
    double *r1 = r, *r2 = r + 3, *r3 = r + 6;
    double uG = -1.;
    if (g.shape != MPMGPU_SHAPE_LCPDI) {
        if (DIM == 3) {
            double ap[3];
            if (face == 1 || face == 3) cross3(r1, r3, ap);
            else if (face == 2 || face == 4) cross3(r2, r3, ap);
            else cross3(r1, r2, ap);
            uG = sqrt(ap[0] * ap[0] + ap[1] * ap[1] + ap[2] * ap[2]);
        } else {
            const double *rs = (face == 1 || face == 3) ? r1 : r2;
            uG = sqrt(rs[0] * rs[0] + rs[1] * rs[1] + rs[2] * rs[2]);
        }
        // the undeformed semi sides (GetUndeformedSemiSides)
        double lpx, lpy, lpz;
        if (g.uniformLp) { lpx = g.lp[0]; lpy = g.lp[1]; lpz = g.lp[2]; }
        else { lpx = P.f[MPF_LP_X][p]; lpy = P.f[MPF_LP_Y][p]; lpz = P.f[MPF_LP_Z][p]; }
        const double ux = ((g.xmin + (double)(ei + 1) * g.dx) - (g.xmin + (double)ei * g.dx)) * (0.5 * lpx);
        const double uy = ((g.ymin + (double)(ej + 1) * g.dy) - (g.ymin + (double)ej * g.dy)) * (0.5 * lpy);
        const double uz = DIM == 3 ? ((g.zmin + (double)(ek + 1) * g.dz) - (g.zmin + (double)ek * g.dz)) * (0.5 * lpz) : 0.;
#pragma unroll
        for (int a = 0; a < 9; a++) r[a] = 0.;
        r1[0] = ux; r2[1] = uy; r3[2] = uz;
    }
    const double pos[3] = {P.f[MPF_POS_X][p], P.f[MPF_POS_Y][p], DIM == 3 ? P.f[MPF_POS_Z][p] : 0.};
    const int nc = DIM == 3 ? 4 : 2;
    for (int k = 0; k < nc; k++) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (DIM == 3) {
                const signed char *sg = k_face_corners3[face - 1][k];
                cx[k][a] = ((pos[a] + (double)sg[0] * r1[a]) + (double)sg[1] * r2[a]) + (double)sg[2] * r3[a];
            } else {
                const signed char *sg = k_face_corners2[face - 1][k];
                cx[k][a] = a < 2 ? (pos[a] + (double)sg[0] * r1[a]) + (double)sg[1] * r2[a] : 0.;
            }
        }
    }
    wn[0] = wn[1] = wn[2] = 0.;
    double r0[3];
#pragma unroll
    for (int a = 0; a < 3; a++) r0[a] = 0.5 * (cx[1][a] - cx[0][a]);
    if (DIM == 3) {
        double r1f[3], ap[3];
#pragma unroll
        for (int a = 0; a < 3; a++) r1f[a] = 0.5 * (cx[3][a] - cx[0][a]);
        cross3(r0, r1f, ap);
        const double am = sqrt(ap[0] * ap[0] + ap[1] * ap[1] + ap[2] * ap[2]);
        const double faceWt = uG < 0. ? am : uG;
        if (dof == 1) wn[0] = faceWt;
        else if (dof == 2) wn[1] = faceWt;
        else if (dof == 11) {
            const double sc = uG > 0. ? uG / am : 1.;
#pragma unroll
            for (int a = 0; a < 3; a++) wn[a] = uG > 0. ? ap[a] * sc : ap[a];
        } else wn[2] = faceWt;
    } else {
        double faceWt = uG < 0. ? sqrt(r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]) : uG;
        faceWt *= thick;
        if (dof == 1) wn[0] = faceWt;
        else if (dof == 2) wn[1] = faceWt;
        else if (dof == 11 || dof == 12) {
            const double ex = dof == 11 ? r0[1] : r0[0], ey = dof == 11 ? -r0[0] : r0[1];
            const double en = sqrt(ex * ex + ey * ey);
            wn[0] = ex * faceWt / en;
            wn[1] = ey * faceWt / en;
        } else wn[2] = faceWt;
    }
    return nc;
}

// the nodal traction forces of traction table t (particle p) at step time tt
template <int DIM>
__device__ void particle_tractions(const KGrid &g, const KPart &P, const KNodes &N, const KLoads &L, int p, int t,
                                   double tt) {
    double *tab = L.ttab + (size_t)3 * t * TRAC_NODES;
    for (int o = 0; o < 3 * TRAC_NODES; o++) tab[o] = 0.;
    int ei, ej, ek;
    elem_ijk(g, P.elem[p], ei, ej, ek);
    const int fld = g.fieldOf[P.mat[p]];
    const KNodes Nf = field_view(g, N, fld);
    for (int e = L.tstart[t]; e < L.tstart[t + 1]; e++) {
        double tmag;
        style_value(L.tstyle[e], L.tvalue[e], L.tftime[e], tt, tmag);
        double cx[4][3], wn[3];
        const int nc = surface_info<DIM>(g, P, p, ei, ej, ek, L.tface[e], L.tdir[e], L.tthick[t], cx, wn);
        for (int k = 0; k < nc; k++) {
            bool ok = true;
            const int ci = find_cell(cx[k][0], g.xmin, g.dx, g.nx, ok);
            const int cj = find_cell(cx[k][1], g.ymin, g.dy, g.ny, ok);
            const int ck = DIM == 3 ? find_cell(cx[k][2], g.zmin, g.dz, g.nz, ok) : 0;
            const int ox = ci - ei, oy = cj - ej, oz = DIM == 3 ? ck - ek : 0;
            if (!ok || ox < -1 || ox > 1 || oy < -1 || oy > 1 || oz < -1 || oz > 1) {
                atomicAdd(L.terr, 1ull);          // a fatal error, reported after the step
                continue;
            }
            double xi[3], invd[3];
            particle_xi<DIM>(g, ci, cj, ck, cx[k][0], cx[k][1], cx[k][2], xi, invd);
            for (int jz = 0; jz < (DIM == 3 ? 2 : 1); jz++)
                for (int jy = 0; jy < 2; jy++)
                    for (int jx = 0; jx < 2; jx++) {
                        const long long n = ((long long)(ck + jz) * g.nny + (cj + jy)) * g.nnx + (ci + jx);
                        if (Nf.cnt[n] <= 0) continue;
                        const double fn = cpdi_cfn<DIM>(xi, jx, jy, jz);
                        const int o = DIM == 3 ? ((oz + jz + 1) * 4 + (oy + jy + 1)) * 4 + (ox + jx + 1)
                                               : (oy + jy + 1) * 4 + (ox + jx + 1);
#pragma unroll
                        for (int c = 0; c < DIM; c++) tab[c * TRAC_NODES + o] += wn[c] * (tmag * fn);
                    }
        }
    }
}

// marks the (field, tile) of every loaded particle in slot par, clears slot par ^ 1 for the next
// step, and fills the traction tables of the step (PostForcesTask.cpp:51, after the grid forces'
// strain update and the mass pass that sets the point counts)
template <int DIM>
__global__ void k_mark_load_tiles(KGrid g, KPart P, KBins B, KNodes N, KLoads L, int par) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < B.ntall) L.tflag[par ^ 1][p] = 0;
    if (p >= P.n) return;
    const int q = L.map[B.home[p]];
    if (q < 0) return;
    L.tflag[par][B.key[p] / TileGeom<DIM>::NC] = 1;
    if (L.nt > 0 && L.tslot[q] >= 0) particle_tractions<DIM>(g, P, N, L, (int)p, L.tslot[q], L.time[par]);
}

void launch_p2g_force(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd, const KLoads &L,
                      int par, hipStream_t s) {
    const KBins tb = tile_bins(b, TILE_SLOT_FORCE);
    const bool cp = g.shape == MPMGPU_SHAPE_LCPDI;
    const int *loaded = nullptr;
    if (L.nl > 0) {
        const long long n = std::max<long long>(p.n, b.ntall);
        if (dim == 3) hipLaunchKernelGGL(k_mark_load_tiles<3>, dim3(blocks_for(n, NT)), dim3(NT), 0, s, g, p, b, nd, L, par);
        else hipLaunchKernelGGL(k_mark_load_tiles<2>, dim3(blocks_for(n, NT)), dim3(NT), 0, s, g, p, b, nd, L, par);
        loaded = L.tflag[par];
        const KBins tl = tile_bins(b, TILE_SLOT_FORCE_LOADS);
        if (dim == 3) {
            if (cp) TILE_LAUNCH((k_p2g_force<3, MPMGPU_SHAPE_LCPDI, true>), tl, s, g, p, tl, nd, L, loaded);
            else TILE_LAUNCH((k_p2g_force<3, MPMGPU_SHAPE_UGIMP, true>), tl, s, g, p, tl, nd, L, loaded);
        } else {
            if (cp) TILE_LAUNCH((k_p2g_force<2, MPMGPU_SHAPE_LCPDI, true>), tl, s, g, p, tl, nd, L, loaded);
            else TILE_LAUNCH((k_p2g_force<2, MPMGPU_SHAPE_UGIMP, true>), tl, s, g, p, tl, nd, L, loaded);
        }
    }
    if (dim == 3) {
        if (cp) TILE_LAUNCH((k_p2g_force<3, MPMGPU_SHAPE_LCPDI, false>), tb, s, g, p, tb, nd, L, loaded);
        else TILE_LAUNCH((k_p2g_force<3, MPMGPU_SHAPE_UGIMP, false>), tb, s, g, p, tb, nd, L, loaded);
    } else {
        if (cp) TILE_LAUNCH((k_p2g_force<2, MPMGPU_SHAPE_LCPDI, false>), tb, s, g, p, tb, nd, L, loaded);
        else TILE_LAUNCH((k_p2g_force<2, MPMGPU_SHAPE_UGIMP, false>), tb, s, g, p, tb, nd, L, loaded);
    }
}

void launch_node_force(int dim, const KGrid &g, const KBins &b, const KNodes &nd, hipStream_t s) {
    if (dim == 3) hipLaunchKernelGGL(k_node_force<3>, dim3(blocks_for(g.nstride * g.nfields, NT)), dim3(NT), 0, s, g, b, nd);
    else hipLaunchKernelGGL(k_node_force<2>, dim3(blocks_for(g.nstride * g.nfields, NT)), dim3(NT), 0, s, g, b, nd);
}

template <int MOVE>
static void g2p_move(int dim, const KGrid &g, const KPart &p, const KBins &tb, const KNodes &nd,
                     const MpmMaterial *mats, int mode, hipStream_t s) {
    const bool cp = g.shape == MPMGPU_SHAPE_LCPDI;
    if (dim == 3) {
        if (cp) TILE_LAUNCH((k_g2p_p2g<3, MPMGPU_SHAPE_LCPDI, MOVE>), tb, s, g, p, tb, nd, mats, mode);
        else TILE_LAUNCH((k_g2p_p2g<3, MPMGPU_SHAPE_UGIMP, MOVE>), tb, s, g, p, tb, nd, mats, mode);
    } else {
        if (cp) TILE_LAUNCH((k_g2p_p2g<2, MPMGPU_SHAPE_LCPDI, MOVE>), tb, s, g, p, tb, nd, mats, mode);
        else TILE_LAUNCH((k_g2p_p2g<2, MPMGPU_SHAPE_UGIMP, MOVE>), tb, s, g, p, tb, nd, mats, mode);
    }
}

void launch_g2p_p2g(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                    const MpmMaterial *mats, int mode, hipStream_t s) {
    const KBins tb = tile_bins(b, TILE_SLOT_G2P);
    // damping and the FMPM/XPIC moves have their own instances, so the plain FLIP update keeps its registers
    if (mode != 0) g2p_move<G2P_SERIES>(dim, g, p, tb, nd, mats, mode, s);
    else if (g.damped) g2p_move<G2P_FLIP_DAMPED>(dim, g, p, tb, nd, mats, mode, s);
    else g2p_move<G2P_FLIP>(dim, g, p, tb, nd, mats, mode, s);
}

void launch_node_mom2(int dim, const KGrid &g, const KBins &b, const KNodes &nd, hipStream_t s) {
    if (dim == 3) hipLaunchKernelGGL(k_node_mom2<3>, dim3(blocks_for(g.nstride * g.nfields, NT)), dim3(NT), 0, s, g, b, nd);
    else hipLaunchKernelGGL(k_node_mom2<2>, dim3(blocks_for(g.nstride * g.nfields, NT)), dim3(NT), 0, s, g, b, nd);
}

template <int DIM, int SH>
static void usl_reset_mat(int mat, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                          const MpmMaterial *mats, const KStats &st, int only, int slot, hipStream_t s) {
    const KBins tb = tile_bins(b, slot);
    switch (mat) {
        case MPMGPU_MAT_ISOTROPIC:
            TILE_LAUNCH((k_usl_reset<DIM, MPMGPU_MAT_ISOTROPIC, SH>), tb, s, g, p, tb, nd, mats, st, only); break;
        case MPMGPU_MAT_ISOPLASTICITY:
            TILE_LAUNCH((k_usl_reset<DIM, MPMGPU_MAT_ISOPLASTICITY, SH>), tb, s, g, p, tb, nd, mats, st, only); break;
        case MPMGPU_MAT_NEOHOOKEAN:
            TILE_LAUNCH((k_usl_reset<DIM, MPMGPU_MAT_NEOHOOKEAN, SH>), tb, s, g, p, tb, nd, mats, st, only); break;
        case MPMGPU_MAT_MOONEY:
            TILE_LAUNCH((k_usl_reset<DIM, MPMGPU_MAT_MOONEY, SH>), tb, s, g, p, tb, nd, mats, st, only); break;
        default:
            TILE_LAUNCH((k_usl_reset<DIM, 0, SH>), tb, s, g, p, tb, nd, mats, st, only); break;
    }
}

void launch_usl_reset(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                      const MpmMaterial *mats, const KStats &st, hipStream_t s) {
    const bool cp = g.shape == MPMGPU_SHAPE_LCPDI;
    by_type(g, TILE_SLOT_TYPED + 4, TILE_SLOT_USL, [&](int mat, int only, int slot) {
        if (dim == 3) {
            if (cp) usl_reset_mat<3, MPMGPU_SHAPE_LCPDI>(mat, g, p, b, nd, mats, st, only, slot, s);
            else usl_reset_mat<3, MPMGPU_SHAPE_UGIMP>(mat, g, p, b, nd, mats, st, only, slot, s);
        } else {
            if (cp) usl_reset_mat<2, MPMGPU_SHAPE_LCPDI>(mat, g, p, b, nd, mats, st, only, slot, s);
            else usl_reset_mat<2, MPMGPU_SHAPE_UGIMP>(mat, g, p, b, nd, mats, st, only, slot, s);
        }
    });
}

void launch_p2g_contact(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                        const MpmMaterial *mats, bool newPositions, hipStream_t s) {
    const KBins tb = tile_bins(b, newPositions ? TILE_SLOT_CONTACT_LAST : TILE_SLOT_CONTACT_FIRST);
    int np = newPositions ? 1 : 0;
    const bool cp = g.shape == MPMGPU_SHAPE_LCPDI;
    if (dim == 3) {
        if (cp) TILE_LAUNCH((k_p2g_contact<3, MPMGPU_SHAPE_LCPDI>), tb, s, g, p, tb, nd, mats, np);
        else TILE_LAUNCH((k_p2g_contact<3, MPMGPU_SHAPE_UGIMP>), tb, s, g, p, tb, nd, mats, np);
    } else {
        if (cp) TILE_LAUNCH((k_p2g_contact<2, MPMGPU_SHAPE_LCPDI>), tb, s, g, p, tb, nd, mats, np);
        else TILE_LAUNCH((k_p2g_contact<2, MPMGPU_SHAPE_UGIMP>), tb, s, g, p, tb, nd, mats, np);
    }
}

void launch_node_contact(int dim, const KGrid &g, const KBins &b, const KNodes &nd, hipStream_t s) {
    if (dim == 3) hipLaunchKernelGGL(k_node_contact<3>, dim3(blocks_for(g.nstride * g.nfields, NT)), dim3(NT), 0, s, g, b, nd);
    else hipLaunchKernelGGL(k_node_contact<2>, dim3(blocks_for(g.nstride * g.nfields, NT)), dim3(NT), 0, s, g, b, nd);
}

// ---------------------------------------------------------------------------------------
// Grid velocity conditions (NodalVelBC::GridVelocityConditions -> VelocityBCLoop,
// NodalVelBC.cpp:321-400; MatVelocityField::ZeroVelocityBC and AddVelocityBC,
// MatVelocityField.cpp:490-575, norm a unit axis). The reference zeroes for every condition of
// the list, then adds for every one; conditions on different nodes touch different values, so
// taking each node's conditions in list order, zeroes first, gives the same numbers.

__device__ __forceinline__ int bc_value(const KBcs &bc, int e, double t, double &v) {
    return style_value(bc.style[e], bc.value[e], bc.ftime[e], t, v);
}

// one pass on momenta p and forces fo of one field of one node (ZeroVelocityBC, then AddVelocityBC)
template <int PASS>
__device__ __forceinline__ void bc_pass(const KBcs &bc, int e0, int e1, double t, double m, double dt,
                                        double p[3], double fo[3]) {
    for (int e = e0; e < e1; e++) {
        double v;
        const int on = PASS == BC_PASS_MASS_MOMENTUM ? bc_value(bc, e, t, v) : bc.act[e];
        if (!on) continue;
        const int a = bc.dir[e];
        if (PASS == BC_PASS_GRID_FORCES) {
            const double dotf = fo[a], dotp = p[a];
            fo[a] += -dotf - dotp / dt;
        } else {
            const double dotn = p[a];
            p[a] += -dotn;
            if (PASS == BC_PASS_UPDATE_MOMENTUM) fo[a] += -dotn / dt;
        }
    }
    for (int e = e0; e < e1; e++) {
        double v;
        int on;
        if (PASS == BC_PASS_MASS_MOMENTUM) on = bc_value(bc, e, t, v);
        else { on = bc.act[e]; v = bc.cur[e]; }
        if (!on) continue;
        const int a = bc.dir[e];
        const double mv = m * v;
        if (PASS == BC_PASS_GRID_FORCES) {
            fo[a] += mv / dt;
        } else {
            p[a] += mv;
            if (PASS == BC_PASS_UPDATE_MOMENTUM) fo[a] += mv / dt;
        }
    }
}

// one thread per (node with conditions, field). MASS_MOMENTUM (after the contact pass, before the
// first strain update) keeps the momenta it starts from; GRID_FORCES starts from those again
// (MatVelocityField::RestoreMomenta), sets the forces and redoes the momentum update of the node
// (pk + ftot*dt, as k_node_force). The mass-momentum pass reads the step's start time from slot
// par and writes the next step's into slot par ^ 1, which no thread of this step reads.
template <int PASS>
__global__ void k_vel_bcs(KGrid g, KNodes N, KBcs bc, int par) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (PASS == BC_PASS_MASS_MOMENTUM && idx == 0) bc.time[par ^ 1] = bc.time[par] + g.timestep;
    if (idx >= (long long)bc.nb * g.nfields) return;
    const int f = (int)(idx / bc.nb);
    const int b = (int)(idx - (long long)f * bc.nb);
    const int e0 = bc.start[b], e1 = bc.start[b + 1];
    double t = 0.;
    if (PASS == BC_PASS_MASS_MOMENTUM) {
        t = bc.time[par];
        if (f == 0)
            for (int e = e0; e < e1; e++) {
                double v;
                bc.act[e] = bc_value(bc, e, t, v);
                bc.cur[e] = v;
            }
    }
    const KNodes Nf = field_view(g, N, f);
    const long long n = bc.node[b], nn = g.nstride;
    if (Nf.cnt[n] <= 0) return;
    const double m = Nf.mass[n], dt = g.timestep;
    if (PASS == BC_PASS_XPIC || PASS == BC_PASS_XPIC_FORCE) {
        // NodalVelBC::GridVelocityConditions for the XPIC passes (NodalVelBC.cpp:327-335) ->
        // MatVelocityField::ZeroVelocityBC (MatVelocityField.cpp:523-534) on the increment
        for (int e = e0; e < e1; e++) {
            if (!bc.act[e]) continue;
            const int a = bc.dir[e];
            const double dotn = Nf.vprev[a * nn + n];
            Nf.vprev[a * nn + n] = 0.;
            Nf.vk[a * nn + n] -= dotn;
            if (PASS == BC_PASS_XPIC_FORCE) Nf.ftot[a * nn + n] += -m * dotn / dt;
        }
        return;
    }
    double p[3] = {Nf.pk[n], Nf.pk[nn + n], Nf.pk[2 * nn + n]};
    double fo[3] = {Nf.ftot[n], Nf.ftot[nn + n], Nf.ftot[2 * nn + n]};
    double *sv = bc.save + ((long long)f * bc.nb + b) * 3;
    if (PASS == BC_PASS_MASS_MOMENTUM) {
        sv[0] = p[0]; sv[1] = p[1]; sv[2] = p[2];
        bc_pass<BC_PASS_MASS_MOMENTUM>(bc, e0, e1, t, m, dt, p, fo);
    } else if (PASS == BC_PASS_GRID_FORCES || PASS == BC_PASS_FORCES_MOMENTA) {
        p[0] = sv[0]; p[1] = sv[1]; p[2] = sv[2];
        bc_pass<BC_PASS_GRID_FORCES>(bc, e0, e1, t, m, dt, p, fo);
        for (int c = 0; c < 3; c++) p[c] = p[c] + fo[c] * dt;
        if (PASS == BC_PASS_FORCES_MOMENTA) bc_pass<BC_PASS_UPDATE_MOMENTUM>(bc, e0, e1, t, m, dt, p, fo);
    } else {
        bc_pass<PASS>(bc, e0, e1, t, m, dt, p, fo);
    }
    Nf.pk[n] = p[0]; Nf.pk[nn + n] = p[1]; Nf.pk[2 * nn + n] = p[2];
    if (PASS != BC_PASS_MASS_MOMENTUM && PASS != BC_PASS_UPDATE_STRAINS_LAST) {
        Nf.ftot[n] = fo[0]; Nf.ftot[nn + n] = fo[1]; Nf.ftot[2 * nn + n] = fo[2];
    }
    if (m != 0.) {
        const double s = 1. / m;
        Nf.vk[n] = s * p[0]; Nf.vk[nn + n] = s * p[1]; Nf.vk[2 * nn + n] = s * p[2];
    }
}

void launch_vel_bcs(const KGrid &g, const KNodes &nd, const KBcs &bc, int pass, int par, hipStream_t s) {
    const dim3 grid(blocks_for((long long)bc.nb * g.nfields, NT)), blk(NT);
    switch (pass) {
        case BC_PASS_MASS_MOMENTUM: hipLaunchKernelGGL(k_vel_bcs<BC_PASS_MASS_MOMENTUM>, grid, blk, 0, s, g, nd, bc, par); break;
        case BC_PASS_GRID_FORCES: hipLaunchKernelGGL(k_vel_bcs<BC_PASS_GRID_FORCES>, grid, blk, 0, s, g, nd, bc, par); break;
        case BC_PASS_UPDATE_MOMENTUM: hipLaunchKernelGGL(k_vel_bcs<BC_PASS_UPDATE_MOMENTUM>, grid, blk, 0, s, g, nd, bc, par); break;
        case BC_PASS_FORCES_MOMENTA: hipLaunchKernelGGL(k_vel_bcs<BC_PASS_FORCES_MOMENTA>, grid, blk, 0, s, g, nd, bc, par); break;
        case BC_PASS_XPIC: hipLaunchKernelGGL(k_vel_bcs<BC_PASS_XPIC>, grid, blk, 0, s, g, nd, bc, par); break;
        case BC_PASS_XPIC_FORCE: hipLaunchKernelGGL(k_vel_bcs<BC_PASS_XPIC_FORCE>, grid, blk, 0, s, g, nd, bc, par); break;
        default: hipLaunchKernelGGL(k_vel_bcs<BC_PASS_UPDATE_STRAINS_LAST>, grid, blk, 0, s, g, nd, bc, par); break;
    }
}

void launch_xpic(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd, const KBcs &bc,
                 int order, bool fmpm, bool forcePass, int par, hipStream_t s) {
    const dim3 ng(blocks_for(g.nstride * g.nfields, NT)), blk(NT);
    const KBins tb = tile_bins(b, TILE_SLOT_XPIC);
    const bool cp = g.shape == MPMGPU_SHAPE_LCPDI;
    if (dim == 3) hipLaunchKernelGGL(k_xpic_init<3>, ng, blk, 0, s, g, b, nd, fmpm ? 1 : 0);
    else hipLaunchKernelGGL(k_xpic_init<2>, ng, blk, 0, s, g, b, nd, fmpm ? 1 : 0);
    for (int it = 2; it <= order; it++) {
        if (dim == 3) {
            if (cp) TILE_LAUNCH((k_xpic_p2g<3, MPMGPU_SHAPE_LCPDI>), tb, s, g, p, tb, nd);
            else TILE_LAUNCH((k_xpic_p2g<3, MPMGPU_SHAPE_UGIMP>), tb, s, g, p, tb, nd);
            hipLaunchKernelGGL(k_xpic_node<3>, ng, blk, 0, s, g, tb, nd);
        } else {
            if (cp) TILE_LAUNCH((k_xpic_p2g<2, MPMGPU_SHAPE_LCPDI>), tb, s, g, p, tb, nd);
            else TILE_LAUNCH((k_xpic_p2g<2, MPMGPU_SHAPE_UGIMP>), tb, s, g, p, tb, nd);
            hipLaunchKernelGGL(k_xpic_node<2>, ng, blk, 0, s, g, tb, nd);
        }
        if (bc.nb) MPM_KNS::launch_vel_bcs(g, nd, bc, forcePass && !fmpm ? BC_PASS_XPIC_FORCE : BC_PASS_XPIC, par, s);
    }
}

// MatPtLoadBC::SetParticleFext (MatPtLoadBC.cpp:210-221, called from InitializationTask.cpp:91):
// one thread per loaded particle zeroes its pFext and adds each of its loads' value at the
// step's start time (BCValue, zero before a constant or linear load starts), in list order.
// Like the mass-momentum velocity pass it writes the next step's start time into slot par ^ 1.
__global__ void k_particle_loads(KGrid g, KLoads L, int par) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const double t = L.time[par];
    if (q == 0) L.time[par ^ 1] = t + g.timestep;
    if (q >= L.nl) return;
    double fe[3] = {0., 0., 0.};
    for (int e = L.start[q]; e < L.start[q + 1]; e++) {
        double v;
        style_value(L.style[e], L.value[e], L.ftime[e], t, v);
        fe[L.dir[e]] += v;
    }
    for (int c = 0; c < 3; c++) L.fext[3 * q + c] = fe[c];
}

void launch_particle_loads(const KGrid &g, const KLoads &L, int par, hipStream_t s) {
    hipLaunchKernelGGL(k_particle_loads, dim3(blocks_for(L.nl, NT)), dim3(NT), 0, s, g, L, par);
}

void launch_contact_normals(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                            const MpmMaterial *mats, const KStats &st, int callType, hipStream_t s) {
    if (g.normals < 5 || !nd.cnorm) return;
    if (callType == CONTACT_MASS_MOMENTUM) {
        if (dim == 3) hipLaunchKernelGGL(k_contact_list<3>, dim3(blocks_for(g.nnodes, NT)), dim3(NT), 0, s, g, nd);
        else hipLaunchKernelGGL(k_contact_list<2>, dim3(blocks_for(g.nnodes, NT)), dim3(NT), 0, s, g, nd);
    }
    const int np = callType == CONTACT_UPDATE_STRAINS_LAST ? 1 : 0;
    const dim3 grid(4 * g_cus);
    const bool cp = g.shape == MPMGPU_SHAPE_LCPDI;
    if (dim == 3) {
        if (cp) hipLaunchKernelGGL((k_contact_normals<3, MPMGPU_SHAPE_LCPDI>), grid, dim3(NT), 0, s, g, p, b, nd, mats, st, np);
        else hipLaunchKernelGGL((k_contact_normals<3, MPMGPU_SHAPE_UGIMP>), grid, dim3(NT), 0, s, g, p, b, nd, mats, st, np);
    } else {
        if (cp) hipLaunchKernelGGL((k_contact_normals<2, MPMGPU_SHAPE_LCPDI>), grid, dim3(NT), 0, s, g, p, b, nd, mats, st, np);
        else hipLaunchKernelGGL((k_contact_normals<2, MPMGPU_SHAPE_UGIMP>), grid, dim3(NT), 0, s, g, p, b, nd, mats, st, np);
    }
}

void launch_contact(int dim, const KGrid &g, const KNodes &nd, int callType, hipStream_t s) {
    if (dim == 3) hipLaunchKernelGGL(k_contact<3>, dim3(blocks_for(g.nnodes, NT)), dim3(NT), 0, s, g, nd, callType);
    else hipLaunchKernelGGL(k_contact<2>, dim3(blocks_for(g.nnodes, NT)), dim3(NT), 0, s, g, nd, callType);
}

// ---------------------------------------------------------------------------------------
// particle runs between contexts: one thread per (particle, row); rows 0..MPF_NUM_FIELDS-1 are
// the double fields, the next MPI_NUM_FIELDS the int fields

__global__ void k_move(KMove m) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (p >= m.cnt) return;
    if (r < MPF_NUM_FIELDS) {
        const double *src = (r >= MPF_POS_X && r <= MPF_POS_Z) ? m.spos[r - MPF_POS_X] : m.sd + r * m.sn;
        m.dd[r * m.dn + m.doff + p] = src[m.soff + p];
    } else {
        // int rows MPI_*, then the upload-index row (MPI_NUM_FIELDS + 2)
        const int q = r - MPF_NUM_FIELDS < MPI_NUM_FIELDS ? r - MPF_NUM_FIELDS : MPI_NUM_FIELDS + 2;
        int v = m.si[q * m.sn + m.soff + p];
        if (q == MPI_MAT) v += m.matShift;
        if (q == MPI_NUM_FIELDS + 2) v += m.homeShift;
        m.di[q * m.dn + m.doff + p] = v;
    }
}

void launch_move(const KMove &m, hipStream_t s) {
    if (m.cnt <= 0) return;
    hipLaunchKernelGGL(k_move, dim3(blocks_for(m.cnt, NT), MPF_NUM_FIELDS + MPI_NUM_FIELDS + 1), dim3(NT), 0, s, m);
}

template <typename T>
__global__ void k_gather_row(const T *src, T *dst, const int *order, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[order[k]];
}

template <typename T>
__global__ void k_scatter_row(const T *src, T *dst, const int *home, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[home[k]] = src[k];
}

__global__ void k_iota(int *a, long long n, int base) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = base + (int)k;
}

void launch_gather_row(const double *src, double *dst, const int *order, long long n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_gather_row<double>, dim3(blocks_for(n, NT)), dim3(NT), 0, s, src, dst, order, n);
}
void launch_gather_irow(const int *src, int *dst, const int *order, long long n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_gather_row<int>, dim3(blocks_for(n, NT)), dim3(NT), 0, s, src, dst, order, n);
}
void launch_scatter_row(const double *src, double *dst, const int *home, long long n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_scatter_row<double>, dim3(blocks_for(n, NT)), dim3(NT), 0, s, src, dst, home, n);
}
void launch_scatter_irow(const int *src, int *dst, const int *home, long long n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_scatter_row<int>, dim3(blocks_for(n, NT)), dim3(NT), 0, s, src, dst, home, n);
}
void launch_iota(int *a, long long n, int base, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, NT)), dim3(NT), 0, s, a, n, base);
}

}  // namespace mpmk_s0 / mpmk_s1
