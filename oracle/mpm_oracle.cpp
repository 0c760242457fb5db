/*
 * mpm_oracle.cpp -- CPU restatement of NairnMPM's explicit time step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker of the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product never links it.
 *
 * It restates, in the reference's order of operations, the hot path of one MPM step for the
 * configurations of this project: single- or multi-material (multimaterial contact is not yet
 * restated here), uGIMP or linear shape functions, USAVG+ with FLIP, no cracks, no transport,
 * no boundary conditions. Every function cites the reference file:line it follows (paths are
 * relative to the reference root). Grid sums are accumulated particle by particle in mpm[]
 * order, which is the reference's summation order when it runs as a single patch (-np 1).
 *
 * Pinned by tests/test_oracle_*.py against closed forms, invariants and brute force (there are
 * no reference golden vectors: the reference needs Xerces-C, absent from this image).
 */
#include <omp.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <stdexcept>
#include <vector>

#include "../include/nairnmpm_gpu.h"

namespace {

const double SQRT_TWOTHIRDS = 0.8164965809277260;   // System/MPMPrefix.hpp:92

struct Vec3 { double x, y, z; };

// ---------------------------------------------------------------------------------------
// 3x3 matrix with the reference's 2D flag (Common/System/Matrix3.cpp:66-160)
struct Mat3 {
    double m[3][3];
    bool is2D;
    Mat3() { zero(); }
    void zero() { for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) m[i][j] = 0.; is2D = true; }
    static Mat3 full(double xx, double xy, double xz, double yx, double yy, double yz,
                     double zx, double zy, double zz) {          // Matrix3.cpp:75-89
        Mat3 a; a.m[0][0] = xx; a.m[0][1] = xy; a.m[0][2] = xz;
        a.m[1][0] = yx; a.m[1][1] = yy; a.m[1][2] = yz;
        a.m[2][0] = zx; a.m[2][1] = zy; a.m[2][2] = zz; a.is2D = false; return a;
    }
    static Mat3 two(double xx, double xy, double yx, double yy, double zz) {   // :92-104
        Mat3 a; a.m[0][0] = xx; a.m[0][1] = xy; a.m[1][0] = yx; a.m[1][1] = yy; a.m[2][2] = zz;
        a.is2D = true; return a;
    }
    void add(const Mat3 &r) {                                    // operator+= :884-892
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) m[i][j] += r.m[i][j];
        if (!r.is2D) is2D = false;
    }
    void scale(double f) {                                       // Scale :410-424
        m[0][0] *= f; m[0][1] *= f; m[1][0] *= f; m[1][1] *= f; m[2][2] *= f;
        if (!is2D) { m[0][2] *= f; m[1][2] *= f; m[2][0] *= f; m[2][1] *= f; }
    }
    Mat3 mul(const Mat3 &r) const {                              // operator*= :920-966
        Mat3 res;
        if (is2D && r.is2D) {
            res.m[0][0] = m[0][0] * r.m[0][0] + m[0][1] * r.m[1][0];
            res.m[0][1] = m[0][0] * r.m[0][1] + m[0][1] * r.m[1][1];
            res.m[1][0] = m[1][0] * r.m[0][0] + m[1][1] * r.m[1][0];
            res.m[1][1] = m[1][0] * r.m[0][1] + m[1][1] * r.m[1][1];
            res.m[2][2] = m[2][2] * r.m[2][2];
            res.is2D = true;
        } else {
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double s = 0.;
                    for (int k = 0; k < 3; k++) s += m[i][k] * r.m[k][j];
                    res.m[i][j] = s;
                }
            res.is2D = false;
        }
        return res;
    }
    double trace() const { return m[0][0] + m[1][1] + m[2][2]; }   // :1192
    double determinant() const {                                    // :1195-1202
        if (is2D) return m[2][2] * (m[0][0] * m[1][1] - m[1][0] * m[0][1]);
        return m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2])
             - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2])
             + m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    }
    // Exponential(kmax) for kmax = 1 or 2 (Matrix3.cpp:312-394)
    Mat3 exponential(int kmax) const {
        if (is2D) {
            if (kmax == 1) return two(1. + m[0][0], m[0][1], m[1][0], 1. + m[1][1], 1. + m[2][2]);
            double c0 = m[0][1] * m[1][0] - m[0][0] * m[1][1];
            double c1 = m[0][0] + m[1][1];
            double beta0 = 0., beta1 = 1., alpha0 = 1., alpha1 = 1.;
            double betaz = m[2][2], ezz = 1. + betaz;
            for (int k = 2; k <= kmax; k++) {
                double factor = 1 / (double)k;
                double temp = beta1;
                beta1 = factor * (c1 * temp + beta0);
                beta0 = factor * c0 * temp;
                betaz *= factor * m[2][2];
                alpha0 += beta0;
                alpha1 += beta1;
                ezz += betaz;
            }
            return two(alpha0 + alpha1 * m[0][0], alpha1 * m[0][1], alpha1 * m[1][0],
                       alpha0 + alpha1 * m[1][1], ezz);
        }
        if (kmax == 1)
            return full(1. + m[0][0], m[0][1], m[0][2], m[1][0], 1. + m[1][1], m[1][2],
                        m[2][0], m[2][1], 1. + m[2][2]);
        Mat3 m2 = *this;
        m2 = m2.mul(m2);
        return full(1. + m[0][0] + 0.5 * m2.m[0][0], m[0][1] + 0.5 * m2.m[0][1], m[0][2] + 0.5 * m2.m[0][2],
                    m[1][0] + 0.5 * m2.m[1][0], 1. + m[1][1] + 0.5 * m2.m[1][1], m[1][2] + 0.5 * m2.m[1][2],
                    m[2][0] + 0.5 * m2.m[2][0], m[2][1] + 0.5 * m2.m[2][1], 1. + m[2][2] + 0.5 * m2.m[2][2]);
    }
};

// ---------------------------------------------------------------------------------------
// uGIMP node tables (Common/Elements/FourNodeIsoparam.cpp:27-31,
// NairnMPM/src/Elements/EightNodeIsoparamBrick.cpp:23-53). Natural coordinate of a candidate
// node is 2*off-1 on each axis.
const int xoff2[16] = {0, 1, 1, 0, -1, 0, 1, 2, 2, 2, 2, 1, 0, -1, -1, -1};
const int yoff2[16] = {0, 0, 1, 1, -1, -1, -1, -1, 0, 1, 2, 2, 2, 2, 1, 0};
const int xoff3[64] = {0,1,1,0,0,1,1,0,
    -1,0,1,2,2,2,2,1,0,-1,-1,-1,
    -1,0,1,2,2,2,2,1,0,-1,-1,-1,
    0,1,1,0,-1,0,1,2,2,2,2,1,0,-1,-1,-1,
    0,1,1,0,-1,0,1,2,2,2,2,1,0,-1,-1,-1};
const int yoff3[64] = {0,0,1,1,0,0,1,1,
    -1,-1,-1,-1,0,1,2,2,2,2,1,0,
    -1,-1,-1,-1,0,1,2,2,2,2,1,0,
    0,0,1,1,-1,-1,-1,-1,0,1,2,2,2,2,1,0,
    0,0,1,1,-1,-1,-1,-1,0,1,2,2,2,2,1,0};
const int zoff3[64] = {0,0,0,0,1,1,1,1,
    0,0,0,0,0,0,0,0,0,0,0,0,
    1,1,1,1,1,1,1,1,1,1,1,1,
    -1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,
    2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2};
// linear element node order (FourNodeIsoparam.cpp:21-22, EightNodeIsoparamBrick.cpp:18-20)
const double xii4[4] = {-1., 1., 1., -1.};
const double eti4[4] = {-1., -1., 1., 1.};
const double xii8[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
const double eti8[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
const double zti8[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};

struct ShapeOut {
    int n;
    long node[64];
    double fn[64], dx[64], dy[64], dz[64];
};

// 1D uGIMP factor of one candidate (shared by both dimensions' closed forms)
inline bool gimp_in_support(double xp, double lp) { return xp < 2. + lp; }

struct Grid {
    MpmGridDesc d;
    int dim;
    long nnx, nny, nnz;     // nodes per axis
    long nnodes, nelems;
    std::vector<double> xpts, ypts, zpts;
    long node_index(long i, long j, long k) const { return (k * nny + j) * nnx + i; }
    void build(const MpmGridDesc &g) {
        d = g;
        dim = g.analysis == MPMGPU_THREED ? 3 : 2;
        nnx = g.nx + 1; nny = g.ny + 1; nnz = dim == 3 ? g.nz + 1 : 1;
        nnodes = nnx * nny * nnz;
        nelems = (long)g.nx * g.ny * (dim == 3 ? g.nz : 1);
        // node coordinates (Read_MPM/Generators.cpp:1823-1833)
        xpts.resize(nnx); ypts.resize(nny); zpts.resize(nnz);
        for (long i = 0; i < nnx; i++) xpts[i] = g.xmin + (double)i * g.dx;
        for (long i = 0; i < nny; i++) ypts[i] = g.ymin + (double)i * g.dy;
        for (long i = 0; i < nnz; i++) zpts[i] = dim == 3 ? g.zmin + (double)i * g.dz : 0.;
    }
    void elem_ijk(long e, long &i, long &j, long &k) const {
        i = e % d.nx; long r = e / d.nx; j = r % d.ny; k = dim == 3 ? r / d.ny : 0;
    }
};

// ElementBase::PtInElement for rectangles (EightNodeIsoparamBrick.cpp:207-213,
// FourNodeIsoparam.cpp:413-423); element extents from node coordinates (ElementBase.cpp:62)
bool pt_in_element(const Grid &g, long e, const double *p) {
    long i, j, k; g.elem_ijk(e, i, j, k);
    if (p[0] < g.xpts[i] || p[0] >= g.xpts[i + 1]) return false;
    if (p[1] < g.ypts[j] || p[1] >= g.ypts[j + 1]) return false;
    if (g.dim == 3 && (p[2] < g.zpts[k] || p[2] >= g.zpts[k + 1])) return false;
    return true;
}

// GetXiPos (EightNodeIsoparamBrick.cpp:276-281, FourNodeIsoparam.cpp:351-356)
void get_xi(const Grid &g, long e, const double *p, double *xi) {
    long i, j, k; g.elem_ijk(e, i, j, k);
    double xmn = g.xpts[i], xmx = g.xpts[i + 1], ymn = g.ypts[j], ymx = g.ypts[j + 1];
    xi[0] = (2. * p[0] - xmn - xmx) / (xmx - xmn);
    xi[1] = (2. * p[1] - ymn - ymx) / (ymx - ymn);
    if (g.dim == 3) {
        double zmn = g.zpts[k], zmx = g.zpts[k + 1];
        xi[2] = (2. * p[2] - zmn - zmx) / (zmx - zmn);
    } else xi[2] = 0.;
}

// 2D uGIMP (Common/Elements/FourNodeIsoparam.cpp:431-521)
void gimp2d(const Grid &g, long e, const double *xi, const double *lp, bool getDeriv, ShapeOut &o) {
    long ei, ej, ek; g.elem_ijk(e, ei, ej, ek);
    double q1x = 2. - lp[0], q2x = 2. + lp[0];
    double q1y = 2. - lp[1], q2y = 2. + lp[1];
    double inv_size_x = 1. / (4. * lp[0]);
    double inv_size_y = 1. / (4. * lp[1]);
    double inv_dx = 0, inv_dy = 0;
    if (getDeriv) {
        inv_dx = 2.0 / (g.xpts[ei + 1] - g.xpts[ei]);
        inv_dy = 2.0 / (g.ypts[ej + 1] - g.ypts[ej]);
    }
    int i = 0;
    for (int id = 0; id < 16; id++) {
        double gx = 2. * xoff2[id] - 1., gy = 2. * yoff2[id] - 1.;
        double xp = fabs(xi[0] - gx);
        if (xp >= q2x) continue;
        double yp = fabs(xi[1] - gy);
        if (yp >= q2y) continue;
        double Svpx, Svpy, argx = 0., argy = 0.;
        if (xp < lp[0]) Svpx = ((4. - lp[0]) * lp[0] - xp * xp) * inv_size_x;
        else if (xp <= q1x) Svpx = (2. - xp) / 2.;
        else { argx = (q2x - xp) * inv_size_x; Svpx = 2. * lp[0] * argx * argx; }
        if (yp < lp[1]) Svpy = ((4. - lp[1]) * lp[1] - yp * yp) * inv_size_y;
        else if (yp <= q1y) Svpy = (2. - yp) / 2.;
        else { argy = (q2y - yp) * inv_size_y; Svpy = 2. * lp[1] * argy * argy; }
        o.fn[i] = Svpx * Svpy;
        if (getDeriv) {
            double xsign = xi[0] > gx ? 1. : -1.;
            double ysign = xi[1] > gy ? 1. : -1.;
            double dSvpx, dSvpy;
            if (xp < lp[0]) dSvpx = -xp * inv_size_x * 2.0;
            else if (xp <= q1x) dSvpx = -0.5;
            else dSvpx = -argx;
            if (yp < lp[1]) dSvpy = -yp * inv_size_y * 2.0;
            else if (yp <= q1y) dSvpy = -0.5;
            else dSvpy = -argy;
            o.dx[i] = xsign * dSvpx * Svpy * inv_dx;
            o.dy[i] = ysign * Svpx * dSvpy * inv_dy;
            o.dz[i] = 0.;
        }
        o.node[i] = g.node_index(ei + xoff2[id], ej + yoff2[id], 0);
        i++;
    }
    o.n = i;
}

// 3D uGIMP (NairnMPM/src/Elements/EightNodeIsoparamBrick.cpp:289-400), including the
// reference's inv_size_z = 1/(4*lp.y) (line 303)
void gimp3d(const Grid &g, long e, const double *xi, const double *lp, bool getDeriv, ShapeOut &o) {
    long ei, ej, ek; g.elem_ijk(e, ei, ej, ek);
    double q1x = 2. - lp[0], q2x = 2. + lp[0];
    double q1y = 2. - lp[1], q2y = 2. + lp[1];
    double q1z = 2. - lp[2], q2z = 2. + lp[2];
    double inv_size_x = 1. / (4. * lp[0]);
    double inv_size_y = 1. / (4. * lp[1]);
    double inv_size_z = 1. / (4. * lp[1]);
    double inv_dx = 0, inv_dy = 0, inv_dz = 0;
    if (getDeriv) {
        inv_dx = 2.0 / (g.xpts[ei + 1] - g.xpts[ei]);
        inv_dy = 2.0 / (g.ypts[ej + 1] - g.ypts[ej]);
        inv_dz = 2.0 / (g.zpts[ek + 1] - g.zpts[ek]);
    }
    int i = 0;
    for (int id = 0; id < 64; id++) {
        double gx = 2. * xoff3[id] - 1., gy = 2. * yoff3[id] - 1., gz = 2. * zoff3[id] - 1.;
        double xp = fabs(xi[0] - gx);
        if (xp >= q2x) continue;
        double yp = fabs(xi[1] - gy);
        if (yp >= q2y) continue;
        double zp = fabs(xi[2] - gz);
        if (zp >= q2z) continue;
        double Svpx, Svpy, Svpz, argx = 0., argy = 0., argz = 0.;
        if (xp < lp[0]) Svpx = ((4. - lp[0]) * lp[0] - xp * xp) * inv_size_x;
        else if (xp <= q1x) Svpx = 0.5 * (2. - xp);
        else { argx = (q2x - xp) * inv_size_x; Svpx = 2. * lp[0] * argx * argx; }
        if (yp < lp[1]) Svpy = ((4. - lp[1]) * lp[1] - yp * yp) * inv_size_y;
        else if (yp <= q1y) Svpy = 0.5 * (2. - yp);
        else { argy = (q2y - yp) * inv_size_y; Svpy = 2. * lp[1] * argy * argy; }
        if (zp < lp[2]) Svpz = ((4. - lp[2]) * lp[2] - zp * zp) * inv_size_z;
        else if (zp <= q1z) Svpz = 0.5 * (2. - zp);
        else { argz = (q2z - zp) * inv_size_z; Svpz = 2. * lp[2] * argz * argz; }
        o.fn[i] = Svpx * Svpy * Svpz;
        if (getDeriv) {
            double xsign = xi[0] > gx ? 1. : -1.;
            double ysign = xi[1] > gy ? 1. : -1.;
            double zsign = xi[2] > gz ? 1. : -1.;
            double dSvpx, dSvpy, dSvpz;
            if (xp < lp[0]) dSvpx = -xp / (2. * lp[0]);
            else if (xp <= q1x) dSvpx = -0.5;
            else dSvpx = -argx;
            if (yp < lp[1]) dSvpy = -yp / (2. * lp[1]);
            else if (yp <= q1y) dSvpy = -0.5;
            else dSvpy = -argy;
            if (zp < lp[2]) dSvpz = -zp / (2. * lp[2]);
            else if (zp <= q1z) dSvpz = -0.5;
            else dSvpz = -argz;
            o.dx[i] = xsign * dSvpx * Svpy * Svpz * inv_dx;
            o.dy[i] = ysign * Svpx * dSvpy * Svpz * inv_dy;
            o.dz[i] = zsign * Svpx * Svpy * dSvpz * inv_dz;
        }
        o.node[i] = g.node_index(ei + xoff3[id], ej + yoff3[id], ek + zoff3[id]);
        i++;
    }
    o.n = i;
}

// classic linear shape functions on the particle's element
// (FourNodeIsoparam.cpp:189-214, EightNodeIsoparamBrick.cpp:87-108)
void linear_shape(const Grid &g, long e, const double *xi, bool getDeriv, ShapeOut &o) {
    long ei, ej, ek; g.elem_ijk(e, ei, ej, ek);
    if (g.dim == 2) {
        double dx = g.xpts[ei + 1] - g.xpts[ei], dy = g.ypts[ej + 1] - g.ypts[ej];
        const int ox[4] = {0, 1, 1, 0}, oy[4] = {0, 0, 1, 1};
        for (int i = 0; i < 4; i++) {
            double temp1 = (1. + xii4[i] * xi[0]);
            double temp2 = (1. + eti4[i] * xi[1]);
            o.fn[i] = 0.25 * temp1 * temp2;
            if (getDeriv) {
                o.dx[i] = 0.5 * xii4[i] * temp2 / dx;
                o.dy[i] = 0.5 * eti4[i] * temp1 / dy;
                o.dz[i] = 0.;
            }
            o.node[i] = g.node_index(ei + ox[i], ej + oy[i], 0);
        }
        o.n = 4;
    } else {
        double dx = g.xpts[ei + 1] - g.xpts[ei], dy = g.ypts[ej + 1] - g.ypts[ej];
        double dz = g.zpts[ek + 1] - g.zpts[ek];
        const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0}, oy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
        const int oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
        for (int i = 0; i < 8; i++) {
            double temp1 = (1. + xii8[i] * xi[0]);
            double temp2 = (1. + eti8[i] * xi[1]);
            double temp3 = (1. + zti8[i] * xi[2]);
            o.fn[i] = 0.125 * temp1 * temp2 * temp3;
            if (getDeriv) {
                o.dx[i] = 0.25 * xii8[i] * temp2 * temp3 / dx;
                o.dy[i] = 0.25 * eti8[i] * temp1 * temp3 / dy;
                o.dz[i] = 0.25 * zti8[i] * temp1 * temp2 / dz;
            }
            o.node[i] = g.node_index(ei + ox[i], ej + oy[i], ek + oz[i]);
        }
        o.n = 8;
    }
}

// ---------------------------------------------------------------------------------------
// Particle state accessor over the SoA buffers
struct PState {
    double *f[MPF_NUM_FIELDS];
    int32_t *iv[MPI_NUM_FIELDS];
};

struct Oracle {
    Grid g;
    std::vector<MpmMaterial> mats;
    MpmRunParams run;
    long np = 0;
    std::vector<double> pd;     // MPF_NUM_FIELDS * np
    std::vector<int32_t> pi;    // MPI_NUM_FIELDS * np
    std::vector<double> xi;     // 3*np: ncpos of the step
    // node fields: one MatVelocityField per material field (MatVelocityField.hpp:40-124),
    // node n of field f at index f*nnodes + n (vectors interleaved x,y,z)
    int nf = 1;
    std::vector<double> mass, pk, ftot, vk, pkc;
    std::vector<int> npts;
    // multimaterial contact extrapolations (ContactTerms: volume, displacement, volume gradient)
    std::vector<double> cvol, cdisp, cgrad;
    std::vector<long> contactNodes;     // MaterialContactNode::materialContactNodes
    int64_t steps = 0, leftGrid = 0, nanCount = 0;
    double mtime = 0.;
    std::string err;

    // Threaded timing mode (nthreads > 1, the CPU baseline only; parity checks run serial):
    // particle loops split into contiguous per-thread ranges; the scatters add into per-thread
    // copies of the node arrays, summed into the shared arrays in thread order afterwards, as the
    // reference's patches sum their ghost nodes after each scatter (GridPatch.cpp:143-195). The
    // grid sums then differ from the serial ones in rounding only.
    int nthreads = 1;
    struct Acc { double *mass, *pk, *ftot, *cvol, *cdisp, *cgrad; int *npts; };
    enum { ACC_MASS = 1, ACC_PK = 2, ACC_FTOT = 4, ACC_NPTS = 8, ACC_CONTACT = 16 };
    std::vector<std::vector<double>> tb;   // per thread: mass, pk(3), ftot(3), cvol, cdisp(3), cgrad(3)
    std::vector<std::vector<int>> tnp;

    // grid velocity conditions in NodalVelBC list order (firstVelocityBC), their values at the
    // step's start time and whether each acts (GetNodeNum > 0)
    std::vector<MpmVelocityBC> vbcs;
    std::vector<double> bcValue;
    std::vector<char> bcOn;
    enum { BC_MASS_MOMENTUM = 0, BC_GRID_FORCES = 1, BC_UPDATE_MOMENTUM = 2, BC_UPDATE_STRAINS_LAST = 3 };
    // external particle loads in MatPtLoadBC list order (firstLoadedPt) and the particles' pFext
    std::vector<MpmParticleLoad> loads;
    std::vector<double> pfext;          // 3 per particle

    // BoundaryCondition::BCValue (BoundaryCondition.cpp:72-104); acts = GetNodeNum != 0 (:124-137)
    static double style_value(int style, double value, double ftime, double t, bool &acts) {
        acts = true;
        switch (style) {
            case 1: acts = t >= ftime; return acts ? value : 0.;
            case 2: acts = t >= ftime; return acts ? 0. + value * (t - ftime) : 0.;
            case 3: return value * sin(t * ftime);
            case 4: return value * cos(t * ftime);
            default: acts = false; return 0.;
        }
    }

    // MatPtLoadBC::SetParticleFext (MatPtLoadBC.cpp:210-221): ZeroMPLoad for every load, then
    // AddMPLoad (:91-102) adds BCValue at the step's start time to its component
    void set_particle_fext() {
        if (loads.empty()) return;
        for (const MpmParticleLoad &l : loads)
            for (int c = 0; c < 3; c++) pfext[3 * l.particle + c] = 0.;
        for (const MpmParticleLoad &l : loads) {
            bool acts;
            pfext[3 * l.particle + (l.dir - 1)] += style_value(l.style, l.value, l.ftime, mtime, acts);
        }
    }

    // traction conditions in MatPtTractionBC list order (firstTractionPt)
    std::vector<MpmParticleTraction> tractions;

    // MatPoint2D::GetSurfaceInfo (MatPoint2D.cpp:1138-1300, planar) and MatPoint3D::GetSurfaceInfo
    // (MatPoint3D.cpp:628-851): the corners of the face (two in 2D, four in 3D, in the reference's
    // order) and hat n times half the face length times the thickness (2D) or a quarter of the
    // face area (3D). uGIMP and classic particles (not CPDI, not finite GIMP) take the corners of the
    // undeformed particle (GetUndeformedSemiSides, MatPoint3D.cpp:528-534) and the size of the
    // deformed face.
    int surface_info(long p, int face, int dof, double thick, double cx[4][3], double wn[3]) {
        Mat3 Fm = get_F(p);
        const double (*pF)[3] = Fm.m;
        long i, j, k; g.elem_ijk(I(MPI_ELEM, p), i, j, k);
        // GetSemiSideVectors with GetParticleSize (MatPoint3D.cpp:413-434, MatPoint2D.cpp:423-440)
        double psx = (g.xpts[i + 1] - g.xpts[i]) * (0.5 * F(MPF_LP_X, p));
        double psy = (g.ypts[j + 1] - g.ypts[j]) * (0.5 * F(MPF_LP_Y, p));
        double psz = g.dim == 3 ? (g.zpts[k + 1] - g.zpts[k]) * (0.5 * F(MPF_LP_Z, p)) : 0.;
        double r1[3] = {pF[0][0] * psx, pF[1][0] * psx, g.dim == 3 ? pF[2][0] * psx : 0.};
        double r2[3] = {pF[0][1] * psy, pF[1][1] * psy, g.dim == 3 ? pF[2][1] * psy : 0.};
        double r3[3] = {pF[0][2] * psz, pF[1][2] * psz, pF[2][2] * psz};
        if (g.dim == 2) r3[0] = r3[1] = r3[2] = 0.;
        auto cross = [](const double *a, const double *b, double *c) {     // CrossProduct
            c[0] = a[1] * b[2] - b[1] * a[2];
            c[1] = b[0] * a[2] - a[0] * b[2];
            c[2] = a[0] * b[1] - b[0] * a[1];
        };
        auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        double uG = -1.;
        if (run.shape != MPMGPU_SHAPE_LCPDI) {
            if (g.dim == 3) {
                double ap[3];
                switch (face) {
                    case 1: case 3: cross(r1, r3, ap); break;
                    case 4: case 2: cross(r2, r3, ap); break;
                    default: cross(r1, r2, ap); break;
                }
                uG = sqrt(dot(ap, ap));
            } else {
                uG = (face == 1 || face == 3) ? sqrt(dot(r1, r1)) : sqrt(dot(r2, r2));
            }
            for (int a = 0; a < 3; a++) r1[a] = r2[a] = r3[a] = 0.;
            r1[0] = psx; r2[1] = psy; r3[2] = psz;
        }
        const double pos[3] = {F(MPF_POS_X, p), F(MPF_POS_Y, p), g.dim == 3 ? F(MPF_POS_Z, p) : 0.};
        // each corner as pos -+ r1 -+ r2 (-+ r3), evaluated left to right as the reference writes it
        auto corner = [&](double *c, double s1, double s2, double s3) {
            for (int a = 0; a < 3; a++) {
                double v = pos[a];
                v = s1 > 0. ? v + r1[a] : v - r1[a];
                v = s2 > 0. ? v + r2[a] : v - r2[a];
                if (g.dim == 3) v = s3 > 0. ? v + r3[a] : v - r3[a];
                c[a] = g.dim == 3 || a < 2 ? v : 0.;
            }
        };
        int nc;
        if (g.dim == 3) {
            nc = 4;
            switch (face) {
                case 1:  corner(cx[0], -1, -1, -1); corner(cx[1], 1, -1, -1); corner(cx[2], 1, -1, 1); corner(cx[3], -1, -1, 1); break;
                case 2:  corner(cx[0], 1, -1, -1); corner(cx[1], 1, 1, -1); corner(cx[2], 1, 1, 1); corner(cx[3], 1, -1, 1); break;
                case 3:  corner(cx[0], 1, 1, -1); corner(cx[1], -1, 1, -1); corner(cx[2], -1, 1, 1); corner(cx[3], 1, 1, 1); break;
                case 4:  corner(cx[0], -1, 1, -1); corner(cx[1], -1, -1, -1); corner(cx[2], -1, -1, 1); corner(cx[3], -1, 1, 1); break;
                case 5:  corner(cx[0], 1, -1, -1); corner(cx[1], -1, -1, -1); corner(cx[2], -1, 1, -1); corner(cx[3], 1, 1, -1); break;
                default: corner(cx[0], -1, -1, 1); corner(cx[1], 1, -1, 1); corner(cx[2], 1, 1, 1); corner(cx[3], -1, 1, 1); break;
            }
        } else {
            nc = 2;
            switch (face) {
                case 1:  corner(cx[0], -1, -1, 0); corner(cx[1], 1, -1, 0); break;
                case 2:  corner(cx[0], 1, -1, 0); corner(cx[1], 1, 1, 0); break;
                case 3:  corner(cx[0], 1, 1, 0); corner(cx[1], -1, 1, 0); break;
                default: corner(cx[0], -1, 1, 0); corner(cx[1], -1, -1, 0); break;
            }
        }
        double r0[3] = {0.5 * (cx[1][0] - cx[0][0]), 0.5 * (cx[1][1] - cx[0][1]), 0.5 * (cx[1][2] - cx[0][2])};
        wn[0] = wn[1] = wn[2] = 0.;
        if (g.dim == 3) {
            double rb[3] = {0.5 * (cx[3][0] - cx[0][0]), 0.5 * (cx[3][1] - cx[0][1]), 0.5 * (cx[3][2] - cx[0][2])};
            double ap[3];
            cross(r0, rb, ap);
            const double faceWt = uG < 0. ? sqrt(dot(ap, ap)) : uG;
            switch (dof) {
                case 1: wn[0] = faceWt; break;
                case 2: wn[1] = faceWt; break;
                case 11:
                    if (uG > 0.) {
                        const double sc = uG / sqrt(dot(ap, ap));
                        for (int a = 0; a < 3; a++) ap[a] *= sc;
                    }
                    for (int a = 0; a < 3; a++) wn[a] = ap[a];
                    break;
                default: wn[2] = faceWt; break;
            }
        } else {
            double faceWt = uG < 0. ? sqrt(dot(r0, r0)) : uG;
            faceWt *= thick;
            double ex, ey, en;
            switch (dof) {
                case 1: wn[0] = faceWt; break;
                case 2: wn[1] = faceWt; break;
                case 11:
                    ex = r0[1]; ey = -r0[0]; en = sqrt(ex * ex + ey * ey);
                    wn[0] = ex * faceWt / en; wn[1] = ey * faceWt / en;
                    break;
                case 12:
                    ex = r0[0]; ey = r0[1]; en = sqrt(ex * ex + ey * ey);
                    wn[0] = ex * faceWt / en; wn[1] = ey * faceWt / en;
                    break;
                default: wn[2] = faceWt; break;
            }
        }
        return nc;
    }

    // MatPtTractionBC::SetParticleSurfaceTractions and AddMPFluxBC (MatPtTractionBC.cpp:64-240,
    // no cracks, not ExactTractions, not axisymmetric), called from PostForcesTask.cpp:51: for each
    // condition in list order and each face corner, the linear shape functions of the corner's
    // element (GetShapeFunctionsForTractions, MoreMPMElementBase.cpp:91-107) times tmag times the
    // weighted normal go to the particle's material field at the element's nodes where that field
    // has points (NodalPoint::AddTractionTask3, NodalPointMPM.cpp:644-650); one rescaling pass
    // only, so the net shape function is not renormalised.
    void add_tractions() {
        for (const MpmParticleTraction &t : tractions) {
            const long p = (long)t.particle;
            bool acts;
            const double tmag = style_value(t.style, t.value, t.ftime, mtime, acts);
            const int fld = field(p);
            double cx[4][3], wn[3];
            const int nc = surface_info(p, t.face, t.dir, t.thickness, cx, wn);
            for (int c = 0; c < nc; c++) {
                const long e = find_element(cx[c]);
                if (e < 0) throw std::runtime_error("A Traction edge node has left the grid");
                double xi[3];
                get_xi(g, e, cx[c], xi);
                ShapeOut o;
                linear_shape(g, e, xi, false, o);
                for (int i = 0; i < o.n; i++) {
                    const long n = fnode(fld, o.node[i]);
                    if (npts[n] <= 0) continue;
                    const double s = tmag * o.fn[i];
                    for (int a = 0; a < 3; a++) ftot[3 * n + a] += s * wn[a];
                }
            }
        }
    }

    // lCPDI corners of the step (MPMBase CPDIDomain: inElem, ncpos, wg; ws = 1/8 or 1/4)
    struct Cpdi { long elem[8]; double xi[8][3]; double wg[8][3]; };
    std::vector<Cpdi> cpdi;

    double &F(int f, long p) { return pd[(size_t)f * np + p]; }
    int32_t &I(int f, long p) { return pi[(size_t)f * np + p]; }
    bool mm() const { return nf > 1; }
    long fnode(int f, long n) const { return (long)f * g.nnodes + n; }
    int field(long p) { return mm() ? mats[I(MPI_MAT, p)].field : 0; }

    // MatPoint3D/MatPoint2D::GetVolume(DEFORMED_AREA) (MatPoint3D.cpp:405-409, MatPoint2D.cpp:407-417)
    double volume(long p) {
        Mat3 pF = get_F(p);
        double rho = mats[I(MPI_MAT, p)].rho;
        double mp = F(MPF_MASS, p);
        const double (*f)[3] = pF.m;
        if (g.dim == 3)
            return (f[0][0] * (f[1][1] * f[2][2] - f[2][1] * f[1][2])
                  - f[0][1] * (f[1][0] * f[2][2] - f[2][0] * f[1][2])
                  + f[0][2] * (f[1][0] * f[2][1] - f[2][0] * f[1][1])) * mp / rho;
        return (f[0][0] * f[1][1] - f[1][0] * f[0][1]) * mp / rho;
    }

    // MeshInfo::FindElementFromPoint for a regular grid (MeshInfo.cpp:593-633); -1 = left the grid
    long find_element(const double *c) {
        const MpmGridDesc &d = g.d;
        long col = (long)(int)((c[0] - d.xmin) / d.dx);
        if (col < 0 || col >= d.nx) { if (c[0] == d.xmin + d.nx * d.dx) col = d.nx - 1; else return -1; }
        long row = (long)(int)((c[1] - d.ymin) / d.dy);
        if (row < 0 || row >= d.ny) { if (c[1] == d.ymin + d.ny * d.dy) row = d.ny - 1; else return -1; }
        if (g.dim == 3) {
            long zr = (long)(int)((c[2] - d.zmin) / d.dz);
            if (zr < 0 || zr >= d.nz) { if (c[2] == d.zmin + d.nz * d.dz) zr = d.nz - 1; else return -1; }
            return (long)d.nx * (zr * d.ny + row) + col;
        }
        return row * d.nx + col;
    }

    // MatPoint3D::GetCPDINodesAndWeights (MatPoint3D.cpp:538-627), MatPoint2D (MatPoint2D.cpp:519-571)
    // with GetSemiSideVectors (MatPoint3D.cpp:413-434, MatPoint2D.cpp:423-440); no rcrit rescaling
    bool cpdi_init(long p) {
        Mat3 pFm = get_F(p);
        const double (*pF)[3] = pFm.m;
        long i, j, k; g.elem_ijk(I(MPI_ELEM, p), i, j, k);
        double pos[3] = {F(MPF_POS_X, p), F(MPF_POS_Y, p), F(MPF_POS_Z, p)};
        Cpdi &cd = cpdi[p];
        double psx = (g.xpts[i + 1] - g.xpts[i]) * (0.5 * F(MPF_LP_X, p));   // GetParticleSize
        double psy = (g.ypts[j + 1] - g.ypts[j]) * (0.5 * F(MPF_LP_Y, p));
        if (g.dim == 3) {
            static const double r1s[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
            static const double r2s[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
            static const double r3s[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};
            double psz = (g.zpts[k + 1] - g.zpts[k]) * (0.5 * F(MPF_LP_Z, p));
            double r1[3] = {pF[0][0] * psx, pF[1][0] * psx, pF[2][0] * psx};
            double r2[3] = {pF[0][1] * psy, pF[1][1] * psy, pF[2][1] * psy};
            double r3[3] = {pF[0][2] * psz, pF[1][2] * psz, pF[2][2] * psz};
            double Vp = 8. * (r1[0] * (r2[1] * r3[2] - r2[2] * r3[1])
                            + r1[1] * (r2[2] * r3[0] - r2[0] * r3[2])
                            + r1[2] * (r2[0] * r3[1] - r2[1] * r3[0]));
            for (int c = 0; c < 8; c++) {
                double cx[3];
                for (int a = 0; a < 3; a++) cx[a] = pos[a] + r1s[c] * r1[a] + r2s[c] * r2[a] + r3s[c] * r3[a];
                cd.elem[c] = find_element(cx);
                if (cd.elem[c] < 0) return false;
                get_xi(g, cd.elem[c], cx, cd.xi[c]);
            }
            Vp = 1. / Vp;
            const double x1 = r1[0], y1 = r1[1], z1 = r1[2], x2 = r2[0], y2 = r2[1], z2 = r2[2];
            const double x3 = r3[0], y3 = r3[1], z3 = r3[2];
            double (*w)[3] = cd.wg;
            w[0][0] = (z1*y2 - y1*z2 - z1*y3 + z2*y3 + y1*z3 - y2*z3)*Vp;
            w[0][1] = (-(z1*x2) + x1*z2 + z1*x3 - z2*x3 - x1*z3 + x2*z3)*Vp;
            w[0][2] = (y1*x2 - x1*y2 - y1*x3 + y2*x3 + x1*y3 - x2*y3)*Vp;
            w[1][0] = (z1*y2 - y1*z2 - z1*y3 - z2*y3 + y1*z3 + y2*z3)*Vp;
            w[1][1] = (-(z1*x2) + x1*z2 + z1*x3 + z2*x3 - x1*z3 - x2*z3)*Vp;
            w[1][2] = (y1*x2 - x1*y2 - y1*x3 - y2*x3 + x1*y3 + x2*y3)*Vp;
            w[2][0] = (z1*y2 - y1*z2 + z1*y3 - z2*y3 - y1*z3 + y2*z3)*Vp;
            w[2][1] = (-(z1*x2) + x1*z2 - z1*x3 + z2*x3 + x1*z3 - x2*z3)*Vp;
            w[2][2] = (y1*x2 - x1*y2 + y1*x3 - y2*x3 - x1*y3 + x2*y3)*Vp;
            w[3][0] = (z1*y2 - y1*z2 + z1*y3 + z2*y3 - y1*z3 - y2*z3)*Vp;
            w[3][1] = (-(z1*x2) + x1*z2 - z1*x3 - z2*x3 + x1*z3 + x2*z3)*Vp;
            w[3][2] = (y1*x2 - x1*y2 + y1*x3 + y2*x3 - x1*y3 - x2*y3)*Vp;
            w[4][0] = (-(z1*y2) + y1*z2 - z1*y3 + z2*y3 + y1*z3 - y2*z3)*Vp;
            w[4][1] = (z1*x2 - x1*z2 + z1*x3 - z2*x3 - x1*z3 + x2*z3)*Vp;
            w[4][2] = (-(y1*x2) + x1*y2 - y1*x3 + y2*x3 + x1*y3 - x2*y3)*Vp;
            w[5][0] = (-(z1*y2) + y1*z2 - z1*y3 - z2*y3 + y1*z3 + y2*z3)*Vp;
            w[5][1] = (z1*x2 - x1*z2 + z1*x3 + z2*x3 - x1*z3 - x2*z3)*Vp;
            w[5][2] = (-(y1*x2) + x1*y2 - y1*x3 - y2*x3 + x1*y3 + x2*y3)*Vp;
            w[6][0] = (-(z1*y2) + y1*z2 + z1*y3 - z2*y3 - y1*z3 + y2*z3)*Vp;
            w[6][1] = (z1*x2 - x1*z2 - z1*x3 + z2*x3 + x1*z3 - x2*z3)*Vp;
            w[6][2] = (-(y1*x2) + x1*y2 + y1*x3 - y2*x3 - x1*y3 + x2*y3)*Vp;
            w[7][0] = (-(z1*y2) + y1*z2 + z1*y3 + z2*y3 - y1*z3 - y2*z3)*Vp;
            w[7][1] = (z1*x2 - x1*z2 - z1*x3 - z2*x3 + x1*z3 + x2*z3)*Vp;
            w[7][2] = (-(y1*x2) + x1*y2 + y1*x3 + y2*x3 - x1*y3 - x2*y3)*Vp;
        } else {
            double r1[2] = {pF[0][0] * psx, pF[1][0] * psx};
            double r2[2] = {pF[0][1] * psy, pF[1][1] * psy};
            double Ap = 4. * (r1[0] * r2[1] - r1[1] * r2[0]);
            double cs[4][3] = {{pos[0] - r1[0] - r2[0], pos[1] - r1[1] - r2[1], 0.},
                               {pos[0] + r1[0] - r2[0], pos[1] + r1[1] - r2[1], 0.},
                               {pos[0] + r1[0] + r2[0], pos[1] + r1[1] + r2[1], 0.},
                               {pos[0] - r1[0] + r2[0], pos[1] - r1[1] + r2[1], 0.}};
            for (int c = 0; c < 4; c++) {
                cd.elem[c] = find_element(cs[c]);
                if (cd.elem[c] < 0) return false;
                get_xi(g, cd.elem[c], cs[c], cd.xi[c]);
            }
            Ap = 1. / Ap;
            cd.wg[0][0] = (r1[1] - r2[1]) * Ap;
            cd.wg[1][0] = (r1[1] + r2[1]) * Ap;
            cd.wg[2][0] = (-r1[1] + r2[1]) * Ap;
            cd.wg[3][0] = (-r1[1] - r2[1]) * Ap;
            cd.wg[0][1] = (-r1[0] + r2[0]) * Ap;
            cd.wg[1][1] = (-r1[0] - r2[0]) * Ap;
            cd.wg[2][1] = (r1[0] - r2[0]) * Ap;
            cd.wg[3][1] = (r1[0] + r2[0]) * Ap;
            for (int c = 0; c < 4; c++) cd.wg[c][2] = 0.;
        }
        return true;
    }

    // ElementBase::GetCPDIFunctions (MoreMPMElementBase.cpp:581-656): linear element functions at
    // each corner (GetShapeFunctionsForTractions, :91-108) weighted by ws and wg, merged per node
    // in order of first appearance; functions below 1e-15 are dropped
    void cpdi_shape(long p, bool deriv, ShapeOut &o) {
        static const double xii[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
        static const double eti[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
        static const double zti[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};
        const Cpdi &cd = cpdi[p];
        int ncorner = g.dim == 3 ? 8 : 4;
        double ws = g.dim == 3 ? 0.125 : 0.25;
        int ind = 0;
        for (int c = 0; c < ncorner; c++) {
            long ei, ej, ek; g.elem_ijk(cd.elem[c], ei, ej, ek);
            const double *x = cd.xi[c];
            for (int jn = 0; jn < ncorner; jn++) {
                double cfn;
                if (g.dim == 3) cfn = 0.125 * (1. + xii[jn] * x[0]) * (1. + eti[jn] * x[1]) * (1. + zti[jn] * x[2]);
                else cfn = 0.25 * (1. + xii[jn] * x[0]) * (1. + eti[jn] * x[1]);
                if (cfn < 1e-15) continue;
                long node = g.node_index(ei + (xii[jn] > 0. ? 1 : 0), ej + (eti[jn] > 0. ? 1 : 0),
                                         ek + (g.dim == 3 && zti[jn] > 0. ? 1 : 0));
                int look = ind - 1;
                for (; look >= 0; look--) if (o.node[look] == node) break;
                if (look >= 0) {
                    o.fn[look] += ws * cfn;
                    if (deriv) { o.dx[look] += cd.wg[c][0] * cfn; o.dy[look] += cd.wg[c][1] * cfn; o.dz[look] += cd.wg[c][2] * cfn; }
                } else {
                    o.node[ind] = node;
                    o.fn[ind] = ws * cfn;
                    if (deriv) { o.dx[ind] = cd.wg[c][0] * cfn; o.dy[ind] = cd.wg[c][1] * cfn; o.dz[ind] = cd.wg[c][2] * cfn; }
                    ind++;
                }
            }
        }
        o.n = ind;
    }

    Acc shared_acc() { return {mass.data(), pk.data(), ftot.data(), cvol.data(), cdisp.data(), cgrad.data(), npts.data()}; }

    // body(p, acc) for every particle, its grid sums going into acc (the `which` arrays)
    template <class Body>
    void scatter(int which, Body body) {
        if (nthreads <= 1) {
            Acc a = shared_acc();
            for (long p = 0; p < np; p++) body(p, a);
            return;
        }
        const size_t nn = (size_t)nf * g.nnodes;
        if ((int)tb.size() < nthreads) {
            tb.resize(nthreads);
            tnp.resize(nthreads);
        }
        int used = 1;
#pragma omp parallel num_threads(nthreads)
        {
            const int t = omp_get_thread_num(), T = omp_get_num_threads();
#pragma omp single
            used = T;
            std::vector<double> &b = tb[t];
            if (b.size() != 14 * nn) b.assign(14 * nn, 0.);
            if (tnp[t].size() != nn) tnp[t].assign(nn, 0);
            Acc a = {b.data(), b.data() + nn, b.data() + 4 * nn, b.data() + 7 * nn, b.data() + 8 * nn,
                     b.data() + 11 * nn, tnp[t].data()};
            if (which & ACC_MASS) std::fill(a.mass, a.mass + nn, 0.);
            if (which & ACC_PK) std::fill(a.pk, a.pk + 3 * nn, 0.);
            if (which & ACC_FTOT) std::fill(a.ftot, a.ftot + 3 * nn, 0.);
            if (which & ACC_NPTS) std::fill(a.npts, a.npts + nn, 0);
            if (which & ACC_CONTACT) std::fill(a.cvol, a.cvol + 7 * nn, 0.);
            const long p0 = np * t / T, p1 = np * (t + 1) / T;
            for (long p = p0; p < p1; p++) body(p, a);
        }
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long i = 0; i < (long)nn; i++) {
            for (int t = 0; t < used; t++) {
                const double *b = tb[t].data();
                if (which & ACC_MASS) mass[i] += b[i];
                if (which & ACC_PK) for (int c = 0; c < 3; c++) pk[3 * i + c] += b[nn + 3 * i + c];
                if (which & ACC_FTOT) for (int c = 0; c < 3; c++) ftot[3 * i + c] += b[4 * nn + 3 * i + c];
                if (which & ACC_NPTS) npts[i] += tnp[t][i];
                if (which & ACC_CONTACT) {
                    cvol[i] += b[7 * nn + i];
                    for (int c = 0; c < 3; c++) {
                        cdisp[3 * i + c] += b[8 * nn + 3 * i + c];
                        cgrad[3 * i + c] += b[11 * nn + 3 * i + c];
                    }
                }
            }
        }
    }

    // body(p) for every particle; bodies that write only particle p's words
    template <class Body>
    void particles(Body body) {
        if (nthreads <= 1) {
            for (long p = 0; p < np; p++) body(p);
            return;
        }
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long p = 0; p < np; p++) body(p);
    }

    void shape(long p, bool deriv, ShapeOut &o) {
        if (run.shape == MPMGPU_SHAPE_LCPDI) { cpdi_shape(p, deriv, o); return; }
        long e = I(MPI_ELEM, p);
        const double *x = &xi[3 * p];
        double lp[3] = {F(MPF_LP_X, p), F(MPF_LP_Y, p), F(MPF_LP_Z, p)};
        if (run.shape == MPMGPU_SHAPE_UGIMP) {
            if (g.dim == 3) gimp3d(g, e, x, lp, deriv, o); else gimp2d(g, e, x, lp, deriv, o);
        } else linear_shape(g, e, x, deriv, o);
    }

    // InitializationTask::Execute (InitializationTask.cpp:43-85): zero the grid, find xi
    void task_initialize() {
        const long nn = (long)mass.size();
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
        for (long n = 0; n < nn; n++) {
            mass[n] = 0.;
            npts[n] = 0;
            for (int c = 0; c < 3; c++) pk[3 * n + c] = ftot[3 * n + c] = vk[3 * n + c] = pkc[3 * n + c] = 0.;
        }
        std::fill(cvol.begin(), cvol.end(), 0.);
        std::fill(cdisp.begin(), cdisp.end(), 0.);
        std::fill(cgrad.begin(), cgrad.end(), 0.);
        contactNodes.clear();
        if (run.shape == MPMGPU_SHAPE_LCPDI && (long)cpdi.size() != np) cpdi.assign(np, Cpdi());
        long bad = 0;
        particles([&](long p) {
            // ElementBase::GetShapeFunctionData (MoreMPMElementBase.cpp:50-58)
            if (run.shape == MPMGPU_SHAPE_LCPDI && !cpdi_init(p)) {
#pragma omp atomic
                bad++;
                return;
            }
            double pos[3] = {F(MPF_POS_X, p), F(MPF_POS_Y, p), F(MPF_POS_Z, p)};
            get_xi(g, I(MPI_ELEM, p), pos, &xi[3 * p]);
        });
        if (bad) {
            err = "A CPDI particle domain node has left the grid";
            throw std::runtime_error(err);
        }
        set_particle_fext();                 // InitializationTask.cpp:91
    }

    // NodalPoint::AddMassMomentum's multimaterial extras: CrackVelocityField::AddVolumeDisplacement
    // (CrackVelocityField.cpp:395-402) and CrackVelocityFieldMulti::AddVolumeGradient
    // (CrackVelocityFieldMulti.cpp:108-112); displacement = pos - origpos, weights shape*Vp and
    // shape*mp, gradient weight Vp
    static void add_contact_terms(Acc &a, long n, double shape, double Vp, double mp, const double *d,
                                  double dNdx, double dNdy, double dNdz) {
        double wtVolume = shape * Vp;
        a.cvol[n] += wtVolume;
        double fnmp = shape * mp;
        for (int c = 0; c < 3; c++) a.cdisp[3 * n + c] += d[c] * fnmp;
        a.cgrad[3 * n] += dNdx * Vp; a.cgrad[3 * n + 1] += dNdy * Vp; a.cgrad[3 * n + 2] += dNdz * Vp;
    }

    // MassAndMomentumTask::Execute (MassAndMomentumTask.cpp:48-127) ->
    // NodalPoint::AddMassMomentum (NodalPointMPM.cpp:419-453); multimaterial mode extrapolates
    // with gradients because volumeGradientIndex >= 0 (MassAndMomentumTask.cpp:137-140)
    void task_mass_momentum() {
        scatter(ACC_MASS | ACC_PK | ACC_NPTS | (mm() ? ACC_CONTACT : 0), [&](long p, Acc &a) {
            ShapeOut o;
            shape(p, mm(), o);
            int f = field(p);
            double mp = F(MPF_MASS, p);
            double v[3] = {F(MPF_VEL_X, p), F(MPF_VEL_Y, p), F(MPF_VEL_Z, p)};
            double Vp = 0., d[3] = {0., 0., 0.};
            if (mm()) {
                Vp = volume(p);
                for (int c = 0; c < 3; c++) d[c] = F(MPF_POS_X + c, p) - F(MPF_ORIG_X + c, p);
            }
            for (int i = 0; i < o.n; i++) {
                long n = fnode(f, o.node[i]);
                double fnmp = o.fn[i] * mp;
                for (int c = 0; c < 3; c++) a.pk[3 * n + c] += fnmp * v[c];   // AddMomentumTask1
                a.npts[n] += 1;
                if (mm()) add_contact_terms(a, n, o.fn[i], Vp, mp, d, o.dx[i], o.dy[i], g.dim == 3 ? o.dz[i] : 0.);
                a.mass[n] += fnmp;                                             // AddMass
            }
        });
    }

    // PostExtrapolationTask::Execute (PostExtrapolationTask.cpp:46-165): CalcTotalMassAndCount
    // copies pk (MatVelocityField.cpp:158-164) and flags nodes with two or more materials
    // (CrackVelocityFieldMulti.cpp:1620-1637); then contact at MASS_MOMENTUM_CALL (:155)
    void task_post_extrapolation() {
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
        for (long k = 0; k < (long)nf * g.nnodes; k++)
            if (npts[k] >= 1) for (int c = 0; c < 3; c++) pkc[3 * k + c] = pk[3 * k + c];
        if (mm()) {
            for (long n = 0; n < g.nnodes; n++) {
                int numberMaterials = 0;
                for (int f = 0; f < nf; f++) numberMaterials += npts[fnode(f, n)] >= 1 ? 1 : 0;
                if (numberMaterials > 1) contactNodes.push_back(n);
            }
            if (run.normalsMethod >= 5) contact_lists();
        }
        material_contact(CALL_MASS_MOMENTUM);
        // ContactAndMomentaBCs(MASS_MOMENTUM_CALL) (PostExtrapolationTask.cpp:153-159): the
        // conditions change pk but not its copy, which RestoreMomenta brings back
        bc_values();
        velocity_bcs(BC_MASS_MOMENTUM);
    }

    // ---- grid velocity conditions -------------------------------------------------------
    // NodalVelBC::GridVelocityBCValues (NodalVelBC.cpp:293-303, called from
    // PostExtrapolationTask.cpp:154) -> BoundaryCondition::GetNodeNum and BCValue
    // (BoundaryCondition.cpp:72-140) at mtime, the step's start time
    void bc_values() {
        bcValue.assign(vbcs.size(), 0.);
        bcOn.assign(vbcs.size(), 1);
        for (size_t e = 0; e < vbcs.size(); e++) {
            const MpmVelocityBC &b = vbcs[e];
            bool acts;
            bcValue[e] = style_value(b.style, b.value, b.ftime, mtime, acts);
            bcOn[e] = acts;
        }
    }

    // NodalVelBC::GridVelocityConditions -> VelocityBCLoop (NodalVelBC.cpp:321-400): every acting
    // condition first zeroes its component on each field of its node that has points
    // (CrackVelocityFieldSingle/Multi::ZeroVelocityBC), then every one adds mass * value
    // (AddVelocityBC); MatVelocityField::ZeroVelocityBC and AddVelocityBC (MatVelocityField.cpp:490-575)
    // with norm the condition's unit axis, so dot products pick the axis component
    void velocity_bcs(int pass) {
        if (vbcs.empty()) return;
        const double dt = run.timestep;
        for (int add = 0; add < 2; add++) {
            for (size_t e = 0; e < vbcs.size(); e++) {
                if (!bcOn[e]) continue;
                const int a = vbcs[e].dir - 1;
                for (int f = 0; f < nf; f++) {
                    const long k = fnode(f, (long)vbcs[e].node);
                    if (npts[k] <= 0) continue;
                    double &p = pk[3 * k + a], &fo = ftot[3 * k + a];
                    if (!add) {
                        if (pass == BC_GRID_FORCES) {
                            // force so that the updated momentum is zero: -dotf - dotp/deltime
                            double dotf = fo, dotp = p;
                            fo += -dotf - dotp / dt;
                        } else {
                            double dotn = p;
                            p += -dotn;
                            if (pass == BC_UPDATE_MOMENTUM) fo += -dotn / dt;
                        }
                    } else {
                        const double mv = mass[k] * bcValue[e];
                        if (pass == BC_GRID_FORCES) {
                            fo += mv / dt;
                        } else {
                            p += mv;
                            if (pass == BC_UPDATE_MOMENTUM) fo += mv / dt;
                        }
                    }
                }
            }
        }
    }

    // ---- multimaterial contact -----------------------------------------------------------
    enum { CALL_MASS_MOMENTUM = 0, CALL_UPDATE_MOMENTUM = 1, CALL_UPDATE_STRAINS_LAST = 2 };
    static double dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

    // MatVelocityField::ChangeMatMomentum (MatVelocityField.cpp:197-224)
    void change_momentum(long k, const double *delP, int callType, double deltime) {
        for (int c = 0; c < 3; c++) pk[3 * k + c] += delP[c];
        if (callType == CALL_UPDATE_MOMENTUM) {
            double sc = 1. / deltime;
            for (int c = 0; c < 3; c++) ftot[3 * k + c] += delP[c] * sc;
        } else if (callType == CALL_MASS_MOMENTUM) {
            for (int c = 0; c < 3; c++) pkc[3 * k + c] = pk[3 * k + c];
        }
    }

    // CoulombFriction::GetFrictionalDeltaMomentum and GetSslideAcDt (CoulombFriction.cpp:146-282);
    // returns false when the law leaves the momenta unchanged
    bool coulomb(const MpmMaterial &law, double *delPi, const double *norm, double dotn, double deltaDotn,
                 double mred, double deltime, const double *delFi) {
        double coeff = law.friction, coeffStatic = law.frictionStatic;
        // frictionStyle (CoulombFriction.cpp:67-85)
        int style = coeff < 0. ? 0 : (coeff == 0. ? 1 : 2);            // STICK, FRICTIONLESS, FRICTIONAL
        if (coeffStatic > 0. && style == 1) style = 2;
        if (style == 0) return true;
        // skipped for the regression normals, whose separation comes from the particles
        // (CoulombFriction.cpp:162)
        if (delFi != nullptr && run.normalsMethod < 5) {
            double fn = dot(delFi, norm);
            deltaDotn += deltime * (dotn - 0.5 * fn * deltime) / mred;
        }
        bool inContact = deltaDotn < 0. && dotn < 0.;
        if (style == 1) {
            if (inContact) { for (int c = 0; c < 3; c++) delPi[c] = norm[c] * dotn; return true; }
            return false;
        }
        if (!inContact) return false;                                    // HasFreeSeparation
        double dott = 0.;
        double tang[3];
        for (int c = 0; c < 3; c++) tang[c] = delPi[c];
        for (int c = 0; c < 3; c++) tang[c] += norm[c] * (-dotn);
        double tangMag = sqrt(dot(tang, tang));
        if (tangMag > 0.) {
            double sc = 1. / tangMag;
            for (int c = 0; c < 3; c++) tang[c] *= sc;
            dott = dot(delPi, tang);
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

    // ---- regression contact normals (normals methods 5 and 6) ----------------------------
    // MaterialContactNode::GetMaterialContactData (MaterialContactNode.cpp:253-350), serial form:
    // for each contact node, the particles whose shape-function node list holds it, in mpm[]
    // order (start-of-step shape functions); built at the mass-momentum call, used by all three
    std::vector<std::vector<long>> contactLists;
    void contact_lists() {
        contactLists.assign(contactNodes.size(), std::vector<long>());
        std::vector<long> slot(g.nnodes, -1);
        for (size_t c = 0; c < contactNodes.size(); c++) slot[contactNodes[c]] = (long)c;
        for (long p = 0; p < np; p++) {
            ShapeOut o;
            shape(p, false, o);
            for (int i = 0; i < o.n; i++)
                if (slot[o.node[i]] >= 0) contactLists[slot[o.node[i]]].push_back(p);
        }
    }

    // Matrix4::cofactor / determinant / Inverse / Times (Matrix4.cpp:96-103,133-140,206-231): the
    // minor's determinant in Matrix3's 3D form (Matrix3.cpp:1195-1202)
    static double cofactor4(const double m[4][4], int row, int col) {
        double c[3][3];
        int crow = 0;
        for (int i = 0; i < 4; i++) {
            if (i == row) continue;
            int ccol = 0;
            for (int j = 0; j < 4; j++) {
                if (j == col) continue;
                c[crow][ccol++] = m[i][j];
            }
            crow++;
        }
        return c[0][0] * (c[1][1] * c[2][2] - c[2][1] * c[1][2]) - c[1][0] * (c[0][1] * c[2][2] - c[2][1] * c[0][2]) +
               c[2][0] * (c[0][1] * c[1][2] - c[1][1] * c[0][2]);
    }
    static void solve4(const double m[4][4], const double *rhs, double *out) {
        double det = m[0][0] * cofactor4(m, 0, 0) - m[1][0] * cofactor4(m, 1, 0) + m[2][0] * cofactor4(m, 2, 0) -
                     m[3][0] * cofactor4(m, 3, 0);
        double f = 1. / det;
        double inv[4][4];
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) inv[r][c] = ((r + c) & 1 ? -cofactor4(m, c, r) : cofactor4(m, c, r)) * f;
        for (int r = 0; r < 4; r++) out[r] = inv[r][0] * rhs[0] + inv[r][1] * rhs[1] + inv[r][2] * rhs[2] + inv[r][3] * rhs[3];
    }
    // Matrix3(xx,xy,xz,yy,yz,zz) symmetric 3D, Inverse and Times (Matrix3.cpp:107-121,394-408,446-460)
    static void solve3(const double s[6], const double *rhs, double *out) {
        double m[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], s[5]}};
        double inv[3][3] = {{m[1][1] * m[2][2] - m[2][1] * m[1][2], -(m[0][1] * m[2][2] - m[2][1] * m[0][2]),
                             m[0][1] * m[1][2] - m[1][1] * m[0][2]},
                            {-(m[1][0] * m[2][2] - m[2][0] * m[1][2]), m[0][0] * m[2][2] - m[2][0] * m[0][2],
                             -(m[0][0] * m[1][2] - m[1][0] * m[0][2])},
                            {m[1][0] * m[2][1] - m[2][0] * m[1][1], -(m[0][0] * m[2][1] - m[2][0] * m[0][1]),
                             m[0][0] * m[1][1] - m[1][0] * m[0][1]}};
        double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                     m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
        double f = 1. / det;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) inv[r][c] *= f;
        for (int r = 0; r < 3; r++) out[r] = inv[r][0] * rhs[0] + inv[r][1] * rhs[1] + inv[r][2] * rhs[2];
    }

    // MatPoint3D::GetDeformedRadius / MatPoint2D::GetDeformedRadius (MatPoint3D.cpp:497-510,
    // MatPoint2D.cpp:482-497), the ellipse methods: the distance from the particle's center to
    // the surface of its deformed domain along the unit vector n. In 2D the reference applies
    // F^-T where 3D applies F^-1.
    double deformed_radius(long p, const double *n) {
        Mat3 Fm = get_F(p);
        long ei, ej, ek;
        g.elem_ijk(I(MPI_ELEM, p), ei, ej, ek);
        double psx = (g.xpts[ei + 1] - g.xpts[ei]) * (0.5 * F(MPF_LP_X, p));     // GetParticleSize
        double psy = (g.ypts[ej + 1] - g.ypts[ej]) * (0.5 * F(MPF_LP_Y, p));
        if (g.dim == 2) {
            const double (*f)[3] = Fm.m;
            double rdet2D = 1. / (f[0][0] * f[1][1] - f[0][1] * f[1][0]);
            double n0x = (f[1][1] * n[0] - f[1][0] * n[1]) * rdet2D;
            double n0y = (-f[0][1] * n[0] + f[0][0] * n[1]) * rdet2D;
            double rd = sqrt(n0x * n0x / (psx * psx) + n0y * n0y / (psy * psy));
            return 1. / rd;
        }
        double psz = (g.zpts[ek + 1] - g.zpts[ek]) * (0.5 * F(MPF_LP_Z, p));
        const double (*m)[3] = Fm.m;
        double inv[3][3] = {{m[1][1] * m[2][2] - m[2][1] * m[1][2], -(m[0][1] * m[2][2] - m[2][1] * m[0][2]),
                             m[0][1] * m[1][2] - m[1][1] * m[0][2]},
                            {-(m[1][0] * m[2][2] - m[2][0] * m[1][2]), m[0][0] * m[2][2] - m[2][0] * m[0][2],
                             -(m[0][0] * m[1][2] - m[1][0] * m[0][2])},
                            {m[1][0] * m[2][1] - m[2][0] * m[1][1], -(m[0][0] * m[2][1] - m[2][0] * m[0][1]),
                             m[0][0] * m[1][1] - m[1][0] * m[0][1]}};
        double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                     m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
        double f = 1. / det;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) inv[r][c] *= f;
        double n0[3];
        for (int r = 0; r < 3; r++) n0[r] = inv[r][0] * n[0] + inv[r][1] * n[1] + inv[r][2] * n[2];
        double rd = sqrt(n0[0] * n0[0] / (psx * psx) + n0[1] * n0[1] / (psy * psy) + n0[2] * n0[2] / (psz * psz));
        return 1. / rd;
    }

    // the separation from the point cloud along norm (MachineLearning.cpp:195-226, 614-633): the
    // gap between the material-a particles' farthest extent along norm and the others' nearest
    double cloud_separation(const std::vector<long> &list, const int *mis, const double *norm) {
        double supportForA = 1.e30, supportForB = 1.e30;
        for (size_t i = 0; i < list.size(); i++) {
            if (mis[i] == 0) continue;
            long p = list[i];
            double xp[3] = {F(MPF_POS_X, p), F(MPF_POS_Y, p), g.dim == 3 ? F(MPF_POS_Z, p) : 0.};
            if (mis[i] < 0) {
                double t = -dot(norm, xp) - deformed_radius(p, norm);
                if (t < supportForA) supportForA = t;
            } else {
                double t = dot(norm, xp) - deformed_radius(p, norm);
                if (t < supportForB) supportForB = t;
            }
        }
        return supportForA + supportForB;
    }

    // CrackVelocityFieldMulti::LinearRegressionNormal and LogisticRegressionNormal
    // (Nodes/MachineLearning.cpp:58-228, 232-633) for material field a against all other
    // materials of the node's list (b < 0); returns the unit normal and the separation
    // (delMats.x), with the reference's penalties (offset 0.1 for the linear fit; 1e-7 cell^2 on
    // the normal's components and 0 on the offset for the logistic fit), the linear fit's beta
    // doubled as the logistic start, at most 15 Gauss-Newton iterations until
    // 1 - cos(change) <= 1e-5, and the linear normal when the iteration ends in NaN
    void regression_normal(long n, const std::vector<long> &list, int a, bool logistic, double norm[3],
                           double &deln) {
        const long nl = (long)list.size();
        std::vector<int> mis(nl);
        long ni = n % g.nnx, nj = (n / g.nnx) % g.nny, nk = g.dim == 3 ? n / (g.nnx * g.nny) : 0;
        double nd[3] = {g.xpts[ni], g.ypts[nj], g.zpts[nk]};
        for (long i = 0; i < nl; i++) mis[i] = mats[I(MPI_MAT, list[i])].field == a ? -1 : 1;
        double linNorm[3] = {0., 0., 0.};
        double cell = std::min(g.d.dx, std::min(g.d.dy, g.dim == 3 ? g.d.dz : g.d.dy));   // GetMinCellDimension
        double cell2 = cell * cell;
        if (g.dim == 3) {
            double mt[10] = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0.}, rhs[4] = {0., 0., 0., 0.}, lambda[4];
            double normPenalty = 1.e-7 * cell2;
            if (logistic) { for (int i = 0; i < 3; i++) lambda[i] = normPenalty; lambda[3] = 0.; }
            else { lambda[0] = lambda[1] = lambda[2] = 0.; lambda[3] = 0.1; }
            for (long i = 0; i < nl; i++) {
                long p = list[i];
                double x = F(MPF_POS_X, p) - nd[0], y = F(MPF_POS_Y, p) - nd[1], z = F(MPF_POS_Z, p) - nd[2];
                mt[0] += x * x; mt[1] += x * y; mt[2] += x * z; mt[3] += x;
                mt[4] += y * y; mt[5] += y * z; mt[6] += y;
                mt[7] += z * z; mt[8] += z; mt[9] += 1.;
                double mi = (double)mis[i];
                rhs[0] += mi * x; rhs[1] += mi * y; rhs[2] += mi * z; rhs[3] += mi;
            }
            auto mat4 = [&](double out[4][4]) {
                double m[4][4] = {{mt[0] + lambda[0], mt[1], mt[2], mt[3]},
                                  {mt[1], mt[4] + lambda[1], mt[5], mt[6]},
                                  {mt[2], mt[5], mt[7] + lambda[2], mt[8]},
                                  {mt[3], mt[6], mt[8], mt[9] + lambda[3]}};
                for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) out[r][c] = m[r][c];
            };
            double M[4][4], beta[4];
            mat4(M);
            solve4(M, rhs, beta);
            if (!logistic) {
                double magi = sqrt(beta[0] * beta[0] + beta[1] * beta[1] + beta[2] * beta[2]);
                for (int c = 0; c < 3; c++) norm[c] = beta[c] * (1. / magi);
            } else {
                for (int i = 0; i < 4; i++) { beta[i] *= 2.; lambda[i] *= 1.; }
                double prevNorm[3] = {beta[0], beta[1], beta[2]};
                double nmag = sqrt(dot(prevNorm, prevNorm));
                for (int c = 0; c < 3; c++) prevNorm[c] *= 1. / nmag;
                for (int c = 0; c < 3; c++) linNorm[c] = prevNorm[c];
                double deltaOneMinusCos = 1.;
                int iter = 0;
                while (deltaOneMinusCos > 1.e-5 && iter < 15) {
                    for (int i = 0; i < 10; i++) mt[i] = 0.;
                    for (int i = 0; i < 4; i++) rhs[i] = -lambda[i] * beta[i];
                    for (long i = 0; i < nl; i++) {
                        if (mis[i] == 0) continue;
                        long p = list[i];
                        double x = F(MPF_POS_X, p) - nd[0], y = F(MPF_POS_Y, p) - nd[1], z = F(MPF_POS_Z, p) - nd[2];
                        double dotp = x * beta[0] + y * beta[1] + z * beta[2] + beta[3];
                        double arg = 1. + exp(-dotp);
                        double phi = 2. * (arg - 1.) / (arg * arg);
                        double phi2 = phi * phi;
                        mt[0] += phi2 * x * x; mt[1] += phi2 * x * y; mt[2] += phi2 * x * z; mt[3] += phi2 * x;
                        mt[4] += phi2 * y * y; mt[5] += phi2 * y * z; mt[6] += phi2 * y;
                        mt[7] += phi2 * z * z; mt[8] += phi2 * z; mt[9] += phi2;
                        double mifi = ((double)mis[i] - 2. / arg + 1.) * phi;
                        rhs[0] += mifi * x; rhs[1] += mifi * y; rhs[2] += mifi * z; rhs[3] += mifi;
                    }
                    mat4(M);
                    double dbeta[4];
                    solve4(M, rhs, dbeta);
                    for (int i = 0; i < 4; i++) beta[i] += dbeta[i];
                    for (int c = 0; c < 3; c++) norm[c] = beta[c];
                    nmag = sqrt(dot(norm, norm));
                    for (int c = 0; c < 3; c++) norm[c] *= 1. / nmag;
                    deltaOneMinusCos = 1. - dot(norm, prevNorm);
                    iter++;
                    for (int c = 0; c < 3; c++) prevNorm[c] = norm[c];
                }
            }
        } else {
            double mt[6] = {0., 0., 0., 0., 0., 0.}, rhs[3] = {0., 0., 0.};
            double lambda[3] = {0., 0., 0.1};
            if (logistic) { double np2 = 1.e-7 * cell2; lambda[0] = np2; lambda[1] = np2; lambda[2] = 0.; }
            for (long i = 0; i < nl; i++) {
                long p = list[i];
                double x = F(MPF_POS_X, p) - nd[0], y = F(MPF_POS_Y, p) - nd[1];
                mt[0] += x * x; mt[1] += x * y; mt[2] += x;
                mt[3] += y * y; mt[4] += y; mt[5] += 1.;
                double mi = (double)mis[i];
                if (logistic) { rhs[0] += x * mi; rhs[1] += y * mi; rhs[2] += 1. * mi; }
                else { rhs[0] += mi * x; rhs[1] += mi * y; rhs[2] += mi; }
            }
            double beta[3];
            double S[6] = {mt[0] + lambda[0], mt[1], mt[2], mt[3] + lambda[1], mt[4], mt[5] + lambda[2]};
            solve3(S, rhs, beta);
            if (!logistic) {
                norm[0] = beta[0]; norm[1] = beta[1]; norm[2] = 0.;
                double magi = sqrt(dot(norm, norm));
                for (int c = 0; c < 3; c++) norm[c] *= 1. / magi;
            } else {
                for (int i = 0; i < 3; i++) { beta[i] *= 2.; lambda[i] *= 1.; }
                double prevNorm[3] = {beta[0], beta[1], beta[2]};
                double nmag = sqrt(beta[0] * beta[0] + beta[1] * beta[1]);
                prevNorm[0] /= nmag; prevNorm[1] /= nmag;
                linNorm[0] = prevNorm[0]; linNorm[1] = prevNorm[1]; linNorm[2] = 0.;
                double deltaOneMinusCos = 1.;
                int iter = 0;
                double nv[3] = {0., 0., 0.};
                while (deltaOneMinusCos > 1.e-5 && iter < 15) {
                    for (int i = 0; i < 6; i++) mt[i] = 0.;
                    rhs[0] = -lambda[0] * beta[0]; rhs[1] = -lambda[1] * beta[1]; rhs[2] = -lambda[2] * beta[2];
                    for (long i = 0; i < nl; i++) {
                        if (mis[i] == 0) continue;
                        long p = list[i];
                        double x = F(MPF_POS_X, p) - nd[0], y = F(MPF_POS_Y, p) - nd[1], z = 1.;
                        double arg = 1. + exp(-(x * beta[0] + y * beta[1] + z * beta[2]));
                        double phi = 2. * (arg - 1.) / (arg * arg);
                        double phi2 = phi * phi;
                        mt[0] += phi2 * x * x; mt[1] += phi2 * x * y; mt[2] += phi2 * x;
                        mt[3] += phi2 * y * y; mt[4] += phi2 * y; mt[5] += phi2;
                        double mifi = ((double)mis[i] - 2. / arg + 1.) * phi;
                        rhs[0] += mifi * x; rhs[1] += mifi * y; rhs[2] += mifi;
                    }
                    double S2[6] = {mt[0] + lambda[0], mt[1], mt[2], mt[3] + lambda[1], mt[4], mt[5] + lambda[2]};
                    double dbeta[3];
                    solve3(S2, rhs, dbeta);
                    for (int i = 0; i < 3; i++) beta[i] += dbeta[i];
                    for (int i = 0; i < 3; i++) nv[i] = beta[i];
                    nmag = sqrt(nv[0] * nv[0] + nv[1] * nv[1]);
                    nv[0] /= nmag; nv[1] /= nmag;
                    deltaOneMinusCos = 1. - nv[0] * prevNorm[0] - nv[1] * prevNorm[1];
                    iter++;
                    for (int i = 0; i < 3; i++) prevNorm[i] = nv[i];
                }
                norm[0] = nv[0]; norm[1] = nv[1]; norm[2] = 0.;
            }
        }
        if (logistic && norm[0] != norm[0]) for (int c = 0; c < 3; c++) norm[c] = linNorm[c];
        deln = cloud_separation(list, mis.data(), norm);
    }

    // MaterialContactNode::ContactOnKnownNodes (MaterialContactNode.cpp:196-232) ->
    // CrackVelocityFieldMulti::MaterialContactOnCVF / MaterialContactOnCVFLumped
    // (CrackVelocityFieldMulti.cpp:302-671) for non-rigid materials, contact by displacements,
    // normals by GetNormalVector (:960-1069) methods 0-3, no symmetry planes
    void material_contact(int callType) {
        if (!mm()) return;
        const MpmMaterial &law = mats[run.contactLaw];
        bool ignoreContact = law.type == MPMGPU_MAT_CONTACT;
        bool postUpdate = callType == CALL_UPDATE_MOMENTUM;
        double deltime = run.timestep;
        int normals = run.normalsMethod;
        for (size_t ci = 0; ci < contactNodes.size(); ci++) {
            const long n = contactNodes[ci];
            int act[MPMGPU_MAX_MATERIALS];
            int numMats = 0;
            double Pc[3] = {0., 0., 0.}, Mc = 0.;
            for (int f = 0; f < nf; f++) {
                long k = fnode(f, n);
                if (npts[k] > 0) {
                    for (int c = 0; c < 3; c++) Pc[c] += pk[3 * k + c];
                    Mc += mass[k];
                    act[numMats++] = f;
                }
            }
            if (numMats <= 1) continue;
            bool doingPairs = numMats == 2 && normals != 3;
            int miMax = doingPairs ? numMats - 1 : numMats;
            // GetCMDisplacement (:1296-1313), scaled by 1/Mc
            double dispc[3] = {0., 0., 0.};
            for (int a = 0; a < numMats; a++)
                for (int c = 0; c < 3; c++) dispc[c] += cdisp[3 * fnode(act[a], n) + c];
            double sMc = 1. / Mc;
            for (int c = 0; c < 3; c++) dispc[c] *= sMc;
            double volAll = 0.;                          // GetContactVolumeNonrigid (:1281-1288)
            for (int a = 0; a < numMats; a++) volAll += cvol[fnode(act[a], n)];
            for (int mi = 0; mi < miMax; mi++) {
                int i = act[mi];
                long ki = fnode(i, n);
                double massi = mass[ki];
                double voli = cvol[ki];
                double volj = volAll - voli;
                int j = -1;
                double gradj[3] = {0., 0., 0.};
                double maxOtherMaterialVolume = 0.;
                for (int kj = 0; kj < numMats; kj++) {
                    if (kj == mi) continue;
                    int jj = act[kj];
                    long kk = fnode(jj, n);
                    double matVolume = cvol[kk];
                    if (matVolume > maxOtherMaterialVolume) { maxOtherMaterialVolume = matVolume; j = jj; }
                    for (int c = 0; c < 3; c++) gradj[c] += cgrad[3 * kk + c] * -1.;   // GetVolumeGradient(-1)
                }
                if (j < 0) continue;
                long kj = fnode(j, n);
                double massRatio = massi / Mc;
                double mred = 1. - massRatio;
                double delPi[3];
                for (int c = 0; c < 3; c++) delPi[c] = pk[3 * ki + c] * -1.;
                for (int c = 0; c < 3; c++) delPi[c] += Pc[c] * massRatio;
                double norm[3] = {0., 0., 0.}, dotn = 0., deln = 0.;
                if (!ignoreContact) {
                    double gradi[3] = {cgrad[3 * ki] * 1., cgrad[3 * ki + 1] * 1., cgrad[3 * ki + 2] * 1.};
                    if (normals == 0) {                                   // MAXIMUM_VOLUME_GRADIENT
                        double magi = sqrt(dot(gradi, gradi));
                        double magj = sqrt(dot(gradj, gradj));
                        if (magi >= 1. * magj) { double sc = 1. / magi; for (int c = 0; c < 3; c++) norm[c] = gradi[c] * sc; }
                        else { double sc = 1. / magj; for (int c = 0; c < 3; c++) norm[c] = gradj[c] * sc; }
                    } else if (normals == 1) {                            // MAXIMUM_VOLUME
                        if (voli >= volj) for (int c = 0; c < 3; c++) norm[c] = gradi[c];
                        else for (int c = 0; c < 3; c++) norm[c] = gradj[c];
                        double sc = 1. / sqrt(dot(norm, norm));
                        for (int c = 0; c < 3; c++) norm[c] = norm[c] * sc;
                    } else if (normals == 2) {                            // AVERAGE_MAT_VOLUME_GRADIENTS
                        for (int c = 0; c < 3; c++) norm[c] = gradi[c];
                        for (int c = 0; c < 3; c++) norm[c] += gradj[c];
                        double magi = dot(norm, norm);
                        double sc = 1. / sqrt(magi);
                        for (int c = 0; c < 3; c++) norm[c] *= sc;
                    } else if (normals == 3) {                            // EACH_MATERIALS_MASS_GRADIENT
                        for (int c = 0; c < 3; c++) norm[c] = gradi[c];
                        double sc = 1. / sqrt(dot(norm, norm));
                        for (int c = 0; c < 3; c++) norm[c] *= sc;
                    } else if (normals == 4) {                            // SPECIFIED_NORMAL
                        for (int c = 0; c < 3; c++) norm[c] = run.contactNormal[c];
                    } else {                                              // LINEAR/LOGISTIC_REGRESSION
                        regression_normal(n, contactLists[ci], i, normals == 6, norm, deln);
                    }
                    if (norm[0] != norm[0] || norm[1] != norm[1] || norm[2] != norm[2]) continue;
                    dotn = dot(delPi, norm);
                    const bool hasDeln = normals >= 5;     // the regression found the separation
                    double dispa[3], sa = 1. / massi;
                    for (int c = 0; c < 3; c++) dispa[c] = cdisp[3 * ki + c] * sa;
                    double delta[3], sr = 1. / mred;
                    for (int c = 0; c < 3; c++) delta[c] = dispc[c];
                    for (int c = 0; c < 3; c++) delta[c] -= dispa[c];
                    for (int c = 0; c < 3; c++) delta[c] *= sr;
                    double dispb[3];
                    for (int c = 0; c < 3; c++) dispb[c] = delta[c];
                    for (int c = 0; c < 3; c++) dispb[c] += dispa[c];
                    // CrackSurfaceContact::MaterialSeparation by displacements (CrackSurfaceContact.cpp:281-287)
                    if (!hasDeln) deln = dot(dispb, norm) - dot(dispa, norm);
                    mred *= massi;
                    double delFi[3], *delFiPtr = nullptr;
                    if (postUpdate) {
                        double fsum[3] = {0., 0., 0.};                    // GetCMatFtot (:1316-1326)
                        for (int a = 0; a < numMats; a++)
                            for (int c = 0; c < 3; c++) fsum[c] += ftot[3 * fnode(act[a], n) + c];
                        double sc = massi / Mc;
                        for (int c = 0; c < 3; c++) delFi[c] = fsum[c] * sc;
                        for (int c = 0; c < 3; c++) delFi[c] += ftot[3 * ki + c] * -1.;
                        delFiPtr = delFi;
                    }
                    if (!coulomb(law, delPi, norm, dotn, deln, mred, deltime, delFiPtr)) continue;
                }
                change_momentum(ki, delPi, callType, deltime);
                if (doingPairs) {
                    for (int c = 0; c < 3; c++) delPi[c] *= -1.;
                    change_momentum(kj, delPi, callType, deltime);
                }
            }
        }
    }

    // ---- XPIC(k) and FMPM(k): the PeriodicXPIC custom task ------------------------------------
    // the order of the running step: run.xpicOrder in the steps the task's schedule selects, else 0
    int xpicStepOrder = 0;
    int64_t xpicNextStep = 0;            // nextPeriodicStep (PeriodicXPIC.cpp:261)
    double xpicNextTime = 0.;            // nextPeriodicTime (:254)
    std::vector<double> vprev, vnext;    // vk[VSTARPREV_VEC], vk[VSTARNEXT_VEC] (MatVelocityField.hpp:25-27)
    enum { XPIC_STRAIN_UPDATE = 32, XPIC_PARTICLE_UPDATE = 64 };

    // PeriodicXPIC::PrepareForStep and StepCalculation (PeriodicXPIC.cpp:283-312,364-373): the
    // decision for a step is made by Initialize before the first step (mstep = 0, mtime = 0) and by
    // the custom-task pass at the end of the step before (CMAnalysis advances mtime after MPMStep),
    // so the step number tested is this step's and the time tested is mtime + timestep of the
    // decision: dt for the first step, this step's start time for later ones
    void xpic_schedule() {
        xpicStepOrder = 0;
        if (run.xpicOrder <= 0) return;
        const int64_t s = steps + 1;
        if (s == 1) {
            xpicNextStep = run.xpicPeriodSteps;
            xpicNextTime = run.xpicPeriodTime;
        }
        if (run.xpicPeriodSteps > 0) {
            if (s >= xpicNextStep) {
                xpicStepOrder = run.xpicOrder;
                xpicNextStep += run.xpicPeriodSteps;
            }
        } else if (run.xpicPeriodTime > 0.) {
            const double t = s == 1 ? 0. + run.timestep : mtime;
            if (t >= xpicNextTime) {
                xpicStepOrder = run.xpicOrder;
                xpicNextTime = t + run.xpicPeriodTime;
            }
        }
    }

    // XPICExtrapolationTask::Execute (XPICExtrapolationTask.cpp:50-164) for mechanics, one material
    // velocity field per material, no cracks, no contact: v(k) in vk, plus a*dt for XPIC
    void xpic_extrapolation(int option) {
        const int m = xpicStepOrder;
        if (m <= 1) return;
        const double dt = run.timestep;
        const bool fmpm = run.xpicFMPM != 0;
        const long nn = (long)nf * g.nnodes;
        vprev.assign(3 * (size_t)nn, 0.);
        vnext.assign(3 * (size_t)nn, 0.);
        // InitializeXPICData -> MatVelocityField::XPICSupport(INITIALIZE_XPIC) (MatVelocityField.cpp:321-345)
        for (long k = 0; k < nn; k++) {
            if (npts[k] == 0) continue;
            double *vp = &vprev[3 * k], *vs = &vk[3 * k];
            const double *p = &pk[3 * k], *f = &ftot[3 * k];
            if (fmpm) {
                const double sc = 1. / mass[k];
                for (int c = 0; c < 3; c++) vp[c] = sc * p[c];
                for (int c = 0; c < 3; c++) vs[c] = vp[c];
            } else {
                for (int c = 0; c < 3; c++) vp[c] = p[c];
                for (int c = 0; c < 3; c++) vp[c] += f[c] * -dt;
                const double sc = 1. / mass[k];
                for (int c = 0; c < 3; c++) vp[c] *= sc;
                for (int c = 0; c < 3; c++) vs[c] = vp[c];
                const double sa = dt / mass[k];
                for (int c = 0; c < 3; c++) vs[c] += f[c] * sa;
            }
        }
        for (int it = 2; it <= m; it++) {
            // XPICDoubleLoop (:182-211): node i gains mp Si Sj / mi * vprev_j over the particle's nodes
            particles([&](long p) {
                ShapeOut o;
                shape(p, false, o);
                const long off = fnode(field(p), 0);
                const double mp = F(MPF_MASS, p);
                for (int i = 0; i < o.n; i++) {
                    const long ni = off + o.node[i];
                    const double mi = mass[ni];
                    for (int j = 0; j < o.n; j++) {
                        const long nj = off + o.node[j];
                        const double weight = o.fn[i] * o.fn[j];
                        const double w = mp * weight / mi;
                        for (int c = 0; c < 3; c++) vnext[3 * ni + c] += w * vprev[3 * nj + c];
                    }
                }
            });
            // GetDeltaV (:222-235): vprev -= vnext (XPICSupport GET_DELTAV, MatVelocityField.cpp:347-363)
            for (long k = 0; k < nn; k++)
                if (npts[k] != 0)
                    for (int c = 0; c < 3; c++) vprev[3 * k + c] -= vnext[3 * k + c];
            // NodalVelBC::GridVelocityConditions for XPIC passes (NodalVelBC.cpp:327-335) ->
            // MatVelocityField::ZeroVelocityBC (MatVelocityField.cpp:523-534): zero the axis
            // component of the increment; the particle pass of XPIC keeps the lumped-mass reaction
            // in the material force
            for (size_t e = 0; e < vbcs.size(); e++) {
                if (!bcOn[e]) continue;
                const int a = vbcs[e].dir - 1;
                for (int f = 0; f < nf; f++) {
                    const long k = fnode(f, (long)vbcs[e].node);
                    if (npts[k] <= 0) continue;
                    const double dotn = vprev[3 * k + a];
                    vprev[3 * k + a] += -dotn;
                    if (option == XPIC_PARTICLE_UPDATE && !fmpm) ftot[3 * k + a] += -mass[k] * dotn / dt;
                }
            }
            // UpdateXStar (XPICSupport UPDATE_VSTAR, :365-378): v* += Delta v, vnext = 0
            for (long k = 0; k < nn; k++)
                if (npts[k] != 0)
                    for (int c = 0; c < 3; c++) {
                        vk[3 * k + c] += vprev[3 * k + c];
                        vnext[3 * k + c] = 0.;
                    }
        }
    }

    // MatVelocityField::GridValueCalculation (MatVelocityField.cpp:239-251)
    void grid_velocity() {
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
        for (long n = 0; n < (long)nf * g.nnodes; n++) {
            if (npts[n] == 0 || mass[n] == 0.) continue;
            double s = 1. / mass[n];
            for (int c = 0; c < 3; c++) vk[3 * n + c] = s * pk[3 * n + c];
        }
    }

    // UpdateStrainsFirstTask::FullStrainUpdate (UpdateStrainsFirstTask.cpp:101-168); FMPM(k > 1)
    // takes its grid velocities from the XPIC extrapolation (:106-114), every other mode lumped
    void full_strain_update(double strainTime) {
        if (run.xpicFMPM && xpicStepOrder > 1) xpic_extrapolation(XPIC_STRAIN_UPDATE);
        else grid_velocity();
        particles([&](long p) {
            ShapeOut o;
            shape(p, true, o);
            long off = fnode(field(p), 0);
            Mat3 dv;
            if (g.dim == 3) {       // MatPoint3D::ExtrapVelocityGradient (MatPoint3D.cpp:62-94)
                for (int i = 0; i < o.n; i++) {
                    const double *v = &vk[3 * (off + o.node[i])];
                    Mat3 t = Mat3::full(v[0] * o.dx[i], v[0] * o.dy[i], v[0] * o.dz[i],
                                        v[1] * o.dx[i], v[1] * o.dy[i], v[1] * o.dz[i],
                                        v[2] * o.dx[i], v[2] * o.dy[i], v[2] * o.dz[i]);
                    dv.add(t);
                }
            } else {                // MatPoint2D::ExtrapVelocityGradient (MatPoint2D.cpp:57-89)
                for (int i = 0; i < o.n; i++) {
                    const double *v = &vk[3 * (off + o.node[i])];
                    dv.add(Mat3::two(v[0] * o.dx[i], v[0] * o.dy[i], v[1] * o.dx[i], v[1] * o.dy[i], 0.));
                }
            }
            dv.scale(strainTime);
            constitutive(p, dv);
        });
    }

    // ---- constitutive laws ------------------------------------------------------------
    Mat3 get_F(long p) {
        // MatPoint3D::GetDeformationGradient (MatPoint3D.cpp:364-377), MatPoint2D.cpp:369-377
        double exx = F(MPF_EP_XX, p), eyy = F(MPF_EP_YY, p), ezz = F(MPF_EP_ZZ, p);
        double exy = F(MPF_EP_XY, p), wxy = F(MPF_WR_XY, p);
        if (g.dim == 2)
            return Mat3::two(1. + exx, 0.5 * (exy - wxy), 0.5 * (exy + wxy), 1. + eyy, 1. + ezz);
        double exz = F(MPF_EP_XZ, p), eyz = F(MPF_EP_YZ, p), wxz = F(MPF_WR_XZ, p), wyz = F(MPF_WR_YZ, p);
        return Mat3::full(1. + exx, 0.5 * (exy - wxy), 0.5 * (exz - wxz),
                          0.5 * (exy + wxy), 1. + eyy, 0.5 * (eyz - wyz),
                          0.5 * (exz + wxz), 0.5 * (eyz + wyz), 1. + ezz);
    }
    void set_F(long p, const Mat3 &Fm) {
        // SetDeformationGradientMatrix (MatPoint3D.cpp:320-336, MatPoint2D.cpp:310-322)
        F(MPF_EP_XX, p) = Fm.m[0][0] - 1.;
        F(MPF_EP_YY, p) = Fm.m[1][1] - 1.;
        F(MPF_EP_ZZ, p) = Fm.m[2][2] - 1.;
        F(MPF_EP_XY, p) = Fm.m[1][0] + Fm.m[0][1];
        F(MPF_WR_XY, p) = Fm.m[1][0] - Fm.m[0][1];
        if (g.dim == 3) {
            F(MPF_EP_XZ, p) = Fm.m[2][0] + Fm.m[0][2];
            F(MPF_EP_YZ, p) = Fm.m[2][1] + Fm.m[1][2];
            F(MPF_WR_XZ, p) = Fm.m[2][0] - Fm.m[0][2];
            F(MPF_WR_YZ, p) = Fm.m[2][1] - Fm.m[1][2];
        }
    }
    // Elastic::HypoIncrementDeformation (ElasticMPM.cpp:392-403)
    void hypo_increment(long p, const Mat3 &du) {
        Mat3 dF = du.exponential(1);
        Mat3 pF = get_F(p);
        set_F(p, dF.mul(pF));
    }

    void constitutive(long p, const Mat3 &du) {
        const MpmMaterial &m = mats[I(MPI_MAT, p)];
        switch (m.type) {
            case MPMGPU_MAT_ISOTROPIC: isotropic(p, du, m); break;
            case MPMGPU_MAT_ISOPLASTICITY: isoplastic(p, du, m); break;
            case MPMGPU_MAT_NEOHOOKEAN: neohookean(p, du, m); break;
            case MPMGPU_MAT_MOONEY: mooney(p, du, m); break;
            default: break;
        }
    }

    // IsotropicMat::MPMConstitutiveLaw (MoreIsotropicMat.cpp:33-47) small-rotation branch
    void isotropic(long p, const Mat3 &du, const MpmMaterial &m) {
        hypo_increment(p, du);
        double *sxx = &F(MPF_SP_XX, p), *syy = &F(MPF_SP_YY, p), *szz = &F(MPF_SP_ZZ, p);
        double *syz = &F(MPF_SP_YZ, p), *sxz = &F(MPF_SP_XZ, p), *sxy = &F(MPF_SP_XY, p);
        if (g.dim == 3) {
            // SRConstitutiveLaw3D (MoreIsotropicMat.cpp:286-351) with no residual strains
            double dvxx = du.m[0][0], dvyy = du.m[1][1], dvzz = du.m[2][2];
            double dgamxy = du.m[0][1] + du.m[1][0];
            double dgamxz = du.m[0][2] + du.m[2][0];
            double dgamyz = du.m[1][2] + du.m[2][1];
            double dwrotxy = du.m[1][0] - du.m[0][1];
            double dwrotxz = du.m[2][0] - du.m[0][2];
            double dwrotyz = du.m[2][1] - du.m[1][2];
            double d[6];
            d[0] = m.C[0][0] * dvxx + m.C[0][1] * dvyy + m.C[0][2] * dvzz;
            d[1] = m.C[1][0] * dvxx + m.C[1][1] * dvyy + m.C[1][2] * dvzz;
            d[2] = m.C[2][0] * dvxx + m.C[2][1] * dvyy + m.C[2][2] * dvzz;
            d[3] = m.C[3][3] * dgamyz;
            d[4] = m.C[4][4] * dgamxz;
            d[5] = m.C[5][5] * dgamxy;
            // MaterialBase::Hypo3DCalculations (MaterialBaseMPM.cpp:1026-1047)
            double sp[6] = {*sxx, *syy, *szz, *syz, *sxz, *sxy};
            double st[6];
            st[0] = -dwrotxy * sp[5] - dwrotxz * sp[4];
            st[1] = dwrotxy * sp[5] - dwrotyz * sp[3];
            st[2] = dwrotxz * sp[4] + dwrotyz * sp[3];
            st[3] = 0.5 * (dwrotxy * sp[4] + dwrotxz * sp[5] + dwrotyz * (sp[1] - sp[2]));
            st[4] = 0.5 * (-dwrotxy * sp[3] + dwrotxz * (sp[0] - sp[2]) + dwrotyz * sp[5]);
            st[5] = 0.5 * (dwrotxy * (sp[0] - sp[1]) - dwrotxz * sp[3] - dwrotyz * sp[4]);
            *sxx += d[0] + st[0]; *syy += d[1] + st[1]; *szz += d[2] + st[2];
            *syz += d[3] + st[3]; *sxz += d[4] + st[4]; *sxy += d[5] + st[5];
        } else {
            // SRConstitutiveLaw2D plane strain (MoreIsotropicMat.cpp:185-283)
            double dvxx = du.m[0][0], dvyy = du.m[1][1];
            double dgam = du.m[0][1] + du.m[1][0];
            double dwrotxy = du.m[1][0] - du.m[0][1];
            double c1 = m.C[1][1] * dvxx + m.C[1][2] * dvyy;
            double c2 = m.C[1][2] * dvxx + m.C[2][2] * dvyy;
            double c3 = m.C[3][3] * dgam;
            // Hypo2DCalculations (MaterialBaseMPM.cpp:1012-1020)
            double dnorm = dwrotxy * (*sxy);
            double dshear = 0.5 * dwrotxy * (*sxx - *syy);
            *sxx += c1 - dnorm;
            *syy += c2 + dnorm;
            *sxy += c3 + dshear;
            *szz += m.C[4][1] * dvxx + m.C[4][2] * dvyy + m.C[4][4] * 0.;
        }
    }

    // IsoPlasticity::MPMConstitutiveLaw + PlasticityConstLaw (IsoPlasticity.cpp:128-451) with
    // LinearHardening (LinearHardening.cpp:93-147), small rotation, no residual strains,
    // plane strain or 3D
    void isoplastic(long p, const Mat3 &de, const MpmMaterial &m) {
        hypo_increment(p, de);
        bool is2D = g.dim == 2;
        double delV = de.trace() - 3. * 0.;
        double dgxy = de.m[0][1] + de.m[1][0], dgxz = 0., dgyz = 0.;
        if (!is2D) { dgxz = de.m[0][2] + de.m[2][0]; dgyz = de.m[1][2] + de.m[2][1]; }
        double dexxr = de.m[0][0], deyyr = de.m[1][1], dezzr = de.m[2][2];
        // UpdatePressure (IsoPlasticity.cpp:462-492)
        double dP = -m.Kred * delV;
        F(MPF_PRESSURE, p) += dP;
        // rotate plastic strain and stress (IsoPlasticity.cpp:224-277)
        double exx = F(MPF_ALT_XX, p), eyy = F(MPF_ALT_YY, p), ezz = F(MPF_ALT_ZZ, p);
        double eyz = F(MPF_ALT_YZ, p), exz = F(MPF_ALT_XZ, p), exy = F(MPF_ALT_XY, p);
        double sxx = F(MPF_SP_XX, p), syy = F(MPF_SP_YY, p), szz = F(MPF_SP_ZZ, p);
        double syz = F(MPF_SP_YZ, p), sxz = F(MPF_SP_XZ, p), sxy = F(MPF_SP_XY, p);
        double dwrotxy = de.m[1][0] - de.m[0][1];
        double st0[6] = {sxx, syy, szz, syz, sxz, sxy};
        if (is2D) {
            double dnorm = 0.5 * dwrotxy * exy;
            F(MPF_ALT_XX, p) -= dnorm;
            F(MPF_ALT_YY, p) += dnorm;
            F(MPF_ALT_XY, p) += dwrotxy * (exx - eyy);
            double dn = dwrotxy * sxy;
            st0[0] -= dn;
            st0[1] += dn;
            st0[5] += 0.5 * dwrotxy * (exx - eyy);     // reference quirk: ep, not sp (:252)
        } else {
            double dwrotxz = de.m[2][0] - de.m[0][2];
            double dwrotyz = de.m[2][1] - de.m[1][2];
            double dxy = 0.5 * dwrotxy * exy;
            double dxz = 0.5 * dwrotxz * exz;
            double dyz = 0.5 * dwrotyz * eyz;
            F(MPF_ALT_XX, p) += -dxy - dxz;
            F(MPF_ALT_YY, p) += dxy - dyz;
            F(MPF_ALT_ZZ, p) += dxz + dyz;
            F(MPF_ALT_YZ, p) += dwrotyz * (eyy - ezz) + 0.5 * (dwrotxz * exy + dwrotxy * exz);
            F(MPF_ALT_XZ, p) += dwrotxz * (exx - ezz) + 0.5 * (dwrotyz * exy - dwrotxy * eyz);
            F(MPF_ALT_XY, p) += dwrotxy * (exx - eyy) - 0.5 * (dwrotyz * exz + dwrotxz * eyz);
            double sdxy = dwrotxy * sxy;
            double sdxz = dwrotxz * sxz;
            double sdyz = dwrotyz * syz;
            st0[0] += -sdxy - sdxz;
            st0[1] += sdxy - sdyz;
            st0[2] += sdxz + sdyz;
            st0[3] += 0.5 * (dwrotyz * (syy - szz) + dwrotxz * sxy + dwrotxy * sxz);
            st0[4] += 0.5 * (dwrotxz * (sxx - szz) + dwrotyz * sxy - dwrotxy * syz);
            st0[5] += 0.5 * (dwrotxy * (sxx - syy) - dwrotyz * sxz - dwrotxz * syz);
        }
        // deviatoric trial stress (IsoPlasticity.cpp:279-301)
        double thirdDelV = delV / 3.;
        double dels[6] = {0., 0., 0., 0., 0., 0.};
        dels[0] = 2. * m.Gred * (dexxr - thirdDelV);
        dels[1] = 2. * m.Gred * (deyyr - thirdDelV);
        dels[5] = m.Gred * dgxy;
        dels[2] = 2. * m.Gred * (dezzr - thirdDelV);
        if (!is2D) { dels[3] = m.Gred * dgyz; dels[4] = m.Gred * dgxz; }
        double str[6];
        for (int c = 0; c < 6; c++) str[c] = st0[c] + dels[c];
        if (is2D) { str[3] = 0.; str[4] = 0.; }
        // GetPlasticPotential: J2 (IsoPlasticity.cpp:513-517, MoreIsotropicMat.cpp:356-373)
        double alpint = F(MPF_HIST0, p);
        double s = str[0] * str[0] + str[1] * str[1] + str[2] * str[2];
        double t = is2D ? str[5] * str[5] : str[5] * str[5] + str[4] * str[4] + str[3] * str[3];
        double strialmag = sqrt(s + t + t);
        double yld = alpint < m.alphaMax ? m.yldred + m.Epred * alpint : m.yldredMin;
        double ftrial = strialmag - SQRT_TWOTHIRDS * yld;
        if (ftrial < 0.) {
            F(MPF_SP_XX, p) = str[0]; F(MPF_SP_YY, p) = str[1]; F(MPF_SP_ZZ, p) = str[2];
            F(MPF_SP_YZ, p) = str[3]; F(MPF_SP_XZ, p) = str[4]; F(MPF_SP_XY, p) = str[5];
            return;
        }
        // LinearHardening::SolveForLambdaBracketed closed form (LinearHardening.cpp:124-147)
        double lambdak = (strialmag - SQRT_TWOTHIRDS * (m.yldred + m.Epred * alpint)) / (2. * (m.Gred + m.Epred / 3.));
        if (alpint + SQRT_TWOTHIRDS * lambdak > m.alphaMax)
            lambdak = (strialmag - SQRT_TWOTHIRDS * m.yldredMin) / (2. * m.Gred);
        double dalpha = SQRT_TWOTHIRDS * lambdak;           // HardeningLawBase.cpp:131-134
        alpint = F(MPF_HIST0, p) + dalpha;
        // GetDfDsigma (IsoPlasticity.cpp:497-508) and plastic increment (:385-420)
        double dfds[6];
        for (int c = 0; c < 6; c++) dfds[c] = str[c] / strialmag;
        double dep[6];
        dep[0] = lambdak * dfds[0];
        dep[1] = lambdak * dfds[1];
        dep[2] = lambdak * dfds[2];
        dep[5] = 2. * lambdak * dfds[5];
        F(MPF_ALT_XX, p) += dep[0]; F(MPF_ALT_YY, p) += dep[1]; F(MPF_ALT_ZZ, p) += dep[2];
        F(MPF_ALT_XY, p) += dep[5];
        if (!is2D) {
            dep[4] = 2. * lambdak * dfds[4];
            dep[3] = 2. * lambdak * dfds[3];
            F(MPF_ALT_XZ, p) += dep[4]; F(MPF_ALT_YZ, p) += dep[3];
        }
        F(MPF_SP_XX, p) = str[0] - 2. * m.Gred * dep[0];
        F(MPF_SP_YY, p) = str[1] - 2. * m.Gred * dep[1];
        F(MPF_SP_ZZ, p) = str[2] - 2. * m.Gred * dep[2];
        F(MPF_SP_XY, p) = str[5] - m.Gred * dep[5];
        if (!is2D) {
            F(MPF_SP_YZ, p) = str[3] - m.Gred * dep[3];
            F(MPF_SP_XZ, p) = str[4] - m.Gred * dep[4];
        }
        F(MPF_HIST0, p) = alpint;                            // UpdatePlasticInternal (:136-139)
    }

    // HyperElastic::IncrementDeformation (HyperElastic.cpp:104-141): F <- dF.F and the left
    // Cauchy-Green tensor B <- dF.B.dF^T kept in the alt strain; returns det(dF)
    double hyper_increment(long p, const Mat3 &du) {
        bool is3 = g.dim == 3;
        Mat3 dF = du.exponential(is3 ? 1 : 2);     // incrementalDefGradTerms (StartOutput.cpp:116-119)
        Mat3 pF = get_F(p);
        set_F(p, dF.mul(pF));
        double bxx = F(MPF_ALT_XX, p), byy = F(MPF_ALT_YY, p), bzz = F(MPF_ALT_ZZ, p);
        double bxy = F(MPF_ALT_XY, p), bxz = F(MPF_ALT_XZ, p), byz = F(MPF_ALT_YZ, p);
        Mat3 pBold = is3 ? Mat3::full(bxx, bxy, bxz, bxy, byy, byz, bxz, byz, bzz)
                         : Mat3::two(bxx, bxy, bxy, byy, bzz);
        Mat3 dFoldB = dF.mul(pBold);
        const double (*a)[3] = dFoldB.m;
        const double (*f)[3] = dF.m;
        F(MPF_ALT_XX, p) = a[0][0] * f[0][0] + a[0][1] * f[0][1] + a[0][2] * f[0][2];
        F(MPF_ALT_XY, p) = a[0][0] * f[1][0] + a[0][1] * f[1][1] + a[0][2] * f[1][2];
        F(MPF_ALT_YY, p) = a[1][0] * f[1][0] + a[1][1] * f[1][1] + a[1][2] * f[1][2];
        F(MPF_ALT_ZZ, p) = a[2][0] * f[2][0] + a[2][1] * f[2][1] + a[2][2] * f[2][2];
        if (is3) {
            F(MPF_ALT_XZ, p) = a[0][0] * f[2][0] + a[0][1] * f[2][1] + a[0][2] * f[2][2];
            F(MPF_ALT_YZ, p) = a[1][0] * f[2][0] + a[1][1] * f[2][1] + a[1][2] * f[2][2];
        }
        return dF.determinant();
    }

    // HyperElastic::GetVolumetricTerms (HyperElastic.cpp:171-207); option numbers from HyperElastic.hpp:21
    static double volumetric_term(int ujOption, double J, double Kred) {
        double Kterm;
        if (ujOption == 1) Kterm = Kred * (J - 1.);                                  // J_MINUS_1_SQUARED
        else if (ujOption == 2) { double lj = log(J); Kterm = Kred * lj; Kterm /= J; }   // LN_J_SQUARED
        else Kterm = 0.5 * Kred * (J - 1. / J);                                      // default (0)
        return Kterm;
    }

    // Neohookean::MPMConstitutiveLaw (Neohookean.cpp:177-335), default U(J) options, no
    // residual strains, no artificial viscosity, plane strain or 3D
    void neohookean(long p, const Mat3 &du, const MpmMaterial &m) {
        bool is3 = g.dim == 3;
        double detDf = hyper_increment(p, du);
        double Bxx = F(MPF_ALT_XX, p), Byy = F(MPF_ALT_YY, p), Bzz = F(MPF_ALT_ZZ, p), Bxy = F(MPF_ALT_XY, p);
        double Bxz = is3 ? F(MPF_ALT_XZ, p) : 0., Byz = is3 ? F(MPF_ALT_YZ, p) : 0.;
        double Jres = F(MPF_HIST1, p);
        double dJres = exp(0.);                        // GetIncrementalResJ (HyperElastic.cpp:89-95)
        Jres *= dJres;
        F(MPF_HIST1, p) = Jres;
        double resStretch = pow(Jres, 1. / 3.);
        double Jres23 = resStretch * resStretch;
        double J = detDf * F(MPF_HIST0, p);
        F(MPF_HIST0, p) = J;
        double Jeff = J / Jres;
        double Kterm = volumetric_term(m.ujOption, Jeff, m.Lamesp);
        double Pterm = J * Kterm + Jres * m.Gsp * ((Bxx + Byy + Bzz) / (3. * Jres23) - 1.);
        F(MPF_PRESSURE, p) = -Pterm + 0.;
        double GJeff = resStretch * m.Gsp;
        double I1third = (Bxx + Byy + Bzz) / 3.;
        F(MPF_SP_XX, p) = GJeff * (Bxx - I1third);
        F(MPF_SP_YY, p) = GJeff * (Byy - I1third);
        F(MPF_SP_ZZ, p) = GJeff * (Bzz - I1third);
        F(MPF_SP_XY, p) = GJeff * Bxy;
        if (is3) { F(MPF_SP_XZ, p) = GJeff * Bxz; F(MPF_SP_YZ, p) = GJeff * Byz; }
    }

    // Mooney::MPMConstitutiveLaw (Mooney.cpp:185-367), plane strain or 3D, no residual strains,
    // no artificial viscosity (QAVred = 0)
    void mooney(long p, const Mat3 &du, const MpmMaterial &m) {
        bool is3 = g.dim == 3;
        double detDf = hyper_increment(p, du);
        double Bxx = F(MPF_ALT_XX, p), Byy = F(MPF_ALT_YY, p), Bzz = F(MPF_ALT_ZZ, p), Bxy = F(MPF_ALT_XY, p);
        double Jres = F(MPF_HIST1, p);
        double dJres = exp(0.);
        Jres *= dJres;
        F(MPF_HIST1, p) = Jres;
        double J = detDf * F(MPF_HIST0, p);
        F(MPF_HIST0, p) = J;
        double Jeff = J / Jres;
        double Kterm = J * volumetric_term(m.ujOption, Jeff, m.Ksp);   // Kirchhoff pressure
        F(MPF_PRESSURE, p) = -Kterm + 0.;
        double J23 = pow(J, 2. / 3.);
        double J43 = J23 * J23;
        double JforG1 = J23 / Jres;
        double JforG2 = J43 / Jres;
        double sxx = (2 * Bxx - Byy - Bzz) * m.G1sp / (3. * JforG1)
                   + (Bxx * (Byy + Bzz) - 2 * Byy * Bzz - Bxy * Bxy) * m.G2sp / (3. * JforG2);
        double syy = (2 * Byy - Bxx - Bzz) * m.G1sp / (3. * JforG1)
                   + (Byy * (Bxx + Bzz) - 2 * Bxx * Bzz - Bxy * Bxy) * m.G2sp / (3. * JforG2);
        double szz = (2 * Bzz - Bxx - Byy) * m.G1sp / (3. * JforG1)
                   + (Bzz * (Bxx + Byy) - 2 * Bxx * Byy + 2. * Bxy * Bxy) * m.G2sp / (3. * JforG2);
        double sxy = Bxy * m.G1sp / JforG1 + (Bzz * Bxy) * m.G2sp / JforG2;
        if (is3) {
            double Bxz = F(MPF_ALT_XZ, p), Byz = F(MPF_ALT_YZ, p);
            sxx += (2. * Byz * Byz - Bxz * Bxz) * m.G2sp / (3. * JforG2);
            syy += (2. * Bxz * Bxz - Byz * Byz) * m.G2sp / (3. * JforG2);
            szz -= (Bxz * Bxz + Byz * Byz) * m.G2sp / (3. * JforG2);
            sxy -= Bxz * Byz * m.G2sp / JforG2;
            F(MPF_SP_XZ, p) = Bxz * m.G1sp / JforG1 + (Byy * Bxz - Bxy * Byz) * m.G2sp / JforG2;
            F(MPF_SP_YZ, p) = Byz * m.G1sp / JforG1 + (Bxx * Byz - Bxy * Bxz) * m.G2sp / JforG2;
        }
        F(MPF_SP_XX, p) = sxx; F(MPF_SP_YY, p) = syy; F(MPF_SP_ZZ, p) = szz; F(MPF_SP_XY, p) = sxy;
    }

    // GridForcesTask::Execute (GridForcesTask.cpp:40-153) -> GetFintPlusFext
    // (MatPoint3D.cpp:248-270, MatPoint2D.cpp:242-263); pFext is zero without particle loads
    void task_grid_forces() {
        const bool fext = !loads.empty();
        scatter(ACC_FTOT, [&](long p, Acc &a) {
            ShapeOut o;
            shape(p, true, o);
            double mp = F(MPF_MASS, p), P = F(MPF_PRESSURE, p);
            double sxx = F(MPF_SP_XX, p), syy = F(MPF_SP_YY, p), szz = F(MPF_SP_ZZ, p);
            double syz = F(MPF_SP_YZ, p), sxz = F(MPF_SP_XZ, p), sxy = F(MPF_SP_XY, p);
            long off = fnode(field(p), 0);
            for (int i = 0; i < o.n; i++) {
                long n = off + o.node[i];
                double fx, fy, fz;
                if (g.dim == 3) {
                    fx = -mp * ((sxx - P) * o.dx[i] + sxy * o.dy[i] + sxz * o.dz[i]);
                    fy = -mp * (sxy * o.dx[i] + (syy - P) * o.dy[i] + syz * o.dz[i]);
                    fz = -mp * (sxz * o.dx[i] + syz * o.dy[i] + (szz - P) * o.dz[i]);
                } else {
                    fx = -mp * ((sxx - P) * o.dx[i] + sxy * o.dy[i]);
                    fy = -mp * (sxy * o.dx[i] + (syy - P) * o.dy[i]);
                    fz = 0.0;
                }
                if (fext) {                 // + fni * pFext
                    const double *pe = &pfext[3 * p];
                    fx = fx + o.fn[i] * pe[0];
                    fy = fy + o.fn[i] * pe[1];
                    if (g.dim == 3) fz = fz + o.fn[i] * pe[2];
                }
                a.ftot[3 * n] += fx; a.ftot[3 * n + 1] += fy; a.ftot[3 * n + 2] += fz;   // AddFtotTask3
            }
        });
    }

    // PostForcesTask::Execute (PostForcesTask.cpp:43-111) -> RestoreMomenta (MatVelocityField.cpp:177)
    void task_post_forces() {
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
        for (long n = 0; n < (long)nf * g.nnodes; n++)
            if (npts[n] > 0) for (int c = 0; c < 3; c++) pk[3 * n + c] = pkc[3 * n + c];
        add_tractions();                                 // PostForcesTask.cpp:51
        // gravity on every active field (PostForcesTask.cpp:62-91, MatVelocityField.cpp:254-258)
        const double *gv = run.gravity;
        if (gv[0] != 0. || gv[1] != 0. || gv[2] != 0.)
            for (long n = 0; n < (long)nf * g.nnodes; n++)
                if (npts[n] > 0) for (int c = 0; c < 3; c++) ftot[3 * n + c] += mass[n] * gv[c];
        velocity_bcs(BC_GRID_FORCES);                    // PostForcesTask.cpp:94
    }

    // UpdateMomentaTask::Execute (UpdateMomentaTask.cpp:44-64) -> UpdateMomentum (MatVelocityField.cpp:301-314)
    void task_update_momenta() {
        double dt = run.timestep;
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
        for (long n = 0; n < (long)nf * g.nnodes; n++)
            if (npts[n] > 0) for (int c = 0; c < 3; c++) pk[3 * n + c] += ftot[3 * n + c] * dt;
        material_contact(CALL_UPDATE_MOMENTUM);          // ContactAndMomentaBCs (UpdateMomentaTask.cpp:58)
        // (UpdateMomentaTask.cpp:93); skipped in XPIC and FMPM steps of order > 1 (NodalVelBC.cpp:365-370)
        if (xpicStepOrder <= 1) velocity_bcs(BC_UPDATE_MOMENTUM);
    }

    // UpdateParticlesTask::Execute (UpdateParticlesTask.cpp:44-298) -> IncrementNodalVelAcc
    // (MatVelocityField.cpp:410-426) -> MoveParticle (MatPoint3D.cpp:104-194): FLIP (m = 0), FMPM(k)
    // (m = k > 0) or XPIC(k) (m = -k < 0); for k > 1 the grid velocities come from the XPIC
    // extrapolation (UpdateParticlesTask.cpp:66-79)
    void task_update_particles() {
        const int order = xpicStepOrder;
        if (order > 1) xpic_extrapolation(XPIC_PARTICLE_UPDATE);
        else grid_velocity();
        const int m = run.xpicFMPM ? order : -order;
        double dt = run.timestep;
        particles([&](long p) {
            ShapeOut o;
            shape(p, false, o);
            long off = fnode(field(p), 0);
            double Svk[3] = {0., 0., 0.}, Sacc[3] = {0., 0., 0.};
            for (int i = 0; i < o.n; i++) {
                long n = off + o.node[i];
                for (int c = 0; c < 3; c++) Svk[c] += vk[3 * n + c] * o.fn[i];
                if (m <= 0) {
                    double mnode = o.fn[i] / mass[n];
                    for (int c = 0; c < 3; c++) Sacc[c] += ftot[3 * n + c] * mnode;
                }
            }
            int nc = g.dim == 3 ? 3 : 2;
            // the material may replace the particle damping (MaterialBase::GetMaterialDamping,
            // MaterialBaseMPM.cpp:742-744, called at UpdateParticlesTask.cpp:118)
            const MpmMaterial &pm = mats[I(MPI_MAT, p)];
            const double particleAlpha = pm.usePDamping ? pm.pdamping : run.particleDamping;
            for (int c = 0; c < nc; c++) {
                const double vold = F(MPF_VEL_X + c, p);
                // vm: S(v - a dt) for FLIP, Sv+(k) for FMPM, S(v(k) + a dt) for XPIC (:106-116)
                double vm = m == 0 ? Svk[c] - Sacc[c] * dt : Svk[c];
                // Adamp0 = particleAlpha vel + gridAlpha vm (MatPoint3D.cpp:117-120)
                double adamp = vold * particleAlpha + vm * run.gridDamping;
                if (m > 0) {
                    // FMPM: v = vm - Adamp0 dt, x += 0.5 (v(n+1) + v(n)) dt (:122-137)
                    double vnew = vm + adamp * -dt;
                    F(MPF_VEL_X + c, p) = vnew;
                    double delXRate = vold + vnew;
                    F(MPF_POS_X + c, p) += delXRate * (0.5 * dt);
                } else if (m == 0) {
                    double delV = (Sacc[c] - adamp) * dt;
                    F(MPF_VEL_X + c, p) += delV;
                    double delXRate = vm + 0.5 * delV;
                    F(MPF_POS_X + c, p) += delXRate * dt;
                } else {
                    // XPIC: v = Svk - Adamp0 dt, x += (vm + 0.5 delV) dt (:152-170)
                    double vnew = Svk[c] - adamp * dt;
                    double delV = vnew - vold;
                    F(MPF_VEL_X + c, p) = vnew;
                    double delXRate = vm + 0.5 * delV;
                    F(MPF_POS_X + c, p) += delXRate * dt;
                }
            }
        });
    }

    // UpdateStrainsLastContactTask::Execute (UpdateStrainsLastContactTask.cpp:60-152):
    // RezeroNodeTask6 + AddMassMomentumLast (NodalPointMPM.cpp:470-486) with start-of-step xi
    void task_update_strains_last() {
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
        for (long n = 0; n < (long)nf * g.nnodes; n++)
            if (npts[n] > 0) {                           // MatVelocityField::RezeroNodeTask6 (:429-444)
                for (int c = 0; c < 3; c++) pk[3 * n + c] = 0.;
                if (mm()) {
                    cvol[n] = 0.;
                    for (int c = 0; c < 3; c++) { cdisp[3 * n + c] = 0.; cgrad[3 * n + c] = 0.; }
                }
            }
        scatter(ACC_PK | (mm() ? ACC_CONTACT : 0), [&](long p, Acc &a) {
            ShapeOut o;
            shape(p, mm(), o);
            int f = field(p);
            double mp = F(MPF_MASS, p);
            double v[3] = {F(MPF_VEL_X, p), F(MPF_VEL_Y, p), F(MPF_VEL_Z, p)};
            double Vp = 0., d[3] = {0., 0., 0.};
            if (mm()) {
                Vp = volume(p);
                for (int c = 0; c < 3; c++) d[c] = F(MPF_POS_X + c, p) - F(MPF_ORIG_X + c, p);
            }
            for (int i = 0; i < o.n; i++) {
                long n = fnode(f, o.node[i]);
                double fnmp = o.fn[i] * mp;
                for (int c = 0; c < 3; c++) a.pk[3 * n + c] += v[c] * fnmp;   // AddMomentumTask6
                if (mm()) add_contact_terms(a, n, o.fn[i], Vp, mp, d, o.dx[i], o.dy[i], g.dim == 3 ? o.dz[i] : 0.);
            }
        });
        material_contact(CALL_UPDATE_STRAINS_LAST);      // UpdateStrainsLastContactTask.cpp:145
        velocity_bcs(BC_UPDATE_STRAINS_LAST);
        full_strain_update(run.strainLast);
    }

    // ResetElementsTask::Execute / ResetElement / ReturnToElement (ResetElementsTask.cpp:44-265),
    // MeshInfo::FindElementFromPoint (MeshInfo.cpp:593-640), EdgeElement2D/3D (MeshInfo.cpp:171-199).
    // Stops the run as the reference's CommonException does (ResetElementsTask.cpp:71-151): at the
    // first NaN position (the DeleteLimit default), when the pushed-back count reaches maxLeaving,
    // or when a leaving particle would be deleted (deleteLeaving, not supported). Serially, the
    // particles after the one that stops the run are not reset, as in the reference's loop.
    std::string abortMsg;
    void task_reset_elements() {
        abortMsg.clear();
        if (nthreads <= 1) {
            for (long p = 0; p < np; p++)
                if (!reset_element(p)) break;
        } else {
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (long p = 0; p < np; p++) reset_element(p);
        }
        if (!abortMsg.empty()) throw std::runtime_error(abortMsg);
    }
    void stop_run(const char *m) {
#pragma omp critical(oracle_abort)
        if (abortMsg.empty()) abortMsg = m;
    }
    bool reset_element(long p) {
        const MpmGridDesc &d = g.d;
        long horiz = d.nx, vert = d.ny, depth = g.dim == 3 ? d.nz : 1;
        long totalElems = horiz * vert * depth;
        {
            double pos[3] = {F(MPF_POS_X, p), F(MPF_POS_Y, p), F(MPF_POS_Z, p)};
            if (pos[0] != pos[0] || pos[1] != pos[1] || pos[2] != pos[2]) {
#pragma omp atomic
                nanCount++;
                stop_run("Particle left grid with position nan.");
                return false;
            }
            long e = I(MPI_ELEM, p);
            if (pt_in_element(g, e, pos)) return true;
            bool left = false;
            long col = (long)(int)((pos[0] - d.xmin) / d.dx);
            if (col < 0 || col >= horiz) {
                if (pos[0] == d.xmin + horiz * d.dx) col = horiz - 1; else left = true;
            }
            long row = (long)(int)((pos[1] - d.ymin) / d.dy);
            if (!left && (row < 0 || row >= vert)) {
                if (pos[1] == d.ymin + vert * d.dy) row = vert - 1; else left = true;
            }
            long zrow = 0;
            if (!left && g.dim == 3) {
                zrow = (long)(int)((pos[2] - d.zmin) / d.dz);
                if (zrow < 0 || zrow >= depth) {
                    if (pos[2] == d.zmin + depth * d.dz) zrow = depth - 1; else left = true;
                }
            }
            long num = 0;
            if (!left) {
                num = (g.dim == 3 ? horiz * (zrow * vert + row) + col : row * horiz + col) + 1;
                if (g.dim == 3) {
                    if (num <= horiz * vert || num > totalElems - horiz * vert) left = true;
                    else if (num % horiz <= 1) left = true;
                    else { long xz = num % (horiz * vert); if (xz <= horiz || xz > horiz * (vert - 1)) left = true; }
                } else {
                    if (num <= horiz || num > totalElems - horiz) left = true;
                    else if (num % horiz <= 1) left = true;
                }
            }
            if (!left) { I(MPI_ELEM, p) = (int32_t)(num - 1); return true; }
            if (!I(MPI_LEFT, p)) {
                int64_t count;
#pragma omp atomic capture
                count = ++leftGrid;
                if (run.maxLeaving > 0 && count >= run.maxLeaving) {
                    stop_run("Too many particles have left the grid");
                    return false;
                }
                I(MPI_LEFT, p) = 1;
            }
            if (run.deleteLeaving) {
                stop_run("A particle has left the grid and would be deleted into the reservoir, which is not "
                         "supported (set <LeaveLimit> above 1 to push leaving particles back)");
                return false;
            }
            // ReturnToElement
            double dt = run.timestep;
            double outside[3] = {pos[0], pos[1], pos[2]};
            double inside[3] = {outside[0] - dt * F(MPF_VEL_X, p), outside[1] - dt * F(MPF_VEL_Y, p),
                                outside[2] - dt * F(MPF_VEL_Z, p)};
            if (!pt_in_element(g, e, inside)) {
                long i, j, k; g.elem_ijk(e, i, j, k);
                inside[0] = (g.xpts[i] + g.xpts[i + 1]) / 2.;
                inside[1] = (g.ypts[j] + g.ypts[j + 1]) / 2.;
                inside[2] = g.dim == 3 ? (g.zpts[k] + g.zpts[k + 1]) / 2. : 0.;
            }
            for (int pass = 1; pass <= 10; pass++) {
                double mid[3] = {(outside[0] + inside[0]) / 2., (outside[1] + inside[1]) / 2.,
                                 (outside[2] + inside[2]) / 2.};
                if (pt_in_element(g, e, mid)) { inside[0] = mid[0]; inside[1] = mid[1]; inside[2] = mid[2]; }
                else { outside[0] = mid[0]; outside[1] = mid[1]; outside[2] = mid[2]; }
            }
            F(MPF_POS_X, p) = inside[0]; F(MPF_POS_Y, p) = inside[1];
            if (g.dim == 3) F(MPF_POS_Z, p) = inside[2];
        }
        return true;
    }

    // NairnMPM::MPMStep (NairnMPM.cpp:284-335) with the task list of CreateTasks (:870-1110)
    void step() {
        xpic_schedule();
        task_initialize();
        task_mass_momentum();
        task_post_extrapolation();
        full_strain_update(run.strainFirst);     // UpdateStrainsFirstTask (UpdateStrainsFirstTask.cpp:52)
        task_grid_forces();
        task_post_forces();
        task_update_momenta();
        task_update_particles();
        task_update_strains_last();
        task_reset_elements();
        steps++;
        mtime += run.timestep;
    }
};

}  // namespace

extern "C" {

void *oracle_create(const MpmGridDesc *grid, const MpmMaterial *mats, int32_t nmat,
                    const MpmRunParams *run, int64_t n, const double *dfields, const int32_t *ifields) {
    Oracle *o = new Oracle();
    o->g.build(*grid);
    o->mats.assign(mats, mats + nmat);
    o->run = *run;
    o->np = n;
    o->pd.assign(dfields, dfields + (size_t)MPF_NUM_FIELDS * n);
    o->pi.assign(ifields, ifields + (size_t)MPI_NUM_FIELDS * n);
    o->xi.assign(3 * (size_t)n, 0.);
    size_t nn = (size_t)o->g.nnodes;
    o->nf = run->multiMaterial && run->numFields > 1 ? run->numFields : 1;
    size_t fnn = (size_t)o->nf * nn;
    o->mass.assign(fnn, 0.); o->npts.assign(fnn, 0);
    o->pk.assign(3 * fnn, 0.); o->ftot.assign(3 * fnn, 0.); o->vk.assign(3 * fnn, 0.); o->pkc.assign(3 * fnn, 0.);
    if (o->nf > 1) { o->cvol.assign(fnn, 0.); o->cdisp.assign(3 * fnn, 0.); o->cgrad.assign(3 * fnn, 0.); }
    return o;
}

void oracle_destroy(void *h) { delete (Oracle *)h; }

int oracle_step(void *h, int32_t nsteps) {
    Oracle *o = (Oracle *)h;
    try {
        for (int s = 0; s < nsteps; s++) o->step();
    } catch (const std::exception &e) {
        o->err = e.what();
        return 1;
    }
    return 0;
}

/* message of the last failed oracle_step / oracle_run_task */
const char *oracle_error(void *h) { return ((Oracle *)h)->err.c_str(); }

/* run one task of the step (same ids as MPMGPU_TASK_*) */
int oracle_run_task(void *h, int32_t task) {
    Oracle *o = (Oracle *)h;
    try {
    switch (task) {
        case MPMGPU_TASK_INITIALIZATION: o->xpic_schedule(); o->task_initialize(); break;
        case MPMGPU_TASK_MASS_MOMENTUM: o->task_mass_momentum(); break;
        case MPMGPU_TASK_POST_EXTRAPOLATION: o->task_post_extrapolation(); break;
        case MPMGPU_TASK_UPDATE_STRAINS_FIRST: o->full_strain_update(o->run.strainFirst); break;
        case MPMGPU_TASK_GRID_FORCES: o->task_grid_forces(); break;
        case MPMGPU_TASK_POST_FORCES: o->task_post_forces(); break;
        case MPMGPU_TASK_UPDATE_MOMENTA: o->task_update_momenta(); break;
        case MPMGPU_TASK_UPDATE_PARTICLES: o->task_update_particles(); break;
        case MPMGPU_TASK_UPDATE_STRAINS_LAST: o->task_update_strains_last(); break;
        case MPMGPU_TASK_RESET_ELEMENTS:
            o->task_reset_elements(); o->steps++; o->mtime += o->run.timestep; break;
        default: return 1;
    }
    } catch (const std::exception &e) {
        o->err = e.what();
        return 1;
    }
    return 0;
}

int oracle_get_particles(void *h, double *dfields, int32_t *ifields) {
    Oracle *o = (Oracle *)h;
    memcpy(dfields, o->pd.data(), o->pd.size() * sizeof(double));
    if (ifields) memcpy(ifields, o->pi.data(), o->pi.size() * sizeof(int32_t));
    return 0;
}

// threads of the timing mode (1: the serial oracle the parity checks use)
/* grid velocity conditions in list order (n = 0 removes them); nodes must be in the grid */
int oracle_set_velocity_bcs(void *h, int64_t n, const MpmVelocityBC *bcs) {
    Oracle *o = (Oracle *)h;
    for (int64_t i = 0; i < n; i++) {
        if (bcs[i].node < 0 || bcs[i].node >= o->g.nnodes) return 1;
        if (bcs[i].dir < 1 || bcs[i].dir > o->g.dim) return 1;
        if (bcs[i].style < 1 || bcs[i].style > 4) return 1;
    }
    o->vbcs.assign(bcs, bcs + n);
    o->bc_values();                      // values at the current time until the next step's
    return 0;
}

/* the XPIC or FMPM order the last step ran with (0: FLIP) */
int32_t oracle_xpic_order(void *h) { return ((Oracle *)h)->xpicStepOrder; }

/* traction conditions in list order (n = 0 removes them) */
int oracle_set_particle_tractions(void *h, int64_t n, const MpmParticleTraction *tr) {
    Oracle *o = (Oracle *)h;
    for (int64_t i = 0; i < n; i++) {
        const MpmParticleTraction &t = tr[i];
        if (t.particle < 0 || t.particle >= o->np) return 1;
        if (o->g.dim == 3 && !((t.dir >= 1 && t.dir <= 3) || t.dir == 11)) return 1;
        if (o->g.dim == 2 && !(t.dir == 1 || t.dir == 2 || t.dir == 11 || t.dir == 12)) return 1;
        if (t.face < 1 || t.face > (o->g.dim == 3 ? 6 : 4)) return 1;
        if (t.style < 1 || t.style > 4) return 1;
    }
    o->tractions.assign(tr, tr + n);
    return 0;
}

/* the face corners and weighted normal of GetSurfaceInfo for particle p; returns the corner count */
int oracle_surface_info(void *h, int64_t p, int32_t face, int32_t dir, double thickness, double *corners,
                        double *wn) {
    Oracle *o = (Oracle *)h;
    double cx[4][3];
    const int nc = o->surface_info((long)p, face, dir, thickness, cx, wn);
    for (int c = 0; c < nc; c++)
        for (int a = 0; a < 3; a++) corners[3 * c + a] = cx[c][a];
    return nc;
}

/* external particle loads in list order (n = 0 removes them) */
int oracle_set_particle_loads(void *h, int64_t n, const MpmParticleLoad *loads) {
    Oracle *o = (Oracle *)h;
    for (int64_t i = 0; i < n; i++) {
        if (loads[i].particle < 0 || loads[i].particle >= o->np) return 1;
        if (loads[i].dir < 1 || loads[i].dir > o->g.dim) return 1;
        if (loads[i].style < 1 || loads[i].style > 4) return 1;
    }
    o->loads.assign(loads, loads + n);
    o->pfext.assign(3 * (size_t)o->np, 0.);
    o->set_particle_fext();
    return 0;
}

int oracle_set_threads(void *h, int32_t n) {
    ((Oracle *)h)->nthreads = n < 1 ? 1 : n;
    return 0;
}

int64_t oracle_num_nodes(void *h) { return ((Oracle *)h)->g.nnodes; }
int32_t oracle_num_fields(void *h) { return ((Oracle *)h)->nf; }

/* node values of every field: field f node n at f*nnodes + n (vectors interleaved) */
int oracle_get_grid(void *h, double *mass, double *pk, double *ftot, int32_t *npts) {
    Oracle *o = (Oracle *)h;
    size_t nn = (size_t)o->nf * o->g.nnodes;
    if (mass) memcpy(mass, o->mass.data(), nn * sizeof(double));
    if (pk) memcpy(pk, o->pk.data(), 3 * nn * sizeof(double));
    if (ftot) memcpy(ftot, o->ftot.data(), 3 * nn * sizeof(double));
    if (npts) memcpy(npts, o->npts.data(), nn * sizeof(int32_t));
    return 0;
}

/* the grid velocities vk of every field (x,y,z per field node) */
int oracle_get_velocities(void *h, double *vk) {
    Oracle *o = (Oracle *)h;
    memcpy(vk, o->vk.data(), 3 * (size_t)o->nf * o->g.nnodes * sizeof(double));
    return 0;
}

/* lCPDI shape functions of one particle (unit tests): state holds pos, lp and F (ep/wr fields);
   returns the node count or -1 when a corner leaves the grid */
int oracle_cpdi_shape(const MpmGridDesc *grid, int64_t elem, const double *state, int32_t getDeriv,
                      int64_t *nodes, double *fn, double *dx, double *dy, double *dz) {
    Oracle o;
    o.g.build(*grid);
    o.run.shape = MPMGPU_SHAPE_LCPDI;
    o.np = 1;
    o.pd.assign(state, state + MPF_NUM_FIELDS);
    o.pi.assign(MPI_NUM_FIELDS, 0);
    o.pi[MPI_ELEM] = (int32_t)elem;
    o.cpdi.assign(1, Oracle::Cpdi());
    if (!o.cpdi_init(0)) return -1;
    ShapeOut so;
    o.cpdi_shape(0, getDeriv != 0, so);
    for (int i = 0; i < so.n; i++) {
        nodes[i] = so.node[i]; fn[i] = so.fn[i];
        if (getDeriv) { dx[i] = so.dx[i]; dy[i] = so.dy[i]; dz[i] = so.dz[i]; }
    }
    return so.n;
}

/* Lumped multimaterial contact on one node (unit tests): nf material velocity fields with mass,
   point count, contact volume, mass-weighted displacement and volume gradient, and momentum,
   force and momentum copy (vectors interleaved x,y,z) updated in place by one contact pass of
   callType (0 mass-momentum, 1 update-momentum, 2 strains-last) with contact law `law`, normals
   method `normals` and time step dt (MaterialContactNode::ContactOnKnownNodes,
   MaterialContactNode.cpp:196-232). */
int oracle_contact_node(const MpmMaterial *law, int32_t normals, int32_t callType, int32_t nf, double dt,
                        const double *mass, const int32_t *npts, const double *cvol, const double *cdisp,
                        const double *cgrad, double *pk, double *ftot, double *pkc) {
    if (nf < 2 || nf > MPMGPU_MAX_MATERIALS || callType < 0 || callType > 2) return 1;
    Oracle o;
    o.g.nnodes = 1;
    o.nf = nf;
    o.mats.assign(law, law + 1);
    memset(&o.run, 0, sizeof(o.run));
    o.run.contactLaw = 0;
    o.run.normalsMethod = normals;
    o.run.timestep = dt;
    o.run.multiMaterial = 1;
    o.run.numFields = nf;
    o.mass.assign(mass, mass + nf);
    o.npts.assign(npts, npts + nf);
    o.cvol.assign(cvol, cvol + nf);
    o.cdisp.assign(cdisp, cdisp + 3 * nf);
    o.cgrad.assign(cgrad, cgrad + 3 * nf);
    o.pk.assign(pk, pk + 3 * nf);
    o.ftot.assign(ftot, ftot + 3 * nf);
    o.pkc.assign(pkc, pkc + 3 * nf);
    o.contactNodes.assign(1, 0L);
    o.material_contact(callType);
    memcpy(pk, o.pk.data(), 3 * nf * sizeof(double));
    memcpy(ftot, o.ftot.data(), 3 * nf * sizeof(double));
    memcpy(pkc, o.pkc.data(), 3 * nf * sizeof(double));
    return 0;
}

/* contact extrapolations of every field (interleaved vectors); 1 in single-material mode */
int oracle_get_contact(void *h, double *vol, double *disp, double *grad) {
    Oracle *o = (Oracle *)h;
    if (o->nf < 2) return 1;
    size_t nn = (size_t)o->nf * o->g.nnodes;
    if (vol) memcpy(vol, o->cvol.data(), nn * sizeof(double));
    if (disp) memcpy(disp, o->cdisp.data(), 3 * nn * sizeof(double));
    if (grad) memcpy(grad, o->cgrad.data(), 3 * nn * sizeof(double));
    return 0;
}

/* Regression normal of one node against the current particles (unit tests of normals methods 5
   and 6): the particles whose start-of-step node lists hold `node` (in mpm[] order), material
   field a against the rest; returns the list length, the unit normal and the separation. */
int64_t oracle_regression_normal(void *h, int64_t node, int32_t a, int32_t logistic, double *norm, double *deln) {
    Oracle *o = (Oracle *)h;
    o->task_initialize();
    o->contactNodes.assign(1, (long)node);
    o->contact_lists();
    double d = 0.;
    o->regression_normal((long)node, o->contactLists[0], a, logistic != 0, norm, d);
    *deln = d;
    return (int64_t)o->contactLists[0].size();
}

/* MatPoint3D/MatPoint2D::GetDeformedRadius of particle p along the unit vector n (unit tests) */
double oracle_deformed_radius(void *h, int64_t p, const double *n) { return ((Oracle *)h)->deformed_radius((long)p, n); }

int oracle_counters(void *h, int64_t *out) {
    Oracle *o = (Oracle *)h;
    out[0] = o->steps; out[1] = o->leftGrid; out[2] = o->nanCount;
    return 0;
}

/* Shape functions of one particle position in element `elem` (unit tests):
   returns the node count; nodes/fn/dx/dy/dz get up to 64 entries. */
int oracle_shape(const MpmGridDesc *grid, int32_t shape, int64_t elem, const double *pos,
                 const double *lp, int32_t getDeriv, int64_t *nodes, double *fn, double *dx,
                 double *dy, double *dz) {
    Grid g; g.build(*grid);
    double xi[3];
    get_xi(g, elem, pos, xi);
    ShapeOut o;
    if (shape == MPMGPU_SHAPE_UGIMP) {
        if (g.dim == 3) gimp3d(g, elem, xi, lp, getDeriv != 0, o); else gimp2d(g, elem, xi, lp, getDeriv != 0, o);
    } else linear_shape(g, elem, xi, getDeriv != 0, o);
    for (int i = 0; i < o.n; i++) {
        nodes[i] = o.node[i]; fn[i] = o.fn[i];
        if (getDeriv) { dx[i] = o.dx[i]; dy[i] = o.dy[i]; dz[i] = o.dz[i]; }
    }
    return o.n;
}

/* One constitutive update of a single particle (unit tests): du is the 3x3 displacement
   gradient increment (row-major), state is MPF_NUM_FIELDS doubles updated in place. */
int oracle_constitutive(int32_t analysis, const MpmMaterial *mat, const double *du, double *state) {
    Oracle o;
    MpmGridDesc gd; memset(&gd, 0, sizeof(gd));
    gd.analysis = analysis; gd.nx = gd.ny = 1; gd.nz = analysis == MPMGPU_THREED ? 1 : 0;
    gd.dx = gd.dy = gd.dz = 1.;
    o.g.build(gd);
    o.mats.assign(mat, mat + 1);
    o.np = 1;
    o.pd.assign(state, state + MPF_NUM_FIELDS);
    o.pi.assign(MPI_NUM_FIELDS, 0);
    Mat3 m;
    if (analysis == MPMGPU_THREED) m = Mat3::full(du[0], du[1], du[2], du[3], du[4], du[5], du[6], du[7], du[8]);
    else m = Mat3::two(du[0], du[1], du[3], du[4], 0.);
    o.constitutive(0, m);
    memcpy(state, o.pd.data(), MPF_NUM_FIELDS * sizeof(double));
    return 0;
}

}  // extern "C"
