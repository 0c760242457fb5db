// mpm_launch.hip -- the launchers the context calls: each runs the kernels compiled for the tile
// shape of its grid argument (namespace mpmk_s0 or mpmk_s1); launchers without a grid argument do
// not depend on the shape and run shape 0's copy.
#include <hip/hip_runtime.h>

#include "mpm_kernels.h"

#define BY_SHAPE(G, CALL) ((G).tileShape ? mpmk_s1::CALL : mpmk_s0::CALL)

void set_tile_cus(int cus) {
    mpmk_s0::set_tile_cus(cus);
    mpmk_s1::set_tile_cus(cus);
}
void launch_bin_keys(int dim, const KGrid &g, const KPart &p, const KBins &b, hipStream_t s) {
    BY_SHAPE(g, launch_bin_keys(dim, g, p, b, s));
}
void launch_bin_finish(int dim, const KGrid &g, const KPart &p, const KBins &b, hipStream_t s) {
    BY_SHAPE(g, launch_bin_finish(dim, g, p, b, s));
}
void launch_p2g_mass(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                     const MpmMaterial *mats, const KStats &st, hipStream_t s) {
    BY_SHAPE(g, launch_p2g_mass(dim, g, p, b, nd, mats, st, s));
}
void launch_node_mass(int dim, const KGrid &g, const KBins &b, const KNodes &nd, hipStream_t s) {
    BY_SHAPE(g, launch_node_mass(dim, g, b, nd, s));
}
void launch_usf(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                const MpmMaterial *mats, hipStream_t s) {
    BY_SHAPE(g, launch_usf(dim, g, p, b, nd, mats, s));
}
void launch_p2g_force(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd, const KLoads &L,
                      int par, hipStream_t s) {
    BY_SHAPE(g, launch_p2g_force(dim, g, p, b, nd, L, par, s));
}
void launch_node_force(int dim, const KGrid &g, const KBins &b, const KNodes &nd, hipStream_t s) {
    BY_SHAPE(g, launch_node_force(dim, g, b, nd, s));
}
void launch_g2p_p2g(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                    const MpmMaterial *mats, int mode, hipStream_t s) {
    BY_SHAPE(g, launch_g2p_p2g(dim, g, p, b, nd, mats, mode, s));
}
void launch_node_mom2(int dim, const KGrid &g, const KBins &b, const KNodes &nd, hipStream_t s) {
    BY_SHAPE(g, launch_node_mom2(dim, g, b, nd, s));
}
void launch_usl_reset(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                      const MpmMaterial *mats, const KStats &st, hipStream_t s) {
    BY_SHAPE(g, launch_usl_reset(dim, g, p, b, nd, mats, st, s));
}
void launch_p2g_contact(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                        const MpmMaterial *mats, bool newPositions, hipStream_t s) {
    BY_SHAPE(g, launch_p2g_contact(dim, g, p, b, nd, mats, newPositions, s));
}
void launch_node_contact(int dim, const KGrid &g, const KBins &b, const KNodes &nd, hipStream_t s) {
    BY_SHAPE(g, launch_node_contact(dim, g, b, nd, s));
}
void launch_contact(int dim, const KGrid &g, const KNodes &nd, int callType, hipStream_t s) {
    BY_SHAPE(g, launch_contact(dim, g, nd, callType, s));
}
void launch_contact_normals(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd,
                            const MpmMaterial *mats, const KStats &st, int callType, hipStream_t s) {
    BY_SHAPE(g, launch_contact_normals(dim, g, p, b, nd, mats, st, callType, s));
}
void launch_vel_bcs(const KGrid &g, const KNodes &nd, const KBcs &bc, int pass, int par, hipStream_t s) {
    BY_SHAPE(g, launch_vel_bcs(g, nd, bc, pass, par, s));
}
void launch_xpic(int dim, const KGrid &g, const KPart &p, const KBins &b, const KNodes &nd, const KBcs &bc,
                 int order, bool fmpm, bool forcePass, int par, hipStream_t s) {
    BY_SHAPE(g, launch_xpic(dim, g, p, b, nd, bc, order, fmpm, forcePass, par, s));
}
void launch_particle_loads(const KGrid &g, const KLoads &L, int par, hipStream_t s) {
    BY_SHAPE(g, launch_particle_loads(g, L, par, s));
}
void launch_move(const KMove &m, hipStream_t s) { mpmk_s0::launch_move(m, s); }
void launch_gather_row(const double *src, double *dst, const int *order, long long n, hipStream_t s) {
    mpmk_s0::launch_gather_row(src, dst, order, n, s);
}
void launch_gather_irow(const int *src, int *dst, const int *order, long long n, hipStream_t s) {
    mpmk_s0::launch_gather_irow(src, dst, order, n, s);
}
void launch_scatter_row(const double *src, double *dst, const int *home, long long n, hipStream_t s) {
    mpmk_s0::launch_scatter_row(src, dst, home, n, s);
}
void launch_scatter_irow(const int *src, int *dst, const int *home, long long n, hipStream_t s) {
    mpmk_s0::launch_scatter_irow(src, dst, home, n, s);
}
void launch_iota(int *a, long long n, int base, hipStream_t s) { mpmk_s0::launch_iota(a, n, base, s); }
