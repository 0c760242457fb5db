// mpm_kernels.h -- launch interface of the step kernels (mpm_kernels.hip) used by the context.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/nairnmpm_gpu.h"

// Tile shapes: cells per tile on each axis. A particle in a cell touches nodes at offsets -1..2,
// so a tile's node footprint is (T+3) nodes per axis. Shape 0 is the default; shape 1 doubles the
// tile along x (fewer footprint nodes per cell, more LDS per workgroup, half the tiles). The step
// kernels are compiled once per shape; a context runs one shape, chosen by the autotuner or set.
template <int TX_, int TY_, int TZ_> struct TileDims {
    static constexpr int TX = TX_, TY = TY_, TZ = TZ_;
    static constexpr int NC = TX * TY * TZ;
    static constexpr int FX = TX + 3, FY = TY + 3, FZ = TZ == 1 ? 1 : TZ + 3;
    static constexpr int FOOT = FX * FY * FZ;
};
template <int DIM, int S> struct TileShape;
template <> struct TileShape<3, 0> : TileDims<4, 4, 4> {};
template <> struct TileShape<3, 1> : TileDims<8, 4, 4> {};
template <> struct TileShape<2, 0> : TileDims<8, 8, 1> {};
template <> struct TileShape<2, 1> : TileDims<16, 8, 1> {};
#define MPM_TILE_SHAPES 2

// the same numbers for host code
struct TileSize {
    int tx, ty, tz, nc, foot;
};
template <class T> constexpr TileSize tile_size_of() { return TileSize{T::TX, T::TY, T::TZ, T::NC, T::FOOT}; }
inline TileSize tile_size(int dim, int shape) {
    if (dim == 3) return shape ? tile_size_of<TileShape<3, 1>>() : tile_size_of<TileShape<3, 0>>();
    return shape ? tile_size_of<TileShape<2, 1>>() : tile_size_of<TileShape<2, 0>>();
}

// kernel arguments (passed by value)
struct KGrid {
    int dim;
    int nx, ny, nz;            // cells (nz = 1 in 2D)
    int nnx, nny, nnz;         // nodes (nnz = 1 in 2D)
    int ntx, nty, ntz;         // tiles
    int ntiles;
    int tileShape;             // 0 or 1 (TileShape): which compiled instance of the kernels runs
    long long nnodes;
    double xmin, ymin, zmin, dx, dy, dz;
    double timestep, strainFirst, strainLast;
    double gridDamping, particleDamping;   // constant damping of the FLIP update (1/s)
    int damped;                // nonzero damping here (a batch: in any instance)
    double gravity[3];         // constant body force per unit mass (mm/s^2)
    int hasGravity;
    int uniformLp;             // 1 when every particle has the same lp (then lp[] is used)
    int matType;               // the one material type of all particles, 0 when mixed
    // a batch whose instances differ in material type: each tile's type (its instance's, 0 when the
    // instance mixes types, -1 when it has no particles) and the types present, so the strain
    // updates run one launch per type with that law's registers (null / 0 otherwise)
    const int *tileType;
    int ntypes, types[4], anyMixed;
    double lp[3];
    // material velocity fields (multimaterial mode when nfields > 1)
    int nfields;
    int fieldOf[MPMGPU_MAX_MATERIALS];   // material -> field
    // contact (multimaterial mode): normals method, law (CoulombFriction or stick-to-center-of-mass)
    int normals;
    int ignoreContact;         // ContactLaw type 60: always move to the center-of-mass velocity
    double friction, frictionStatic;
    double contactNormal[3];   // normals method 4 (specified)
    int shape;                 // MPMGPU_SHAPE_UGIMP or MPMGPU_SHAPE_LCPDI
    // ResetElementsTask's stops (one simulation; a batch sets all three to 0): the run stops when
    // maxLeaving particles have left the grid (0: never), at the first leaving particle when
    // deleteLeaving is set, and at the first NaN position when stopOnNan is set
    int maxLeaving, deleteLeaving, stopOnNan;
    // Several independent grids (a batch of instances) share the context's arrays: instance i
    // owns nodes [nodeBase, nodeBase+nnodes) and tiles [tileBase, tileBase+ntiles) of every field.
    // For one grid: nodeBase = tileBase = 0, nstride = nnodes, tilesTotal = ntiles, inst = null.
    long long nstride;         // nodes per field in the context (vector component stride)
    long long nodeBase;
    int tileBase;
    int tilesTotal;            // tiles per field in the context
    const struct KGrid *inst;  // device table of the instances' grids (null for one grid)
    const int *tileInst;       // context tile -> instance
    const int *nodeBlockInst;  // block of NT context nodes -> instance (ranges padded to NT)
    const int *matInst;        // material -> instance
};

struct KPart {
    double *f[MPF_NUM_FIELDS]; // f[MPF_POS_*] is the start-of-step position
    double *posn[3];           // end-of-step position (written by the particle update)
    double *cpr[9];            // lCPDI: semi-side vectors r1, r2, r3 of the step (null otherwise)
    int *elem, *mat, *id, *left;
    long long n;
};

struct KBins {
    int *key;                  // tile-major cell key per particle
    int *order;                // sorted slot -> particle
    int *cellStart;            // nkeys + 1
    int *count;                // nkeys histogram
    int *cursor;               // nkeys scatter cursor
    int *blockSums;            // scan scratch
    const int *home;           // particle -> its index at upload: the order inside a cell
    int *shome;                // sorted slot -> upload index of its particle
    int *tiles;                // the non-empty (field, tile) pairs of the binning, in no fixed order
    int *ntiles;               // length of tiles
    int *work;                 // tile kernels' list cursors: 8 per launch of a step, 32 ints apart
    // per (field, tile): a non-empty pair's footprint reaches the pair's nodes at this binning
    // (need), at this or the previous one (needAny: the pairs whose nodes the node passes visit)
    int *need;
    int *needAny;
    int *tileOn;               // per (field, tile): non-empty at this binning (k_tile_list)
    // run stop (ResetElementsTask's exceptions): stop[0] the reason and its rank (stop_run; 0 while running),
    // stop[1] the steps completed before the step that stopped, stop[2] the steps completed since
    // the upload (counted by the tile list). Once stop[0] is set, binning builds an empty tile
    // list, so every later tile kernel has nothing to do and the particle state stays as the
    // stopping step left it.
    unsigned long long *stop;
    int nkeys;
    int ntall;                 // (field, tile) pairs: nfields * tilesTotal
    int slot;                  // the launch's cursor (set by the launch functions)
};

// launches of the tile kernels in one step, each with its own cursors (cleared when the list is built)
// (the XPIC iterations share one slot, whose cursors each iteration's node pass clears again)
// (per-type strain launches of a batch take TILE_SLOT_TYPED + j: j < 4 the first half step's,
// 4 + j the second's)
enum { TILE_SLOT_MASS, TILE_SLOT_CONTACT_FIRST, TILE_SLOT_USF, TILE_SLOT_FORCE, TILE_SLOT_G2P,
       TILE_SLOT_CONTACT_LAST, TILE_SLOT_USL, TILE_SLOT_FORCE_LOADS, TILE_SLOT_XPIC, TILE_SLOT_TYPED,
       TILE_SLOTS = TILE_SLOT_TYPED + 8 };
#define TILE_SLOT_STRIDE 32

struct KNodes {
    double *mass;              // nnodes
    double *pk;                // 3*nnodes (component-major)
    double *vk;                // 3*nnodes
    double *ftot;              // 3*nnodes
    double *vprev;             // 3*nnodes: the XPIC series' latest increment (null without XPIC)
    int *cnt;                  // nnodes: particle-node pairs (numberPoints)
    double *part;              // per-(field, tile) partial sums, nfields * ntiles * NQMAX * FOOT
    // multimaterial contact extrapolations: volume, mass-weighted displacement, volume gradient
    double *cvol;              // nnodes
    double *cdisp;             // 3*nnodes
    double *cgrad;             // 3*nnodes
    // regression normals (methods 5 and 6): per field the unit normal and the separation of the
    // field's material against the rest at each contact node, 4*nnodes (component-major); and
    // the list of contact nodes of the step (clist[0] their count, the nodes from clist[1])
    double *cnorm;
    int *clist;
};
// Node arrays hold nfields consecutive copies (field f at offset f*nnodes for scalars and
// f*3*nnodes for vectors); a kernel working on one field takes a view shifted to it.

struct KStats {
    unsigned long long *leftGrid;
    unsigned long long *nanCount;
    unsigned long long *cpdiErr;   // lCPDI corners outside the grid or beyond the neighbour cells
    unsigned long long *lrOverflow;  // contact nodes seen by more particles than a regression list holds
    unsigned long long *tracErr;     // traction corners outside the grid or beyond the neighbour cells
};

enum { NQ_MASS = 4, NQ_FORCE = 3, NQ_MOM = 3, NQ_CONTACT = 7, NQMAX = 7 };

// reasons a run stops on the device
enum { STOP_NONE = 0, STOP_TOO_MANY_LEFT = 1, STOP_NAN = 2, STOP_DELETE = 3, STOP_BAD_INPUT = 4 };

// The run stops (ResetElementsTask's exceptions, ResetElementsTask.cpp:71-151). The reference raises
// the first exception in particle order; here the particles of a step run in parallel, so every
// stop carries a rank and the lowest rank wins, whatever the timing: a NaN position ranks first
// (rank -1), a leaving particle by its place among the step's new leavers (the value of the
// left-grid count it incremented), so the decision falls to the leaver whose count is lowest, as
// it does to the first leaver in particle order. stop[0] holds ~(((rank + 1) << 8) | why), nonzero
// once stopped; stop[1] the steps completed before this one.
__device__ __forceinline__ void stop_run(const KBins &B, unsigned long long why, long long rank) {
    const unsigned long long key = ((unsigned long long)(rank + 1) << 8) | why;
    atomicMax(&B.stop[0], ~key);
    atomicExch(&B.stop[1], B.stop[2]);
}
// the reason of a stop word (0 while running)
__host__ __device__ __forceinline__ unsigned long long stop_reason(unsigned long long word) {
    return word ? (~word) & 0xffull : 0ull;
}


// contact call types (UpdateMomentaTask::ContactAndMomentaBCs passType)
enum { CONTACT_MASS_MOMENTUM = 0, CONTACT_UPDATE_MOMENTUM = 1, CONTACT_UPDATE_STRAINS_LAST = 2 };


// Grid velocity conditions of one simulation (single grid), grouped by node: node b's conditions
// are entries start[b]..start[b+1]-1, in the order of the condition list. One thread per
// (node, field) applies them, so no two threads touch the same node values.
struct KBcs {
    int nb;                    // nodes with conditions
    const long long *node;     // nb node indices
    const int *start;          // nb + 1
    const int *dir;            // per entry: axis 0..2
    const int *style;          // 1 constant, 2 linear, 3 sine, 4 cosine
    const double *value;       // per entry (internal units)
    const double *ftime;
    double *cur;               // per entry: value of the step (written by the mass-momentum pass)
    int *act;                  // per entry: acts in this step
    double *save;              // nb * nfields * 3: momenta before the mass-momentum pass
    double *time;              // two slots: slot par is the step's start time, slot par ^ 1 the next's
};
// passes of NodalVelBC::GridVelocityConditions; GRID_FORCES runs the force pass and the momentum
// update of the nodes with conditions, FORCES_MOMENTA adds the UPDATE_MOMENTUM pass in the same
// launch (single-material mode, where no contact pass comes between)
// The XPIC passes zero the held components of the series increment vprev (and take them back
// out of vk); XPIC_FORCE, the particle-update pass of XPIC, also moves the lumped reaction into ftot.
enum { BC_PASS_MASS_MOMENTUM = 0, BC_PASS_GRID_FORCES, BC_PASS_UPDATE_MOMENTUM, BC_PASS_FORCES_MOMENTA,
       BC_PASS_UPDATE_STRAINS_LAST, BC_PASS_XPIC, BC_PASS_XPIC_FORCE };

// External particle loads of one simulation: map[upload index] is -1 or the loaded particle's
// slot q, whose loads are entries start[q]..start[q+1]-1 in list order; fext holds the step's
// pFext of each slot (3 doubles), read by the force scatter through the particle's upload index.
// Particles with external loads or tractions ("loaded" particles, nl of them): map takes a
// particle's upload index to its loaded index q (-1: neither). Loads: entries [start[q],
// start[q+1]) give pFext (fext[3q..3q+2], zero without loads). Tractions: tslot[q] is the
// particle's traction table (-1: none); table t holds entries [tstart[t], tstart[t+1]) and, each
// step, the nodal traction forces on the particle's 4^DIM candidate nodes (offsets -1..2 from its
// cell), component-major: ttab[(3t + c) * TRAC_NODES + o], o = ((oz*4 + oy)*4 + ox) (2D: oy*4 + ox).
#define TRAC_NODES 64
struct KLoads {
    int nl;                    // loaded particles (0: no loads and no tractions)
    const int *map;
    const int *start;
    const int *dir, *style;
    const double *value, *ftime;
    double *fext;
    double *time;              // the two step start time slots shared with KBcs
    int *tflag[2];             // per (field, tile): holds a loaded particle this step (slot par)
    int nt;                    // traction tables (0: no tractions)
    const int *tslot;
    const int *tstart;
    const int *tdir, *tface, *tstyle;
    const double *tvalue, *tftime, *tthick;
    double *ttab;
    unsigned long long *terr;  // traction corners outside the grid or beyond the neighbour cells
};

// A run of particles copied between the particle buffers of two contexts (the solve server moves
// a resident solve into the next batch layout). Double rows are read from sd (row stride sn)
// except the positions, which come from the source's current position rows spos; int rows MPI_*
// likewise from si; material numbers shift by matShift.
struct KMove {
    const double *sd;
    const double *spos[3];
    const int *si;
    long long sn, soff;
    double *dd;
    int *di;
    long long dn, doff;
    long long cnt;
    int matShift;
    int homeShift;             // added to the upload-index row (int row MPI_NUM_FIELDS + 2)
};

// Launchers (mpm_launch.inc), all on the context's stream. The step kernels are compiled once per
// tile shape, into namespaces mpmk_s0 and mpmk_s1; the launchers outside them run the instance of
// g.tileShape (those without a grid argument do not depend on the shape).
namespace mpmk_s0 {
#include "mpm_launch.inc"
}
namespace mpmk_s1 {
#include "mpm_launch.inc"
}
#include "mpm_launch.inc"
