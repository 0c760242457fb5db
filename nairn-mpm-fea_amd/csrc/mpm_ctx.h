// mpm_ctx.h -- the device context shared by the C ABI (mpm_ctx.hip) and the solve server
// (mpm_server.hip). Internal: not part of the public header.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "mpm_events.h"
#include "mpm_kernels.h"

// the HIP events of the kernel-group timing
struct HipEvents {
    typedef hipEvent_t Event;
    typedef hipStream_t Stream;
    static std::string status(hipError_t e) { return e == hipSuccess ? std::string() : hipGetErrorString(e); }
    static std::string create(Event *e) { return status(hipEventCreate(e)); }
    static std::string record(Event e, Stream s) { return status(hipEventRecord(e, s)); }
    static std::string elapsed(float *ms, Event a, Event b) { return status(hipEventElapsedTime(ms, a, b)); }
    static void destroy(Event e) { (void)hipEventDestroy(e); }
};

// Particle rows: PROWS double rows (the MPF_* fields, then the second position buffer) and IROWS
// int rows (the MPI_* fields, the cell key, the sorted slot -> particle order, the index at
// upload and the sorted slot -> upload index). Rows are n entries apart; the memory order of
// particles is tile-major after a sort.
constexpr int PROWS = MPF_NUM_FIELDS + 3;
constexpr int IROWS = MPI_NUM_FIELDS + 4;
constexpr int IROW_KEY = MPI_NUM_FIELDS, IROW_ORDER = MPI_NUM_FIELDS + 1, IROW_HOME = MPI_NUM_FIELDS + 2,
              IROW_SHOME = MPI_NUM_FIELDS + 3;
constexpr int64_t PARTICLE_BYTES = PROWS * sizeof(double) + IROWS * sizeof(int) + sizeof(double);
// device counters: left grid, NaN positions, lCPDI corner faults, the run stop (KBins::stop, three
// words), regression-list overflows, traction corner faults
constexpr int STATS_WORDS = 8;

struct MpmCtx {
    int dim = 3;
    MpmGridDesc gd{};
    MpmRunParams run{};
    std::vector<MpmMaterial> mats;
    KGrid kg{};
    KPart kp{};
    KBins kb{};
    KNodes kn{};
    KStats ks{};
    MpmMaterial *dmats = nullptr;
    double *pbuf = nullptr;           // particle double fields + second position buffer
    int *ibuf = nullptr;              // particle int fields
    double *cbuf = nullptr;           // lCPDI semi-side vectors of the step
    double *scr = nullptr;            // one particle row of scratch (row permutations)
    int sortInterval = 50;           // steps between sorts of the rows into tile order (0: never)
    int sinceSort = 0;
    // batch of independent instances (empty for one simulation)
    std::vector<KGrid> inst;          // host copies of the instance grids
    std::vector<int> matInstH;        // material -> instance
    KGrid *dinst = nullptr;
    int *dtables = nullptr;           // tileInst, nodeBlockInst, matInst
    int *bbuf = nullptr;              // binning arrays (bbufBytes: enough for either tile shape)
    size_t bbufBytes = 0;
    int tileCap = 0;                  // batch: entries of the tile -> instance table dtables holds
    std::vector<int> instType;        // batch: each instance's material type (0 mixed, -1 no particles)
    int *dtileType = nullptr;         // batch: tile -> its instance's type (tileCap entries)
    unsigned long long *sbuf = nullptr;
    long long ncap = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // graphs of graphSteps steps, by position parity
    int graphPhase[2] = {0, 0};      // the first step's number mod the XPIC period each was captured for
    int graphSteps = 2;              // even: a graph ends on the parity it started from
    int parity = 0;                  // which position buffer holds the current positions
    double *posA[3] = {nullptr, nullptr, nullptr}, *posB[3] = {nullptr, nullptr, nullptr};
    int nextTask = 0;                // task sequence check for mpmgpu_run_task
    int64_t steps = 0;
    double mtime = 0.;
    bool stopped = false;            // the device stopped the run (ResetElementsTask exception)
    double *lrbuf = nullptr;         // regression normals: per-field normals and separations
    int *lrlist = nullptr;           // regression normals: the step's contact nodes
    // grid velocity conditions (kbc.nb = 0: none) in one device buffer; bcTime is the step start
    // time as the device counts it (one addition of the time step per step)
    KBcs kbc{};
    char *bcbuf = nullptr;
    double bcTime = 0.;
    // external particle loads and tractions (kld.nl = 0: none) in one device buffer, built from
    // the host copies of the two lists
    KLoads kld{};
    char *ldbuf = nullptr;
    std::vector<MpmParticleLoad> loads;
    std::vector<MpmParticleTraction> tractions;
    double *tbuf = nullptr;          // two step start time slots (by position parity)
    // PeriodicXPIC schedule, kept on the host: the order of the step being enqueued (0: FLIP), the
    // next step number and time that run with run.xpicOrder, and the start time of the next step
    // to enqueue (one addition of the time step per step)
    int xpicStep = 0;
    int64_t xpicNextStep = 0;
    double xpicNextTime = 0., xpicClock = 0.;
    float lastStepMs = 0.f;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // kernel timing: events around each kernel group of a step, on this context's stream
    bool profiling = false;
    EventMarks<HipEvents, MPMGPU_NUM_KERNEL_GROUPS> ev;
    std::string err;

    int fail(const std::string &m) {
        err = m;
        return 1;
    }
};

#define HIPCHECK(ctx, call)                                                               \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return (ctx)->fail(std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

namespace mpmi {
// message for calls without a context (mpmgpu_global_error); returns 1
int fail_global(const std::string &m);
// particle buffers for n particles (kept when big enough) and the field pointers into them
int place_particles(MpmCtx *c, int64_t n);
// after the particle rows are filled (the upload-index row included): second position buffer,
// counters, bins, rows sorted into tile order; waits for the stream
int prepare_particles(MpmCtx *c);
// rows into the order of the current bins (tile-major), enqueued on the context's stream
int sort_rows(MpmCtx *c);
// current and next position rows by parity
void set_positions(MpmCtx *c);
// the kernels of one whole step on the context's stream (with profiling marks when on)
void enqueue_step(MpmCtx *c);
// input checks of mpmgpu_create (nullptr when the input is usable, else the message)
const char *check_run(const MpmGridDesc *grid, const MpmRunParams *run);
const char *check_materials(const MpmMaterial *mats, int nmat);
// geometry, tiles (shape 0), time step and run stops of one grid
void fill_grid(KGrid &g, int dim, const MpmGridDesc *grid, const MpmRunParams *run);
// the tile counts of grid g for a tile shape
void set_tiles(KGrid &g, int shape);
// tiles per field of a context (summed over a batch's instances) for a tile shape
long long tiles_total(const MpmCtx *c, int shape);
// switch the context to another tile shape: tile layout, bins and instance tables, then the
// particles binned again and their rows sorted (results do not depend on the row order)
int set_tile_shape(MpmCtx *c, int shape);
// a batch's instance material types (0: the instance mixes types, -1: no particles): the types the
// strain updates launch for and the tile -> type table of the current layout
int set_instance_types(MpmCtx *c, const std::vector<int> &types);
// ints of the binning arrays for nk cell keys and ntall (field, tile) pairs
size_t bin_words(long long nk, long long ntall);
// byte sizes of a context's grid, bin and counter arrays for its grid c->kg and nmat materials
struct GridArrays {
    size_t nodeBytes = 0, cntBytes = 0, partBytes = 0, matBytes = 0, statBytes = 0, binBytes = 0;
    size_t normBytes = 0, listBytes = 0;      // regression normals (methods 5 and 6) only
};
GridArrays grid_arrays(const MpmCtx *c, int nmat);
// the context's node, tile-partial, bin and counter pointers into arrays of those sizes
void assign_grid_arrays(MpmCtx *c, const GridArrays &a, double *nodeMem, int *cntMem, double *partMem, int *bbuf,
                        unsigned long long *sbuf, double *normMem, int *listMem);
// the bin array pointers into c->bbuf for the current tile shape
void assign_bins(MpmCtx *c);
// tile kernel grids sized for the current device's compute units
void size_tile_grids();
// the grid of one simulation (mpmgpu_create's checks and settings); nullptr or the message
const char *single_grid(MpmCtx *c, const MpmGridDesc *grid, const MpmMaterial *mats, int nmat,
                        const MpmRunParams *run);
}  // namespace mpmi
