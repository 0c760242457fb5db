// mpm_static.hip -- mpmgpu_solve_static: the explicit step on particles held in the caller's device
// memory, with every other device array the step needs (grid, tile partials, bins, counters,
// second position buffer, lCPDI corners) carved from one file-scope device array. Nothing is
// allocated, so callers that cannot allocate device memory can step a solve.
//
// The array is the device's own copy of `scratch`; its address on a device is looked up once per
// device (hipGetSymbolAddress) and only that address is handed to kernels and copies. Every call
// on a device shares the array: calls that overlap on one device race, and each call clears what
// must start at zero before it steps.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mpm_ctx.h"

static __device__ __align__(256) unsigned char scratch[MPMGPU_STATIC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratchLock;
unsigned char *g_scratch[MAX_DEVICES] = {};    // the array's device address, per device

// the address of `scratch` on the current device (looked up on the first call on each device)
int scratch_address(unsigned char **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return mpmi::fail_global(std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return mpmi::fail_global("device index beyond the static scratch table");
    std::lock_guard<std::mutex> lk(g_scratchLock);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return mpmi::fail_global(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char *>(p);
        mpmi::size_tile_grids();
    }
    *out = g_scratch[dev];
    return 0;
}

// first failed input check of the particles (the host checks of mpmgpu_upload_particles), plus
// whether the material types and particle sizes differ from particle 0's
enum { BAD_NONE = 0, BAD_MATERIAL, BAD_CONTACT_LAW, BAD_FIELD, BAD_ELEMENT, BAD_MASS, BAD_SIZE, BAD_PLASTIC };
enum { CHK_BAD = 0, CHK_MIXED_TYPE = 1, CHK_MIXED_LP = 2, CHK_WORDS = 4 };

__global__ void k_check_particles(const double *d, const int *iv, long long n, const MpmMaterial *mats, int nmat,
                                  int dim, int nfields, long long nelem, unsigned long long *chk) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto bad = [&](unsigned long long why) { atomicCAS(&chk[CHK_BAD], 0ull, why); };
    const int m = iv[MPI_MAT * n + i];
    const long long e = iv[MPI_ELEM * n + i];
    if (m < 0 || m >= nmat) { bad(BAD_MATERIAL); return; }
    const int t = mats[m].type;
    if (t == MPMGPU_MAT_COULOMB || t == MPMGPU_MAT_CONTACT) { bad(BAD_CONTACT_LAW); return; }
    if (nfields > 1 && (mats[m].field < 0 || mats[m].field >= nfields)) bad(BAD_FIELD);
    if (e < 0 || e >= nelem) bad(BAD_ELEMENT);
    if (!(d[MPF_MASS * n + i] > 0.)) bad(BAD_MASS);
    for (int a = 0; a < dim; a++) {
        const double l = d[(MPF_LP_X + a) * n + i];
        if (!(l > 0. && l <= 1.)) bad(BAD_SIZE);
    }
    if (t == MPMGPU_MAT_ISOPLASTICITY && d[MPF_HIST0 * n + i] == 0.)
        for (int f = MPF_ALT_XX; f <= MPF_ALT_XY; f++)
            if (d[f * n + i] != 0.) bad(BAD_PLASTIC);
    const int m0 = iv[MPI_MAT * n];
    if (m0 >= 0 && m0 < nmat && mats[m0].type != t) atomicOr(&chk[CHK_MIXED_TYPE], 1ull);
    for (int a = 0; a < 3; a++)
        if (d[(MPF_LP_X + a) * n + i] != d[(MPF_LP_X + a) * n]) atomicOr(&chk[CHK_MIXED_LP], 1ull);
}

const char *bad_message(unsigned long long why) {
    switch (why) {
        case BAD_MATERIAL: return "particle with an undefined material";
        case BAD_CONTACT_LAW: return "particle with a contact-law material";
        case BAD_FIELD: return "particle material without a velocity field";
        case BAD_ELEMENT: return "particle element outside the grid";
        case BAD_MASS: return "particle with a mass that is not positive";
        case BAD_SIZE: return "particle half size must be in (0, 1] cell half widths";
        case BAD_PLASTIC: return "IsoPlasticity particle with plastic strain but zero cumulative plastic strain";
        default: return "bad particle input";
    }
}

__global__ void k_zero_words(unsigned long long *a, long long n) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        a[k] = 0ull;
}

__global__ void k_copy_rows(const double *src, double *dst, long long n) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

// a run of the scratch array, 256-byte aligned
struct Carve {
    unsigned char *base;
    size_t used = 0;
    template <typename T> T *take(size_t bytes) {
        T *p = reinterpret_cast<T *>(base + used);
        used += (bytes + 255) / 256 * 256;
        return p;
    }
};

// bytes of scratch a solve takes: grid arrays, counters, check words, the second position rows,
// the bin rows (key, order, upload index, sorted upload index) and the lCPDI corner rows
size_t scratch_bytes(const mpmi::GridArrays &a, long long n, bool cpdi) {
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    size_t s = up(a.nodeBytes) + up(a.cntBytes) + up(a.partBytes) + up(a.matBytes) + up(a.statBytes) + up(a.binBytes);
    s += up(a.normBytes) + up(a.listBytes);
    s += up(CHK_WORDS * sizeof(unsigned long long));
    s += up(3 * (size_t)n * sizeof(double)) + up(4 * (size_t)n * sizeof(int));
    if (cpdi) s += up(9 * (size_t)n * sizeof(double));
    return s;
}

}  // namespace

extern "C" {

int mpmgpu_solve_static(const MpmGridDesc *grid, const MpmMaterial *mats, int32_t nmat, const MpmRunParams *run,
                        int64_t n, double *dfields, int32_t *ifields, int32_t nsteps, void *stream) {
    using mpmi::fail_global;
    if (!grid || !mats || !run) return fail_global("null grid, materials or run parameters");
    if (n < 0 || n > 0x7fffffffLL) return fail_global("bad particle count");
    if (nsteps < 0) return fail_global("bad step count");
    if (n > 0 && (!dfields || !ifields)) return fail_global("null particle rows");
    if (run->xpicOrder > 0) return fail_global("XPIC and FMPM are not available in the static-scratch solve");
    MpmCtx c;
    if (const char *m = mpmi::single_grid(&c, grid, mats, nmat, run)) return fail_global(m);
    const mpmi::GridArrays a = mpmi::grid_arrays(&c, nmat);
    const bool cpdi = run->shape == MPMGPU_SHAPE_LCPDI;
    const size_t need = scratch_bytes(a, n, cpdi);
    if (need > MPMGPU_STATIC_SCRATCH_BYTES)
        return fail_global("the solve needs " + std::to_string(need) + " bytes of static scratch, more than its " +
                           std::to_string((unsigned long long)MPMGPU_STATIC_SCRATCH_BYTES) + " bytes");
    if (n == 0 || nsteps == 0) return 0;
    unsigned char *base = nullptr;
    if (scratch_address(&base)) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    c.stream = s;
    auto hipfail = [&](const char *what, hipError_t e) {
        return fail_global(std::string(what) + ": " + hipGetErrorString(e));
    };

    // carve the arrays and clear them (the previous call's values stay in the array)
    Carve cv{base};
    double *nodeMem = cv.take<double>(a.nodeBytes);
    int *cntMem = cv.take<int>(a.cntBytes);
    double *partMem = cv.take<double>(a.partBytes);
    c.dmats = cv.take<MpmMaterial>(a.matBytes);
    unsigned long long *sbuf = cv.take<unsigned long long>(a.statBytes);
    int *bbuf = cv.take<int>(a.binBytes);
    unsigned long long *chk = cv.take<unsigned long long>(CHK_WORDS * sizeof(unsigned long long));
    double *posB = cv.take<double>(3 * (size_t)n * sizeof(double));
    int *irows = cv.take<int>(4 * (size_t)n * sizeof(int));
    double *cbuf = cpdi ? cv.take<double>(9 * (size_t)n * sizeof(double)) : nullptr;
    double *normMem = a.normBytes ? cv.take<double>(a.normBytes) : nullptr;
    int *listMem = a.listBytes ? cv.take<int>(a.listBytes) : nullptr;
    mpmi::assign_grid_arrays(&c, a, nodeMem, cntMem, partMem, bbuf, sbuf, normMem, listMem);
    c.sbuf = nullptr;
    c.bbuf = nullptr;                     // nothing here is the context's to free
    const long long words = (long long)(cv.used / sizeof(unsigned long long));
    hipLaunchKernelGGL(k_zero_words, dim3(1024), dim3(256), 0, s, reinterpret_cast<unsigned long long *>(base), words);
    hipError_t e = hipMemcpyAsync(c.dmats, mats, a.matBytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hipfail("material table copy", e);

    // the particle rows are the caller's; the second position buffer and the bin rows are scratch
    const size_t nn = (size_t)n;
    KPart &P = c.kp;
    for (int f = 0; f < MPF_NUM_FIELDS; f++) P.f[f] = dfields + f * nn;
    for (int k = 0; k < 3; k++) {
        c.posA[k] = dfields + (MPF_POS_X + k) * nn;
        c.posB[k] = posB + k * nn;
    }
    for (int k = 0; k < 9; k++) P.cpr[k] = cbuf ? cbuf + k * nn : nullptr;
    P.elem = ifields + MPI_ELEM * nn;
    P.mat = ifields + MPI_MAT * nn;
    P.id = ifields + MPI_ID * nn;
    P.left = ifields + MPI_LEFT * nn;
    P.n = n;
    c.kb.key = irows;
    c.kb.order = irows + nn;
    c.kb.home = irows + 2 * nn;
    c.kb.shome = irows + 3 * nn;
    c.parity = 0;

    // the upload's input checks, on the device
    const long long nelem = (long long)grid->nx * grid->ny * (c.dim == 3 ? grid->nz : 1);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_check_particles, dim3(blocks), dim3(256), 0, s, dfields, ifields, (long long)n, c.dmats,
                       (int)nmat, c.dim, c.kg.nfields, nelem, chk);
    unsigned long long hchk[CHK_WORDS] = {};
    double lp0[3] = {0., 0., 0.};
    int mat0 = 0;
    if ((e = hipMemcpyAsync(hchk, chk, sizeof(hchk), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipMemcpyAsync(&mat0, P.mat, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return hipfail("input check copy", e);
    for (int k = 0; k < 3; k++)
        if ((e = hipMemcpyAsync(&lp0[k], dfields + (MPF_LP_X + k) * nn, sizeof(double), hipMemcpyDeviceToHost, s)) !=
            hipSuccess)
            return hipfail("input check copy", e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hipfail("input check", e);
    if (hchk[CHK_BAD]) return fail_global(bad_message(hchk[CHK_BAD]));
    // one material type lets the strain kernels compile a single law; one particle size lets the
    // kernels skip the size rows (as mpmgpu_upload_particles decides)
    c.kg.matType = hchk[CHK_MIXED_TYPE] ? 0 : mats[mat0].type;
    c.kg.uniformLp = hchk[CHK_MIXED_LP] ? 0 : 1;
    for (int k = 0; k < 3; k++) c.kg.lp[k] = lp0[k];

    // prepare as an upload does (the rows keep the caller's order: they are not sorted into tiles,
    // which changes no result), then step
    launch_iota(irows + 2 * nn, (long long)n, 0, s);
    for (int k = 0; k < 3; k++)
        hipLaunchKernelGGL(k_copy_rows, dim3(1024), dim3(256), 0, s, c.posA[k], c.posB[k], (long long)n);
    mpmi::set_positions(&c);
    launch_bin_keys(c.dim, c.kg, c.kp, c.kb, s);
    launch_bin_finish(c.dim, c.kg, c.kp, c.kb, s);
    // the tile list counts a completed step at every binning; this one is the set-up's, not a step's
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, c.kb.stop + 2, 1ll);
    for (int32_t k = 0; k < nsteps; k++) mpmi::enqueue_step(&c);
    if ((e = hipGetLastError()) != hipSuccess) return hipfail("step launch", e);

    // a run stopped on the device keeps the stopping step's state, as mpmgpu_step does
    unsigned long long st[STATS_WORDS] = {};
    if ((e = hipMemcpyAsync(st, sbuf, sizeof(st), hipMemcpyDeviceToHost, s)) != hipSuccess) return hipfail("counter copy", e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hipfail("step", e);
    const unsigned long long why = stop_reason(st[3]);
    const int parity = why ? (int)((st[4] + 1) & 1ull) : (int)(nsteps & 1);
    if (parity == 1) {
        for (int k = 0; k < 3; k++)
            hipLaunchKernelGGL(k_copy_rows, dim3(1024), dim3(256), 0, s, c.posB[k], c.posA[k], (long long)n);
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return hipfail("position copy", e);
    }
    if (st[6])
        return fail_global("a contact node is seen by more than 1024 particles, more than the regression normals' "
                           "list holds");
    if (cpdi && st[2])
        return fail_global("A CPDI particle domain node has left the grid, or a particle domain reaches beyond the "
                           "neighbouring elements (" + std::to_string(st[2]) + " particle passes)");
    if (why) {
        const char *msg = why == STOP_TOO_MANY_LEFT ? "Too many particles have left the grid"
                          : why == STOP_NAN         ? "Particle left grid with position nan."
                          : why == STOP_DELETE      ? "A particle has left the grid and would be deleted into the reservoir"
                                                    : "the run stopped";
        return fail_global(std::string(msg) + " (step " + std::to_string(st[4] + 1) + ")");
    }
    return 0;
}

}  // extern "C"
