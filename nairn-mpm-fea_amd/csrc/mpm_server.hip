// mpm_server.hip -- the solve server: continuous batching of independent solves on one device.
//
// Requests wait in arrival order until their device footprint fits the memory budget. Admitted
// solves are resident: their particle rows live in one batch context (mpmgpu_batch_create layout)
// and every batch step advances each of them by one of its own steps. When the membership
// changes (a solve finishes, or queued ones fit), a new batch context is built for the new set;
// the survivors' particle rows move into it on the device (k_move), and the old context is
// released. Between layouts the steps replay from captured graphs. Uploads and downloads stay off
// the request loop where memory allows: a submitted solve's rows are copied to a device staging
// buffer by the submitting thread, and a finished solve's rows are copied to a device result
// buffer that mpmgpu_server_collect brings to the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mpm_ctx.h"

namespace {

struct Request {
    int64_t id = 0;
    int state = MPMGPU_REQ_QUEUED;
    MpmGridDesc grid{};
    MpmRunParams run{};
    std::vector<MpmMaterial> mats;
    int64_t n = 0;
    std::vector<double> d;          // initial rows until admitted; final rows once done
    std::vector<int32_t> iv;
    int32_t steps = 0;
    int32_t done = 0;
    int64_t footprint = 0;
    int64_t particleBytes = 0;
    int matType = 0;                // the one material type of its particles, 0 when mixed
    int uniformLp = 1;
    double lp[3] = {0., 0., 0.};
    double tSubmit = 0., tAdmit = -1., tFinish = -1.;
    // place in the current layout (request-loop thread only)
    bool placed = false;
    int64_t off = 0;
    int firstMat = 0;
    // device rows: the staged upload (before admission) or the result (after it finishes); int
    // rows laid out as a context's (MPI_* rows, then the upload-index row at IROW_HOME)
    double *dd = nullptr;
    int *di = nullptr;
    int64_t devBytes = 0;
    int64_t homeBase = 0;          // result rows: upload index of the solve's first particle
    int resultFirstMat = 0;        // result rows: batch number of the solve's first material
};

int64_t row_bytes(int64_t n) { return n * (int64_t)(MPF_NUM_FIELDS * sizeof(double) + IROWS * sizeof(int)); }

void free_rows(Request *r) {
    if (r->dd) (void)hipFree(r->dd);
    if (r->di) (void)hipFree(r->di);
    r->dd = nullptr;
    r->di = nullptr;
}

}  // namespace

struct MpmServer {
    MpmServerParams p{};
    int device = 0;
    std::chrono::steady_clock::time_point t0;
    std::mutex m;                              // guards everything below except ctx and placement
    std::condition_variable cvWork, cvDone;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request *> queue;
    std::vector<Request *> resident;           // admitted, in layout order; newcomers at the end
    bool dirty = false;
    int64_t nextId = 1;
    int64_t layouts = 0;
    int64_t residentBytes = 0;
    double layoutSeconds = 0., stepSeconds = 0.;   // request-loop time in layouts and in batch steps
    int64_t deviceRowBytes = 0;                  // staged uploads and results held on the device
    hipStream_t upStream = nullptr;              // staging copies of submitting threads
    std::string err;
    bool failed = false;
    std::thread loop;
    bool running = false, stopFlag = false;
    MpmCtx *ctx = nullptr;                     // current layout (request-loop thread only)

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

thread_local std::string g_srvErr;

int create_fail(const std::string &m) {
    g_srvErr = m;
    return 1;
}

int srv_fail(MpmServer *s, const std::string &msg) {
    std::lock_guard<std::mutex> lk(s->m);
    s->err = msg;
    s->failed = true;
    s->cvDone.notify_all();
    return 1;
}

int submit_fail(MpmServer *s, const std::string &msg) {
    std::lock_guard<std::mutex> lk(s->m);
    s->err = msg;
    return 1;
}

// A new batch context for `members` (survivors first, in their old order, then newcomers).
int relayout(MpmServer *s, const std::vector<Request *> &members) {
    MpmCtx *old = s->ctx;
    if (members.empty()) {
        if (old) mpmgpu_destroy(old);
        s->ctx = nullptr;
        return 0;
    }
    const size_t ni = members.size();
    std::vector<MpmInstanceDesc> descs(ni);
    std::vector<MpmMaterial> mats;
    std::vector<int64_t> offs(ni);
    std::vector<int> firsts(ni);
    int64_t N = 0;
    for (size_t k = 0; k < ni; k++) {
        const Request *r = members[k];
        descs[k].grid = r->grid;
        descs[k].run = r->run;
        descs[k].firstMaterial = (int32_t)mats.size();
        descs[k].numMaterials = (int32_t)r->mats.size();
        offs[k] = N;
        firsts[k] = (int)mats.size();
        mats.insert(mats.end(), r->mats.begin(), r->mats.end());
        N += r->n;
    }
    MpmCtx *nc = nullptr;
    if (mpmgpu_batch_create((int32_t)ni, descs.data(), mats.data(), (int32_t)mats.size(), &nc))
        return srv_fail(s, std::string("batch layout: ") + mpmgpu_global_error());
    auto bail = [&](const char *what) {
        std::string e = std::string(what) + ": " + nc->err;
        mpmgpu_destroy(nc);
        return srv_fail(s, e);
    };
    if (mpmi::place_particles(nc, N)) return bail("particle buffers");
    // instance grids: the particle size test and material type measured at submission
    int common = members[0]->matType;
    for (const Request *r : members)
        if (r->matType != common) common = 0;
    nc->kg.matType = common;
    for (size_t k = 0; k < ni; k++) {
        nc->inst[k].uniformLp = members[k]->uniformLp;
        for (int a = 0; a < 3; a++) nc->inst[k].lp[a] = members[k]->lp[a];
        nc->inst[k].matType = common;
    }
    if (hipMemcpyAsync(nc->dinst, nc->inst.data(), ni * sizeof(KGrid), hipMemcpyHostToDevice, nc->stream) !=
        hipSuccess)
        return bail("instance table");
    {
        std::vector<int> types(ni);
        for (size_t k = 0; k < ni; k++) types[k] = members[k]->matType;
        if (mpmi::set_instance_types(nc, types)) return bail("instance types");
    }
    if (old) mpmi::set_positions(old);
    std::vector<std::vector<int32_t>> staged;   // newcomers' int rows with batch material numbers
    staged.reserve(ni);
    for (size_t k = 0; k < ni; k++) {
        Request *r = members[k];
        if (r->placed) {
            KMove mv{};
            mv.sd = old->pbuf;
            for (int a = 0; a < 3; a++) mv.spos[a] = old->kp.f[MPF_POS_X + a];
            mv.si = old->ibuf;
            mv.sn = old->kp.n;
            mv.soff = r->off;
            mv.dd = nc->pbuf;
            mv.di = nc->ibuf;
            mv.dn = N;
            mv.doff = offs[k];
            mv.cnt = r->n;
            mv.matShift = firsts[k] - r->firstMat;
            mv.homeShift = (int)(offs[k] - r->off);
            launch_move(mv, nc->stream);
        } else if (r->dd) {
            // staged on the device at submission: rows in the solve's own order and numbering
            KMove mv{};
            mv.sd = r->dd;
            for (int a = 0; a < 3; a++) mv.spos[a] = r->dd + (size_t)(MPF_POS_X + a) * r->n;
            mv.si = r->di;
            mv.sn = r->n;
            mv.soff = 0;
            mv.dd = nc->pbuf;
            mv.di = nc->ibuf;
            mv.dn = N;
            mv.doff = offs[k];
            mv.cnt = r->n;
            mv.matShift = firsts[k];
            mv.homeShift = (int)offs[k];
            launch_move(mv, nc->stream);
        } else {
            const size_t n = (size_t)r->n;
            for (int f = 0; f < MPF_NUM_FIELDS; f++)
                if (hipMemcpyAsync(nc->pbuf + f * N + offs[k], r->d.data() + f * n, n * sizeof(double),
                                   hipMemcpyHostToDevice, nc->stream) != hipSuccess)
                    return bail("particle upload");
            staged.emplace_back(r->iv);
            std::vector<int32_t> &iv = staged.back();
            for (size_t i = 0; i < n; i++) iv[MPI_MAT * n + i] += firsts[k];
            for (int q = 0; q < MPI_NUM_FIELDS; q++)
                if (hipMemcpyAsync(nc->ibuf + q * N + offs[k], iv.data() + q * n, n * sizeof(int32_t),
                                   hipMemcpyHostToDevice, nc->stream) != hipSuccess)
                    return bail("particle upload");
            launch_iota(nc->ibuf + IROW_HOME * N + offs[k], (long long)n, (int)offs[k], nc->stream);
        }
    }
    if (hipGetLastError() != hipSuccess) return bail("particle move");
    if (mpmi::prepare_particles(nc)) return bail("layout");     // waits for the copies
    if (mpmgpu_set_graph_steps(nc, s->p.graphSteps)) return bail("graph length");
    if (old) mpmgpu_destroy(old);
    s->ctx = nc;
    int64_t freed = 0;
    for (size_t k = 0; k < ni; k++) {
        Request *r = members[k];
        if (!r->placed) {
            std::vector<double>().swap(r->d);       // the layout holds the state now
            std::vector<int32_t>().swap(r->iv);
            if (r->dd) {
                free_rows(r);
                freed += r->devBytes;
                r->devBytes = 0;
            }
        }
        r->placed = true;
        r->off = offs[k];
        r->firstMat = firsts[k];
    }
    std::lock_guard<std::mutex> lk(s->m);
    s->deviceRowBytes -= freed;
    s->layouts++;
    return 0;
}

// Final rows of a finished solve, copied on the device to its result buffer (in layout order;
// mpmgpu_server_collect puts them in the solve's own order on the host).
int keep_result(MpmServer *s, Request *r) {
    MpmCtx *c = s->ctx;
    mpmi::set_positions(c);
    const size_t n = (size_t)r->n, N = (size_t)c->kp.n;
    if (hipMalloc(&r->dd, (size_t)MPF_NUM_FIELDS * n * sizeof(double)) != hipSuccess ||
        hipMalloc(&r->di, (size_t)IROWS * n * sizeof(int)) != hipSuccess) {
        free_rows(r);
        return srv_fail(s, "out of device memory keeping a result");
    }
    for (int f = 0; f < MPF_NUM_FIELDS; f++)
        if (hipMemcpyAsync(r->dd + f * n, c->kp.f[f] + r->off, n * sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream) != hipSuccess)
            return srv_fail(s, "result copy failed");
    for (int q = 0; q < IROWS; q++) {
        if (q != IROW_HOME && q >= MPI_NUM_FIELDS) continue;
        if (hipMemcpyAsync(r->di + q * n, c->ibuf + q * N + r->off, n * sizeof(int), hipMemcpyDeviceToDevice,
                           c->stream) != hipSuccess)
            return srv_fail(s, "result copy failed");
    }
    r->devBytes = row_bytes(r->n);
    r->homeBase = r->off;
    r->resultFirstMat = r->firstMat;
    return 0;
}

int advance(MpmServer *s, int32_t maxSteps, int32_t *completed) {
    if (completed) *completed = 0;
    std::vector<Request *> members;
    bool relay;
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (s->failed) return 1;
        const size_t before = s->resident.size();
        s->resident.erase(std::remove_if(s->resident.begin(), s->resident.end(),
                                         [](const Request *r) { return r->state != MPMGPU_REQ_RESIDENT; }),
                          s->resident.end());
        if (s->resident.size() != before) s->dirty = true;
        int64_t bytes = 0, moving = 0;
        for (const Request *r : s->resident) {
            bytes += r->footprint;
            moving += r->particleBytes;        // survivors exist twice while they move
        }
        while (!s->queue.empty()) {
            Request *r = s->queue.front();
            if (s->p.maxResident > 0 && (int)s->resident.size() >= s->p.maxResident) break;
            // device bytes held for other solves' staged uploads and kept results count too; with
            // nothing resident the head is admitted anyway, so staged rows never block progress
            const int64_t held = s->deviceRowBytes - r->devBytes;
            if (!s->resident.empty() && bytes + r->footprint + moving + held > s->p.memoryBudget) break;
            s->queue.pop_front();
            r->state = MPMGPU_REQ_RESIDENT;
            r->tAdmit = s->now();
            s->resident.push_back(r);
            bytes += r->footprint;
            s->dirty = true;
        }
        s->residentBytes = bytes;
        members = s->resident;
        relay = s->dirty;
        s->dirty = false;
    }
    if (relay) {
        const double t0 = s->now();
        if (relayout(s, members)) return 1;
        const double t1 = s->now();
        std::lock_guard<std::mutex> lk(s->m);
        s->layoutSeconds += t1 - t0;
    }
    if (members.empty()) return 0;
    int k = maxSteps;
    for (const Request *r : members) k = std::min(k, r->steps - r->done);
    const double t2 = s->now();
    if (k > 0 && (mpmgpu_step(s->ctx, k) || mpmgpu_sync(s->ctx)))
        return srv_fail(s, std::string("batch step: ") + s->ctx->err);
    const double t3 = s->now();
    std::vector<Request *> fin;
    for (Request *r : members)
        if (r->done + k >= r->steps) fin.push_back(r);
    int64_t kept = 0;
    for (Request *r : fin) {
        if (keep_result(s, r)) return 1;
        kept += r->devBytes;
    }
    if (!fin.empty() && mpmgpu_sync(s->ctx)) return srv_fail(s, std::string("result copy: ") + s->ctx->err);
    {
        std::lock_guard<std::mutex> lk(s->m);
        const double t = s->now();
        s->stepSeconds += t3 - t2;
        s->deviceRowBytes += kept;
        for (Request *r : members) r->done += k;
        for (Request *r : fin) {
            r->state = MPMGPU_REQ_DONE;
            r->tFinish = t;
        }
    }
    if (!fin.empty()) s->cvDone.notify_all();
    if (completed) *completed = (int32_t)fin.size();
    return 0;
}

void loop_main(MpmServer *s) {
    (void)hipSetDevice(s->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->m);
            s->cvWork.wait(lk, [&] { return s->stopFlag || !s->queue.empty() || !s->resident.empty(); });
            if (s->stopFlag || s->failed) break;
        }
        int32_t done = 0;
        if (advance(s, s->p.graphSteps, &done)) break;
    }
}

}  // namespace

extern "C" {

int mpmgpu_server_create(const MpmServerParams *params, MpmServer **out) {
    *out = nullptr;
    MpmServerParams p = *params;
    if (p.analysis != MPMGPU_PLANE_STRAIN && p.analysis != MPMGPU_THREED)
        return create_fail("server analysis type must be 10 (plane strain) or 12 (3D)");
    if (p.shape != MPMGPU_SHAPE_UGIMP && p.shape != MPMGPU_SHAPE_LCPDI && p.shape != MPMGPU_SHAPE_LINEAR)
        return create_fail("server shape functions must be classic (0), uGIMP (1) or lCPDI (10)");
    if (p.graphSteps == 0) p.graphSteps = 8;
    if (p.graphSteps < 2 || p.graphSteps > 64 || (p.graphSteps & 1))
        return create_fail("graph steps must be even, 2 to 64");
    if (p.maxResident < 0) return create_fail("maxResident must be >= 0");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return create_fail("no HIP device");
    if (p.memoryBudget <= 0) {
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) return create_fail("hipMemGetInfo failed");
        p.memoryBudget = (int64_t)(0.9 * (double)freeB);
    }
    MpmServer *s = new MpmServer();
    if (hipStreamCreateWithFlags(&s->upStream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return create_fail("hipStreamCreate failed");
    }
    s->p = p;
    s->device = dev;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return 0;
}

int mpmgpu_server_submit(MpmServer *s, const MpmGridDesc *grid, const MpmMaterial *mats, int32_t nmat,
                         const MpmRunParams *run, int64_t n, const double *dfields, const int32_t *ifields,
                         int32_t steps, int64_t *id) {
    if (grid->analysis != s->p.analysis || run->shape != s->p.shape)
        return submit_fail(s, "request analysis type or shape functions differ from the server's");
    if (run->multiMaterial) return submit_fail(s, "multimaterial requests cannot be served");
    if (run->xpicOrder > 0) return submit_fail(s, "XPIC and FMPM requests cannot be served");
    if (nmat < 1 || nmat > MPMGPU_MAX_MATERIALS) return submit_fail(s, "bad material count");
    if (steps < 1) return submit_fail(s, "a request needs at least one step");
    if (n < 1 || n > 0x7fffffffLL) return submit_fail(s, "bad particle count");
    const int64_t fp = mpmgpu_footprint(grid, run, n);
    if (fp < 0) return submit_fail(s, "unsupported grid or run parameters");
    for (int i = 0; i < nmat; i++) {
        int t = mats[i].type;
        if (t != MPMGPU_MAT_ISOTROPIC && t != MPMGPU_MAT_ISOPLASTICITY && t != MPMGPU_MAT_NEOHOOKEAN &&
            t != MPMGPU_MAT_MOONEY)
            return submit_fail(s, "a request material type is not served");
    }
    const size_t nn = (size_t)n;
    const long long nelem = (long long)grid->nx * grid->ny * (grid->analysis == MPMGPU_THREED ? grid->nz : 1);
    auto r = std::make_unique<Request>();
    for (size_t i = 0; i < nn; i++) {
        int m = ifields[MPI_MAT * nn + i];
        long long e = ifields[MPI_ELEM * nn + i];
        if (m < 0 || m >= nmat) return submit_fail(s, "particle with an undefined material");
        if (e < 0 || e >= nelem) return submit_fail(s, "particle element outside the grid");
        if (!(dfields[MPF_MASS * nn + i] > 0.)) return submit_fail(s, "particle with a mass that is not positive");
        for (int a = 0; a < (grid->analysis == MPMGPU_THREED ? 3 : 2); a++) {
            const double l = dfields[(MPF_LP_X + a) * nn + i];
            if (!(l > 0. && l <= 1.)) return submit_fail(s, "particle half size must be in (0, 1] cell half widths");
        }
        if (mats[m].type == MPMGPU_MAT_ISOPLASTICITY && dfields[MPF_HIST0 * nn + i] == 0.)
            for (int f = MPF_ALT_XX; f <= MPF_ALT_XY; f++)
                if (dfields[f * nn + i] != 0.)
                    return submit_fail(s, "IsoPlasticity particle with plastic strain but zero cumulative plastic strain");
    }
    r->matType = mats[ifields[MPI_MAT * nn]].type;
    for (size_t i = 1; i < nn && r->matType; i++)
        if (mats[ifields[MPI_MAT * nn + i]].type != r->matType) r->matType = 0;
    for (int a = 0; a < 3; a++) r->lp[a] = dfields[(MPF_LP_X + a) * nn];
    for (size_t i = 0; i < nn && r->uniformLp; i++)
        for (int a = 0; a < 3; a++)
            if (dfields[(MPF_LP_X + a) * nn + i] != r->lp[a]) { r->uniformLp = 0; break; }
    r->grid = *grid;
    r->run = *run;
    r->mats.assign(mats, mats + nmat);
    r->n = n;
    r->steps = steps;
    r->footprint = fp;
    r->particleBytes = n * PARTICLE_BYTES;
    if (fp > s->p.memoryBudget) return submit_fail(s, "request footprint exceeds the server's memory budget");
    // stage the rows on the device now if the budget has room beside what is resident
    const int64_t sb = row_bytes(n);
    bool stage = false;
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (s->failed) return 1;
        if (s->residentBytes + s->deviceRowBytes + sb + fp <= s->p.memoryBudget) {
            s->deviceRowBytes += sb;
            stage = true;
        }
    }
    if (stage) {
        bool ok = hipMalloc(&r->dd, (size_t)MPF_NUM_FIELDS * nn * sizeof(double)) == hipSuccess &&
                  hipMalloc(&r->di, (size_t)IROWS * nn * sizeof(int)) == hipSuccess;
        ok = ok && hipMemcpyAsync(r->dd, dfields, (size_t)MPF_NUM_FIELDS * nn * sizeof(double), hipMemcpyHostToDevice,
                                  s->upStream) == hipSuccess;
        ok = ok && hipMemcpyAsync(r->di, ifields, (size_t)MPI_NUM_FIELDS * nn * sizeof(int), hipMemcpyHostToDevice,
                                  s->upStream) == hipSuccess;
        if (ok) launch_iota(r->di + IROW_HOME * nn, n, 0, s->upStream);
        ok = ok && hipStreamSynchronize(s->upStream) == hipSuccess;
        if (ok) {
            r->devBytes = sb;
        } else {                               // fall back to a host copy that the loop uploads
            free_rows(r.get());
            std::lock_guard<std::mutex> lk(s->m);
            s->deviceRowBytes -= sb;
            stage = false;
        }
    }
    if (!stage) {
        r->d.assign(dfields, dfields + (size_t)MPF_NUM_FIELDS * nn);
        r->iv.assign(ifields, ifields + (size_t)MPI_NUM_FIELDS * nn);
    }
    std::lock_guard<std::mutex> lk(s->m);
    if (s->failed) return 1;
    r->id = s->nextId++;
    r->tSubmit = s->now();
    *id = r->id;
    s->queue.push_back(r.get());
    s->reqs[r->id] = std::move(r);
    s->cvWork.notify_one();
    return 0;
}

int mpmgpu_server_advance(MpmServer *s, int32_t maxSteps, int32_t *completed) {
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (s->running) {
            s->err = "the request loop runs on the server's thread";
            return 1;
        }
    }
    if (maxSteps < 1) return submit_fail(s, "maxSteps must be >= 1");
    return advance(s, maxSteps, completed);
}

int mpmgpu_server_start(MpmServer *s) {
    std::lock_guard<std::mutex> lk(s->m);
    if (s->running) return 0;
    s->stopFlag = false;
    s->running = true;
    s->loop = std::thread(loop_main, s);
    return 0;
}

int mpmgpu_server_stop(MpmServer *s) {
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (!s->running) return 0;
        s->stopFlag = true;
        s->cvWork.notify_all();
    }
    s->loop.join();
    std::lock_guard<std::mutex> lk(s->m);
    s->running = false;
    return s->failed ? 1 : 0;
}

int mpmgpu_server_wait(MpmServer *s, int64_t id, double timeoutSec, int32_t *done) {
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) {
        s->err = "unknown request id";
        return 1;
    }
    Request *r = it->second.get();
    auto pred = [&] { return s->failed || r->state >= MPMGPU_REQ_DONE; };
    if (timeoutSec < 0) s->cvDone.wait(lk, pred);
    else s->cvDone.wait_for(lk, std::chrono::duration<double>(timeoutSec), pred);
    if (done) *done = r->state >= MPMGPU_REQ_DONE ? 1 : 0;
    return s->failed ? 1 : 0;
}

int mpmgpu_server_info(MpmServer *s, int64_t id, MpmRequestInfo *info) {
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) {
        s->err = "unknown request id";
        return 1;
    }
    const Request *r = it->second.get();
    info->state = r->state;
    info->stepsTotal = r->steps;
    info->stepsDone = r->done;
    info->pad = 0;
    info->nparticles = r->n;
    info->footprint = r->footprint;
    info->submitted = r->tSubmit;
    info->admitted = r->tAdmit;
    info->finished = r->tFinish;
    return 0;
}

int mpmgpu_server_collect(MpmServer *s, int64_t id, double *dfields, int32_t *ifields) {
    Request *r;
    {
        std::lock_guard<std::mutex> lk(s->m);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) {
            s->err = "unknown request id";
            return 1;
        }
        r = it->second.get();
        if (r->state != MPMGPU_REQ_DONE) {
            s->err = r->state == MPMGPU_REQ_COLLECTED ? "request already collected" : "request not done";
            return 1;
        }
        r->state = MPMGPU_REQ_COLLECTED;       // the result rows are this thread's now
    }
    const size_t n = (size_t)r->n;
    std::vector<double> d((size_t)MPF_NUM_FIELDS * n);
    std::vector<int32_t> iv((size_t)IROWS * n);
    bool ok = hipMemcpy(d.data(), r->dd, d.size() * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
              hipMemcpy(iv.data(), r->di, iv.size() * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    free_rows(r);
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->deviceRowBytes -= r->devBytes;
        r->devBytes = 0;
        if (!ok) {
            s->err = "result download failed";
            return 1;
        }
    }
    // back to the solve's own order (upload-index row) and material numbering
    const int32_t *home = iv.data() + (size_t)IROW_HOME * n;
    for (size_t i = 0; i < n; i++) {
        const long long h = (long long)home[i] - r->homeBase;
        if (h < 0 || h >= (long long)n) {
            std::lock_guard<std::mutex> lk(s->m);
            s->err = "particle rows of a solve left its range";
            return 1;
        }
        for (int f = 0; f < MPF_NUM_FIELDS; f++) dfields[f * n + h] = d[f * n + i];
        if (ifields)
            for (int q = 0; q < MPI_NUM_FIELDS; q++)
                ifields[q * n + h] = iv[q * n + i] - (q == MPI_MAT ? r->resultFirstMat : 0);
    }
    return 0;
}

int mpmgpu_server_counts(MpmServer *s, int32_t *queued, int32_t *resident, int64_t *residentBytes,
                         int64_t *layouts, double *layoutSeconds, double *stepSeconds) {
    std::lock_guard<std::mutex> lk(s->m);
    if (queued) *queued = (int32_t)s->queue.size();
    if (resident) {
        int k = 0;
        for (const Request *r : s->resident) k += r->state == MPMGPU_REQ_RESIDENT;
        *resident = k;
    }
    if (residentBytes) *residentBytes = s->residentBytes;
    if (layouts) *layouts = s->layouts;
    if (layoutSeconds) *layoutSeconds = s->layoutSeconds;
    if (stepSeconds) *stepSeconds = s->stepSeconds;
    return 0;
}

double mpmgpu_server_clock(const MpmServer *s) { return s->now(); }

const char *mpmgpu_server_error(const MpmServer *s) { return s ? s->err.c_str() : g_srvErr.c_str(); }

void mpmgpu_server_destroy(MpmServer *s) {
    if (!s) return;
    mpmgpu_server_stop(s);
    if (s->ctx) mpmgpu_destroy(s->ctx);
    for (auto &kv : s->reqs) free_rows(kv.second.get());
    if (s->upStream) (void)hipStreamDestroy(s->upStream);
    delete s;
}

}  // extern "C"
