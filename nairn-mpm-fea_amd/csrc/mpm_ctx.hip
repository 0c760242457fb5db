// mpm_ctx.hip -- the C ABI: problem building (input boundary) and the device context that runs
// the step. The context owns every device buffer of one simulation; all work goes on its own
// stream (default priority), and whole steps are replayed from a captured HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mpm_ctx.h"
#include "mpm_problem.h"

static thread_local std::string g_err;

int mpmi::fail_global(const std::string &m) {
    g_err = m;
    return 1;
}
using mpmi::fail_global;

void mpmi::set_positions(MpmCtx *c) {
    double **cur = c->parity == 0 ? c->posA : c->posB;
    double **nxt = c->parity == 0 ? c->posB : c->posA;
    for (int a = 0; a < 3; a++) {
        c->kp.f[MPF_POS_X + a] = cur[a];
        c->kp.posn[a] = nxt[a];
    }
}
using mpmi::set_positions;
using mpmi::check_run;
using mpmi::check_materials;
using mpmi::fill_grid;
using mpmi::set_tiles;
using mpmi::bin_words;

static void destroy_graphs(MpmCtx *c) {
    for (int i = 0; i < 2; i++)
        if (c->graph[i]) { (void)hipGraphExecDestroy(c->graph[i]); c->graph[i] = nullptr; }
}

// the device's copy of the step start time of the velocity conditions and loads, in the slot of
// the current position parity (none allocated: nothing to do)
static int put_bc_time(MpmCtx *c) {
    if (!c->tbuf) return 0;
    HIPCHECK(c, hipMemcpyAsync(c->tbuf + c->parity, &c->bcTime, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int ensure_time_slots(MpmCtx *c) {
    if (c->tbuf) return 0;
    HIPCHECK(c, hipMalloc(&c->tbuf, 2 * sizeof(double)));
    return 0;
}

static int build_particle_bcs(MpmCtx *c);

// every (field, tile) pair's nodes are visited by the next step's node passes (after an upload or a
// restored state, when nodes may hold values of particles that are no longer there)
static hipError_t mark_all_needed(MpmCtx *c) {
    return hipMemsetAsync(c->kb.needAny, 1, (size_t)c->kb.ntall * sizeof(int), c->stream);
}

// PeriodicXPIC::PrepareForStep and StepCalculation (PeriodicXPIC.cpp:283-312,364-373): the order
// of step number st (1 for the first step). The reference decides a step's order before the first
// step (step 0, time 0) and at the end of the step before it, so the step tested is st and the
// time tested is dt for the first step and the step's start time after that; a step that gets
// the order moves the next one on by the period (steps), or to the tested time plus the period.
static void xpic_schedule(MpmCtx *c, int64_t st) {
    const MpmRunParams &r = c->run;
    c->xpicStep = 0;
    if (r.xpicOrder <= 0) return;
    if (st == 1) {
        c->xpicNextStep = r.xpicPeriodSteps;
        c->xpicNextTime = r.xpicPeriodTime;
    }
    if (r.xpicPeriodSteps > 0) {
        if (st >= c->xpicNextStep) {
            c->xpicStep = r.xpicOrder;
            c->xpicNextStep += r.xpicPeriodSteps;
        }
    } else if (r.xpicPeriodTime > 0.) {
        const double t = st == 1 ? 0. + r.timestep : c->xpicClock;
        if (t >= c->xpicNextTime) {
            c->xpicStep = r.xpicOrder;
            c->xpicNextTime = t + r.xpicPeriodTime;
        }
    }
}

// Graphs hold whole steps with their XPIC orders fixed at capture. A schedule by step count
// selects exactly the steps s with s % P == 0 (the next step starts at P and moves on by P), so a
// graph of G steps, G a multiple of P, has the same orders at every launch that starts at the same
// s % P: a graph is kept with that phase and captured again when a call starts at another one. A
// schedule by time runs step by step.
static bool xpic_graphable(const MpmCtx *c) {
    const MpmRunParams &r = c->run;
    if (r.xpicOrder <= 0 || r.xpicPeriodSteps == 1) return true;
    return r.xpicPeriodSteps > 1 && c->graphSteps % r.xpicPeriodSteps == 0;
}
static int xpic_phase(const MpmCtx *c, int64_t st) {
    const int P = c->run.xpicPeriodSteps;
    return c->run.xpicOrder > 0 && P > 1 ? (int)(st % P) : 0;
}
// the order of step st under a schedule that xpic_graphable accepts
static int xpic_order_at(const MpmCtx *c, int64_t st) {
    const int P = c->run.xpicPeriodSteps;
    if (c->run.xpicOrder <= 0) return 0;
    return P == 1 || st % P == 0 ? c->run.xpicOrder : 0;
}

// the kernels of one whole step (NairnMPM::MPMStep task list, NairnMPM.cpp:952-1108) with the
// XPIC order in c->xpicStep; with profiling on, an event is recorded between kernel groups
void mpmi::enqueue_step(MpmCtx *c) {
    hipStream_t s = c->stream;
    set_positions(c);
    if (c->profiling) c->ev.begin(s);
    auto done = [&](int group) {
        if (c->profiling) c->ev.end(group, s);
    };
    const bool mm = c->kg.nfields > 1;
    launch_p2g_mass(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, s);
    done(MPMGPU_KG_P2G_MASS);
    // grid velocity conditions follow each contact pass (UpdateMomentaTask::ContactAndMomentaBCs)
    const bool bcs = c->kbc.nb > 0;
    const int par = c->parity;
    auto vel_bcs = [&](int pass) {
        if (bcs) launch_vel_bcs(c->kg, c->kn, c->kbc, pass, par, s);
    };
    launch_node_mass(c->dim, c->kg, c->kb, c->kn, s);
    if (!mm) vel_bcs(BC_PASS_MASS_MOMENTUM);
    done(MPMGPU_KG_NODE_MASS);
    if (mm) {      // PostExtrapolationTask: contact at MASS_MOMENTUM_CALL
        launch_p2g_contact(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, false, s);
        launch_node_contact(c->dim, c->kg, c->kb, c->kn, s);
        launch_contact_normals(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, CONTACT_MASS_MOMENTUM, s);
        launch_contact(c->dim, c->kg, c->kn, CONTACT_MASS_MOMENTUM, s);
        vel_bcs(BC_PASS_MASS_MOMENTUM);
        done(MPMGPU_KG_CONTACT);
    }
    // FMPM(k > 1) strain updates take the extrapolated grid velocities (UpdateStrainsFirstTask.cpp:106-114)
    const int xo = c->xpicStep;
    const bool fmpm = c->run.xpicFMPM != 0;
    auto xpic = [&](bool forcePass) {
        launch_xpic(c->dim, c->kg, c->kp, c->kb, c->kn, c->kbc, xo, fmpm, forcePass, par, s);
    };
    if (fmpm && xo > 1) xpic(false);
    launch_usf(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, s);
    done(MPMGPU_KG_USF);
    if (c->kld.nl) launch_particle_loads(c->kg, c->kld, par, s);    // InitializationTask's pFext
    launch_p2g_force(c->dim, c->kg, c->kp, c->kb, c->kn, c->kld, par, s);
    done(MPMGPU_KG_P2G_FORCE);
    launch_node_force(c->dim, c->kg, c->kb, c->kn, s);
    // the UPDATE_MOMENTUM pass is left out in XPIC and FMPM steps of order > 1 (NodalVelBC.cpp:365-370)
    vel_bcs(mm || xo > 1 ? BC_PASS_GRID_FORCES : BC_PASS_FORCES_MOMENTA);
    done(MPMGPU_KG_NODE_FORCE);
    if (mm) {      // UpdateMomentaTask: contact at UPDATE_MOMENTUM_CALL
        launch_contact_normals(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, CONTACT_UPDATE_MOMENTUM, s);
        launch_contact(c->dim, c->kg, c->kn, CONTACT_UPDATE_MOMENTUM, s);
        vel_bcs(BC_PASS_UPDATE_MOMENTUM);
        done(MPMGPU_KG_CONTACT);
    }
    // the particle update of XPIC(k > 1) and FMPM(k > 1) takes the extrapolated grid velocities
    // (UpdateParticlesTask.cpp:66-79)
    if (xo > 1) xpic(true);
    launch_g2p_p2g(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, xo == 0 ? 0 : (fmpm ? xo : -xo), s);
    done(MPMGPU_KG_G2P_P2G);
    launch_node_mom2(c->dim, c->kg, c->kb, c->kn, s);
    if (!mm) vel_bcs(BC_PASS_UPDATE_STRAINS_LAST);
    done(MPMGPU_KG_NODE_MOM2);
    if (mm) {      // UpdateStrainsLastContactTask: contact at UPDATE_STRAINS_LAST_CALL
        launch_p2g_contact(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, true, s);
        launch_node_contact(c->dim, c->kg, c->kb, c->kn, s);
        launch_contact_normals(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, CONTACT_UPDATE_STRAINS_LAST, s);
        launch_contact(c->dim, c->kg, c->kn, CONTACT_UPDATE_STRAINS_LAST, s);
        vel_bcs(BC_PASS_UPDATE_STRAINS_LAST);
        done(MPMGPU_KG_CONTACT);
    }
    if (fmpm && xo > 1) xpic(false);
    launch_usl_reset(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, s);
    done(MPMGPU_KG_USL_RESET);
    launch_bin_finish(c->dim, c->kg, c->kp, c->kb, s);
    done(MPMGPU_KG_BINNING);
    c->parity ^= 1;
}

// events recorded between two folds of the kernel-group timings (profiling mode)
#define MAX_LIVE_EVENTS 1024

// fold the recorded kernel-group events into per-group totals (after a stream sync); a failed
// event call of profiling mode is returned as the context's error
static int collect_marks(MpmCtx *c) {
    c->ev.fold();
    if (!c->ev.err.empty()) {
        std::string m = "kernel-group timing: " + c->ev.err;
        c->ev.clear_error();
        return c->fail(m);
    }
    return 0;
}

// The device's run stop (ResetElementsTask's exceptions), read after a wait for the stream. A
// stopped run keeps the particle state of the step that stopped it: the positions that step wrote
// are the current ones and the step count is that of the steps before it.
static int read_stop(MpmCtx *c, unsigned long long *why) {
    unsigned long long s[3] = {0, 0, 0};
    HIPCHECK(c, hipMemcpy(s, c->kb.stop, sizeof(s), hipMemcpyDeviceToHost));
    *why = stop_reason(s[0]);
    if (s[0] && !c->stopped) {
        c->stopped = true;
        c->parity = (int)((s[1] + 1) & 1ull);
        c->steps = (int64_t)s[1];
        c->mtime = (double)s[1] * c->run.timestep;
        c->nextTask = 0;
    }
    return 0;
}

static const char *stop_message(unsigned long long why) {
    switch (why) {
        case STOP_TOO_MANY_LEFT: return "Too many particles have left the grid";
        case STOP_NAN: return "Particle left grid with position nan.";
        case STOP_DELETE:
            return "A particle has left the grid and would be deleted into the reservoir, which is not supported "
                   "(set <LeaveLimit> above 1 to push leaving particles back)";
        default: return "the run stopped";
    }
}

// the run's error once the device has stopped it
static int check_stop(MpmCtx *c) {
    unsigned long long why = 0;
    if (read_stop(c, &why)) return 1;
    if (why) return c->fail(std::string(stop_message(why)) + " (step " + std::to_string(c->steps + 1) + ")");
    return 0;
}

// ---- input checks and grid geometry (shared with mpm_static.hip) ----------------------

const char *mpmi::check_run(const MpmGridDesc *grid, const MpmRunParams *run) {
    if (grid->analysis != MPMGPU_PLANE_STRAIN && grid->analysis != MPMGPU_THREED)
        return "analysis type must be 10 (plane strain) or 12 (3D)";
    if (run->shape != MPMGPU_SHAPE_UGIMP && run->shape != MPMGPU_SHAPE_LCPDI && run->shape != MPMGPU_SHAPE_LINEAR)
        return "shape functions must be classic (0), uGIMP (1) or lCPDI (10) on the GPU path";
    if (run->mpmMethod != 2) return "only the USAVG method is implemented";
    if (run->xpicOrder < 0 || run->xpicOrder > MPMGPU_MAX_XPIC_ORDER)
        return "the XPIC or FMPM order must be 0 (FLIP) to 32";
    if (run->xpicOrder > 0 && run->xpicPeriodSteps <= 0 && !(run->xpicPeriodTime > 0.))
        return "an XPIC or FMPM order needs a period in steps or a period time above 0";
    return nullptr;
}

const char *mpmi::check_materials(const MpmMaterial *mats, int nmat) {
    for (int i = 0; i < nmat; i++) {
        int t = mats[i].type;
        if (t != MPMGPU_MAT_ISOTROPIC && t != MPMGPU_MAT_ISOPLASTICITY && t != MPMGPU_MAT_NEOHOOKEAN &&
            t != MPMGPU_MAT_MOONEY && t != MPMGPU_MAT_COULOMB && t != MPMGPU_MAT_CONTACT)
            return "a material type is not implemented on the GPU path";
    }
    return nullptr;
}

// the tiles of grid g (cells set) for tile shape `shape`
void mpmi::set_tiles(KGrid &g, int shape) {
    const TileSize t = tile_size(g.dim, shape);
    g.tileShape = shape;
    g.ntx = (g.nx + t.tx - 1) / t.tx; g.nty = (g.ny + t.ty - 1) / t.ty; g.ntz = (g.nz + t.tz - 1) / t.tz;
    g.ntiles = g.ntx * g.nty * g.ntz;
}

// geometry, tiles (shape 0) and time step of one grid
void mpmi::fill_grid(KGrid &g, int dim, const MpmGridDesc *grid, const MpmRunParams *run) {
    memset(&g, 0, sizeof(g));
    g.dim = dim;
    g.nx = grid->nx; g.ny = grid->ny; g.nz = dim == 3 ? grid->nz : 1;
    g.nnx = g.nx + 1; g.nny = g.ny + 1; g.nnz = dim == 3 ? g.nz + 1 : 1;
    set_tiles(g, 0);
    g.nnodes = (long long)g.nnx * g.nny * g.nnz;
    g.nstride = g.nnodes;
    g.tilesTotal = g.ntiles;
    g.xmin = grid->xmin; g.ymin = grid->ymin; g.zmin = grid->zmin;
    g.dx = grid->dx; g.dy = grid->dy; g.dz = grid->dz;
    g.timestep = run->timestep; g.strainFirst = run->strainFirst; g.strainLast = run->strainLast;
    g.gridDamping = run->gridDamping; g.particleDamping = run->particleDamping;
    g.damped = run->gridDamping != 0. || run->particleDamping != 0.;
    for (int a = 0; a < 3; a++) g.gravity[a] = run->gravity[a];
    g.hasGravity = run->gravity[0] != 0. || run->gravity[1] != 0. || run->gravity[2] != 0.;
    g.nfields = 1;
    g.shape = run->shape;
    g.uniformLp = 0;
    g.maxLeaving = run->maxLeaving > 0 ? run->maxLeaving : 0;
    g.deleteLeaving = run->deleteLeaving ? 1 : 0;
    g.stopOnNan = 1;
}

// ints of the binning arrays for nk cell keys and ntall (field, tile) pairs
size_t mpmi::bin_words(long long nk, long long ntall) {
    return (size_t)(3 * nk + 2 + (nk + 1 + 1023) / 1024 + 1 + ntall + 1 + TILE_SLOTS * 8 * TILE_SLOT_STRIDE + 3 * ntall);
}

extern "C" {

const char *mpmgpu_global_error(void) { return g_err.c_str(); }
const char *mpmgpu_last_error(const MpmCtx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

// ---- problem building ------------------------------------------------------------------

int mpmgpu_problem_create(int32_t analysis, MpmProblem **out) {
    if (analysis != MPMGPU_PLANE_STRAIN && analysis != MPMGPU_THREED)
        return fail_global("analysis type " + std::to_string(analysis) + " is not supported (10 or 12)");
    MpmProblem *p = new MpmProblem();
    p->analysis = analysis;
    p->dim = analysis == MPMGPU_THREED ? 3 : 2;
    if (p->dim == 3) p->ptsPerSide = 2;
    p->run.shape = MPMGPU_SHAPE_UGIMP;
    p->run.mpmMethod = 2;
    p->run.contactLaw = -1;
    *out = p;
    return 0;
}

// <Grid xmin= xmax= ...> with <Horiz cellsize=>/<Vert>/<Depth> (or nx,ny,nz counts when the
// cell size is <= 0)
int mpmgpu_problem_set_grid(MpmProblem *p, const double *ext6, const double *cell3, const int32_t *n3,
                            double thickness) {
    p->haveGrid = true;
    p->gxmin = ext6[0]; p->gxmax = ext6[1]; p->gymin = ext6[2]; p->gymax = ext6[3];
    p->gzmin = ext6[4]; p->gzmax = ext6[5];
    p->cellx = cell3[0]; p->celly = cell3[1]; p->cellz = cell3[2];
    p->nhoriz = n3 ? n3[0] : 0; p->nvert = n3 ? n3[1] : 0; p->ndepth = n3 ? n3[2] : 0;
    if (thickness > 0.) p->thickness = thickness;
    p->gridBuilt = false;
    if (!mpmsetup::build_grid(*p)) return fail_global(p->err);
    return 0;
}

// MPMHeader settings by element name: MatlPtsPerElement, CFL (Timing CFL=), TimeStep, MaxTime,
// GIMP (1 uGIMP, 0 classic, 10 lCPDI)
int mpmgpu_problem_set(MpmProblem *p, const char *key, double value) {
    std::string k(key);
    if (k == "MatlPtsPerElement") {
        int n = (int)(value + 0.5);
        int side = p->dim == 3 ? (int)(cbrt((double)n) + 0.5) : (int)(sqrt((double)n) + 0.5);
        if ((p->dim == 3 && side * side * side != n) || (p->dim == 2 && side * side != n))
            return fail_global("MatlPtsPerElement must be a square (2D) or cube (3D)");
        p->ptsPerSide = side;
    } else if (k == "CFL") p->cfl = value;
    else if (k == "TimeStep") p->timestepIn = value;
    else if (k == "MaxTime") p->maxTime = value;
    else if (k == "LeaveLimit") p->leaveLimit = (int)value;
    else if (k == "GIMP") p->run.shape = (int)value;
    else if (k == "MultiMaterialMode") {       // value = Normals method
        p->mmRequested = true;
        p->normalsGiven = true;
        p->run.normalsMethod = (int)value;
    } else if (k == "ContactLaw") p->contactLawNum = (int)value;   // 1-based material id
    else return fail_global("unknown setting " + k);
    p->finalized = false;
    return 0;
}

// <Material Type= Name=> (Read_MPM/MaterialController.cpp:105-190); returns the 1-based id
int mpmgpu_problem_add_material(MpmProblem *p, int32_t type, const char *name, int32_t *id) {
    MaterialInput m;
    m.type = type;
    m.name = name ? name : "";
    mpmsetup::add_material(*p, m);
    *id = (int32_t)p->matIn.size();
    p->finalized = false;
    return 0;
}

// one property element of a <Material> (<E>, <nu>, <G>, <K>, <rho>, <yield>, <Ep>, ...), in
// the XML's units
int mpmgpu_problem_material_property(MpmProblem *p, int32_t id, const char *prop, double value) {
    if (id < 1 || id > (int)p->matIn.size()) return fail_global("bad material id");
    MaterialInput &m = p->matIn[id - 1];
    std::string k(prop);
    if (k == "rho") m.rho = value;
    else if (k == "E" || k == "Etens") { m.E = value; m.readE = true; }
    else if (k == "nu") { m.nu = value; m.readNu = true; }
    else if (k == "G" || k == "G1") { m.G = value; m.readG = true; }
    else if (k == "K" || k == "Kbulk") m.K = value;
    else if (k == "Lame") m.Lame = value;
    else if (k == "G2") m.G2 = value;
    else if (k == "yield") m.yield = value;
    else if (k == "yieldMin") m.yieldMin = value;
    else if (k == "Ep") m.Ep = value;
    else if (k == "Khard") m.Khard = value;
    else if (k == "coeff") m.friction = value;
    else if (k == "coeffStatic") m.frictionStatic = value;
    else if (k == "UJOption") m.ujOption = (int)value;
    else if (k == "alpha" || k == "beta" || k == "Cv" || k == "kCond" || k == "D" || k == "csat") {
        if (value != 0. && (k == "alpha" || k == "beta")) {
            // thermal and moisture expansion only act with transport tasks, which are not on
            // this path; the expansion coefficients are accepted and unused
        }
    } else return fail_global("material property " + k + " is not supported on this path");
    p->finalized = false;
    return 0;
}

// <Body mat= vx= vy= vz=> with one shape (<Box>/<Cylinder axis=>, <Rect>, <Oval>, <Sphere>)
int mpmgpu_problem_add_body(MpmProblem *p, int32_t mat, const char *shape, const double *ext6,
                            int32_t axis, const double *vel3, double thick) {
    std::string s(shape);
    int kind;
    if (s == "Box" || s == "Cylinder") kind = 0;
    else if (s == "Rect") kind = 1;
    else if (s == "Oval") kind = 2;
    else if (s == "Sphere") kind = 3;
    else return fail_global("shape " + s + " is not supported on this path");
    if (mat < 1) return fail_global("Positive material ID needed in <Body> element.");
    if (!mpmsetup::add_body(*p, mat, kind, ext6, s == "Cylinder" && axis == 0 ? 3 : axis, vel3, thick))
        return fail_global(p->err);
    p->finalized = false;
    return 0;
}

// explicit <mp> particle with <pt>, <vel> (and optional mass, <0 for rho*volume)
int mpmgpu_problem_add_point(MpmProblem *p, int32_t mat, const double *pos3, const double *vel3,
                             const double *lp3, double thick, double mass) {
    if (!mpmsetup::add_point(*p, mat, pos3, vel3, lp3, thick, mass)) return fail_global(p->err);
    p->finalized = false;
    return 0;
}

int mpmgpu_problem_from_xml(const char *xml_text, MpmProblem **out) {
    MpmProblem *p = new MpmProblem();
    if (!mpm_read_xml(xml_text, *p)) {
        g_err = p->err;
        delete p;
        return 1;
    }
    *out = p;
    return 0;
}

int mpmgpu_problem_describe(const MpmProblem *pc, MpmGridDesc *grid, MpmMaterial *mats, int32_t *nmat,
                            MpmRunParams *run, int64_t *nparticles, double *maxTime) {
    MpmProblem *p = const_cast<MpmProblem *>(pc);
    if (!mpmsetup::finalize(*p)) return fail_global(p->err);
    if (grid) *grid = p->grid;
    if (mats) for (size_t i = 0; i < p->mats.size(); i++) mats[i] = p->mats[i];
    if (nmat) *nmat = (int32_t)p->mats.size();
    if (run) *run = p->run;
    if (nparticles) *nparticles = p->nparticles();
    if (maxTime) *maxTime = p->maxTime;
    return 0;
}

int mpmgpu_problem_particles(const MpmProblem *pc, double *dfields, int32_t *ifields) {
    MpmProblem *p = const_cast<MpmProblem *>(pc);
    if (!mpmsetup::finalize(*p)) return fail_global(p->err);
    size_t n = (size_t)p->nparticles();
    for (int f = 0; f < MPF_NUM_FIELDS; f++)
        if (n) memcpy(dfields + f * n, p->d[f].data(), n * sizeof(double));
    if (ifields)
        for (int f = 0; f < MPI_NUM_FIELDS; f++)
            if (n) memcpy(ifields + f * n, p->iv[f].data(), n * sizeof(int32_t));
    return 0;
}

int mpmgpu_problem_archive(const MpmProblem *pc, char *order, double *archiveTime, double *thickness) {
    MpmProblem *p = const_cast<MpmProblem *>(pc);
    if (!mpmsetup::finalize(*p)) return fail_global(p->err);
    if (order) {
        std::string o = p->archiveOrder;
        if (o.size() > 25) o.resize(25);
        memcpy(order, o.c_str(), o.size() + 1);
    }
    if (archiveTime) *archiveTime = p->archiveTime;
    if (thickness)
        for (size_t i = 0; i < (size_t)p->nparticles(); i++)
            thickness[i] = i < p->pthick.size() ? p->pthick[i] : p->thickness;
    return 0;
}

int mpmgpu_problem_velocity_bcs(const MpmProblem *pc, MpmVelocityBC *bcs, int64_t *n) {
    MpmProblem *p = const_cast<MpmProblem *>(pc);
    if (!n) return fail_global("null count");
    if (!mpmsetup::finalize(*p)) return fail_global(p->err);
    const int64_t have = (int64_t)p->velBCs.size();
    if (bcs) {
        const int64_t k = std::min(*n, have);
        for (int64_t i = 0; i < k; i++) bcs[i] = p->velBCs[i];
    }
    *n = have;
    return 0;
}

int mpmgpu_problem_particle_loads(const MpmProblem *pc, MpmParticleLoad *loads, int64_t *n) {
    MpmProblem *p = const_cast<MpmProblem *>(pc);
    if (!n) return fail_global("null count");
    if (!mpmsetup::finalize(*p)) return fail_global(p->err);
    const int64_t have = (int64_t)p->loads.size();
    if (loads) {
        const int64_t k = std::min(*n, have);
        for (int64_t i = 0; i < k; i++) loads[i] = p->loads[i];
    }
    *n = have;
    return 0;
}

int mpmgpu_problem_particle_tractions(const MpmProblem *pc, MpmParticleTraction *tr, int64_t *n) {
    MpmProblem *p = const_cast<MpmProblem *>(pc);
    if (!n) return fail_global("null count");
    if (!mpmsetup::finalize(*p)) return fail_global(p->err);
    const int64_t have = (int64_t)p->tractions.size();
    if (tr) {
        const int64_t k = std::min(*n, have);
        for (int64_t i = 0; i < k; i++) tr[i] = p->tractions[i];
    }
    *n = have;
    return 0;
}

void mpmgpu_problem_destroy(MpmProblem *p) { delete p; }

// ---- device context --------------------------------------------------------------------

}  // extern "C"

// tiles per field of the context (the sum over a batch's instances) for tile shape `shape`
long long mpmi::tiles_total(const MpmCtx *c, int shape) {
    if (c->inst.empty()) {
        KGrid g = c->kg;
        set_tiles(g, shape);
        return g.ntiles;
    }
    long long t = 0;
    for (KGrid g : c->inst) {
        set_tiles(g, shape);
        t += g.ntiles;
    }
    return t;
}

// sizes of the grid, bin and counter arrays of a context for c->kg (nstride nodes; partials and
// bins for the larger of the tile shapes)
mpmi::GridArrays mpmi::grid_arrays(const MpmCtx *c, int nmat) {
    const KGrid &g = c->kg;
    const int nf = g.nfields;
    const long long nn = g.nstride;
    GridArrays a;
    // per field: mass, pk, vk, ftot (+ contact volume, displacement, gradient in multimaterial mode;
    // + the XPIC increment with XPIC or FMPM of order > 1)
    const int perField = 1 + 3 + 3 + 3 + (nf > 1 ? 1 + 3 + 3 : 0) + (c->run.xpicOrder > 1 ? 3 : 0);
    a.nodeBytes = (size_t)nf * nn * perField * sizeof(double);
    a.cntBytes = (size_t)nf * nn * sizeof(int);
    a.matBytes = (size_t)nmat * sizeof(MpmMaterial);
    a.statBytes = STATS_WORDS * sizeof(unsigned long long);
    // tile partials and bins: enough for either tile shape, so the shape can change in place
    for (int shape = 0; shape < MPM_TILE_SHAPES; shape++) {
        const TileSize t = tile_size(c->dim, shape);
        const long long tiles = tiles_total(c, shape);
        a.partBytes = std::max<size_t>(a.partBytes, (size_t)nf * tiles * NQMAX * t.foot * sizeof(double));
        a.binBytes = std::max<size_t>(a.binBytes, bin_words((long long)nf * tiles * t.nc, (long long)nf * tiles) * sizeof(int));
    }
    if (nf > 1 && g.normals >= 5) {
        a.normBytes = (size_t)nf * 4 * nn * sizeof(double);
        a.listBytes = (size_t)(nn + 1) * sizeof(int);
    }
    return a;
}

// the context's grid, bin and counter pointers into the given arrays (sized by grid_arrays)
void mpmi::assign_grid_arrays(MpmCtx *c, const GridArrays &a, double *nodeMem, int *cntMem, double *partMem,
                              int *bbuf, unsigned long long *sbuf, double *normMem, int *listMem) {
    const KGrid &g = c->kg;
    const int nf = g.nfields;
    const long long fnn = (long long)nf * g.nstride;
    c->kn.mass = nodeMem;
    c->kn.pk = nodeMem + fnn;
    c->kn.vk = nodeMem + 4 * fnn;
    c->kn.ftot = nodeMem + 7 * fnn;
    c->kn.cnt = cntMem;
    c->kn.part = partMem;
    c->kn.cvol = c->kn.cdisp = c->kn.cgrad = nullptr;
    if (nf > 1) {
        c->kn.cvol = nodeMem + 10 * fnn;
        c->kn.cdisp = nodeMem + 11 * fnn;
        c->kn.cgrad = nodeMem + 14 * fnn;
    }
    c->kn.vprev = c->run.xpicOrder > 1 ? nodeMem + (nf > 1 ? 17 : 10) * fnn : nullptr;
    c->sbuf = sbuf;
    c->ks.leftGrid = sbuf;
    c->ks.nanCount = sbuf + 1;
    c->ks.cpdiErr = sbuf + 2;
    c->kb.stop = sbuf + 3;
    c->ks.lrOverflow = sbuf + 6;
    c->ks.tracErr = sbuf + 7;
    c->kn.cnorm = a.normBytes ? normMem : nullptr;
    c->kn.clist = a.listBytes ? listMem : nullptr;
    c->bbuf = bbuf;
    c->bbufBytes = a.binBytes;
    assign_bins(c);
}

// the binning arrays in c->bbuf for the current tile shape: cellStart(nkeys+1), count(nkeys+1),
// cursor(nkeys), blockSums, tile list (nfields * tilesTotal) and its length, work counters, and
// the node passes' pair flags
void mpmi::assign_bins(MpmCtx *c) {
    const KGrid &g = c->kg;
    const int nf = g.nfields;
    const long long nk = (long long)nf * g.tilesTotal * tile_size(c->dim, g.tileShape).nc;
    const long long nblk = (nk + 1 + 1023) / 1024 + 1;
    int *bbuf = c->bbuf;
    c->kb.nkeys = (int)nk;
    c->kb.ntall = nf * g.tilesTotal;
    c->kb.cellStart = bbuf;
    c->kb.count = bbuf + nk + 1;
    c->kb.cursor = bbuf + 2 * nk + 2;
    c->kb.blockSums = bbuf + 3 * nk + 2;
    c->kb.tiles = c->kb.blockSums + nblk;
    c->kb.ntiles = c->kb.tiles + c->kb.ntall;
    c->kb.work = c->kb.ntiles + 1;
    c->kb.need = c->kb.work + TILE_SLOTS * 8 * TILE_SLOT_STRIDE;
    c->kb.needAny = c->kb.need + c->kb.ntall;
    c->kb.tileOn = c->kb.needAny + c->kb.ntall;
    c->kb.slot = 0;
}

// the tile kernels' grids fill every CU of the current device at the kernels' occupancy
void mpmi::size_tile_grids() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0)
        set_tile_cus(cus);
}

// device buffers of the grid, bins and counters for c->kg (nstride nodes, tilesTotal tiles per field)
static int alloc_context(MpmCtx *c, const MpmMaterial *mats, int nmat) {
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail_global(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    const mpmi::GridArrays a = mpmi::grid_arrays(c, nmat);
    double *nodeMem = nullptr;
    int *cntMem = nullptr;
    double *partMem = nullptr;
    unsigned long long *sbuf = nullptr;
    int *bbuf = nullptr;
    if (hipMalloc(&nodeMem, a.nodeBytes) != hipSuccess || hipMalloc(&cntMem, a.cntBytes) != hipSuccess ||
        hipMalloc(&partMem, a.partBytes) != hipSuccess || hipMalloc(&c->dmats, a.matBytes) != hipSuccess ||
        hipMalloc(&sbuf, a.statBytes) != hipSuccess) {
        if (nodeMem) (void)hipFree(nodeMem);
        if (cntMem) (void)hipFree(cntMem);
        if (partMem) (void)hipFree(partMem);
        if (sbuf) (void)hipFree(sbuf);
        return fail_global("out of device memory allocating the grid");
    }
    if (hipMalloc(&bbuf, a.binBytes) != hipSuccess) {
        (void)hipFree(nodeMem);
        (void)hipFree(cntMem);
        (void)hipFree(partMem);
        (void)hipFree(sbuf);
        return fail_global("out of device memory allocating the bins");
    }
    if (a.normBytes && (hipMalloc(&c->lrbuf, a.normBytes) != hipSuccess ||
                        hipMalloc(&c->lrlist, a.listBytes) != hipSuccess)) {
        (void)hipFree(nodeMem);
        (void)hipFree(cntMem);
        (void)hipFree(partMem);
        (void)hipFree(sbuf);
        (void)hipFree(bbuf);
        return fail_global("out of device memory allocating the contact normals");
    }
    mpmi::assign_grid_arrays(c, a, nodeMem, cntMem, partMem, bbuf, sbuf, c->lrbuf, c->lrlist);
    if (a.listBytes) (void)hipMemsetAsync(c->lrlist, 0, a.listBytes, c->stream);
    (void)hipMemsetAsync(nodeMem, 0, a.nodeBytes, c->stream);
    (void)hipMemsetAsync(cntMem, 0, a.cntBytes, c->stream);
    (void)hipMemsetAsync(sbuf, 0, a.statBytes, c->stream);
    (void)hipMemcpyAsync(c->dmats, mats, a.matBytes, hipMemcpyHostToDevice, c->stream);
    (void)hipMemsetAsync(bbuf, 0, a.binBytes, c->stream);
    mpmi::size_tile_grids();
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail_global("device initialisation failed");
    return 0;
}

const char *mpmi::single_grid(MpmCtx *c, const MpmGridDesc *grid, const MpmMaterial *mats, int nmat,
                              const MpmRunParams *run) {
    if (const char *m = check_run(grid, run)) return m;
    int nf = run->multiMaterial ? run->numFields : 1;
    if (nf < 1 || nf > MPMGPU_MAX_MATERIALS) return "bad number of material velocity fields";
    if (nf > 1 && run->xpicOrder > 0) return "XPIC and FMPM are not available in multimaterial mode on this path";
    if (nmat < 1 || nmat > MPMGPU_MAX_MATERIALS) return "bad material count";
    if (nf > 1) {
        if (run->contactLaw < 0 || run->contactLaw >= nmat ||
            (mats[run->contactLaw].type != MPMGPU_MAT_COULOMB && mats[run->contactLaw].type != MPMGPU_MAT_CONTACT))
            return "multimaterial mode needs a contact law material";
        if (run->normalsMethod < 0 || run->normalsMethod > 6)
            return "contact normals method must be 0 (MAXG), 1 (MAXV), 2 (AVGG), 3 (OWNG), 4 (specified), "
                   "5 (linear regression) or 6 (logistic regression)";
    }
    if (const char *m = check_materials(mats, nmat)) return m;
    c->dim = grid->analysis == MPMGPU_THREED ? 3 : 2;
    c->gd = *grid;
    c->run = *run;
    c->mats.assign(mats, mats + nmat);
    KGrid &g = c->kg;
    fill_grid(g, c->dim, grid, run);
    for (int m = 0; m < nmat; m++) g.damped |= mats[m].usePDamping;   // a material's own particle damping
    g.nfields = nf;
    for (int m = 0; m < MPMGPU_MAX_MATERIALS; m++)
        g.fieldOf[m] = (nf > 1 && m < nmat && mats[m].field >= 0 && mats[m].field < nf) ? mats[m].field : 0;
    if (nf > 1) {
        const MpmMaterial &law = mats[run->contactLaw];
        g.normals = run->normalsMethod;
        g.ignoreContact = law.type == MPMGPU_MAT_CONTACT ? 1 : 0;
        g.friction = law.friction;
        g.frictionStatic = law.frictionStatic;
        for (int a = 0; a < 3; a++) g.contactNormal[a] = run->contactNormal[a];
    }
    return nullptr;
}

extern "C" {

int mpmgpu_create(const MpmGridDesc *grid, const MpmMaterial *mats, int32_t nmat, const MpmRunParams *run,
                  MpmCtx **out) {
    *out = nullptr;
    MpmCtx *c = new MpmCtx();
    if (const char *m = mpmi::single_grid(c, grid, mats, nmat, run)) {
        delete c;
        return fail_global(m);
    }
    if (alloc_context(c, mats, nmat)) { mpmgpu_destroy(c); return 1; }
    *out = c;
    return 0;
}

int64_t mpmgpu_footprint(const MpmGridDesc *grid, const MpmRunParams *run, int64_t n) {
    if (check_run(grid, run)) return -1;
    const int dim = grid->analysis == MPMGPU_THREED ? 3 : 2;
    KGrid g;
    fill_grid(g, dim, grid, run);
    const long long nodes = (g.nnodes + 255) / 256 * 256;        // as laid out in a batch
    int64_t bytes = nodes * (10 * sizeof(double) + sizeof(int));   // mass, pk, vk, ftot, cnt
    if (run->xpicOrder > 1) bytes += nodes * 3 * (int64_t)sizeof(double);   // the XPIC increment
    int64_t tileBytes = 0;          // tile partials and bins, for the larger of the tile shapes
    for (int shape = 0; shape < MPM_TILE_SHAPES; shape++) {
        set_tiles(g, shape);
        const TileSize t = tile_size(dim, shape);
        tileBytes = std::max<int64_t>(tileBytes, (int64_t)g.ntiles * NQMAX * t.foot * (int64_t)sizeof(double) +
                                                     (int64_t)bin_words((long long)g.ntiles * t.nc, g.ntiles) * (int64_t)sizeof(int));
    }
    bytes += tileBytes;
    bytes += n * PARTICLE_BYTES;
    if (run->shape == MPMGPU_SHAPE_LCPDI) bytes += n * 9 * (int64_t)sizeof(double);
    return bytes;
}

}  // extern "C"

// A batch's instance grids and its tables (tile -> instance for the current tile layout, node
// block -> instance, material -> instance) on the device, and the context grid's pointers to them.
static int upload_batch_tables(MpmCtx *c) {
    const int ninst = (int)c->inst.size();
    KGrid &g = c->kg;
    const int tileAcc = g.tilesTotal;
    const long long nodeAcc = g.nstride;
    if (tileAcc > c->tileCap) return c->fail("batch tile table too small for the tile layout");
    std::vector<int> tables;
    tables.reserve(tileAcc + nodeAcc / 256 + c->matInstH.size());
    for (int i = 0; i < ninst; i++) tables.insert(tables.end(), c->inst[i].ntiles, i);
    for (int i = 0; i < ninst; i++) tables.insert(tables.end(), (c->inst[i].nnodes + 255) / 256, i);
    tables.insert(tables.end(), c->matInstH.begin(), c->matInstH.end());
    HIPCHECK(c, hipMemcpy(c->dtables, tables.data(), tables.size() * sizeof(int), hipMemcpyHostToDevice));
    g.inst = c->dinst;
    g.tileInst = c->dtables;
    g.nodeBlockInst = c->dtables + tileAcc;
    g.matInst = c->dtables + tileAcc + nodeAcc / 256;
    HIPCHECK(c, hipMemcpy(c->dinst, c->inst.data(), ninst * sizeof(KGrid), hipMemcpyHostToDevice));
    return 0;
}

// A batch's tile -> material type table for the current layout (kg.tileType), from the instance
// types of the last upload
static int upload_tile_types(MpmCtx *c) {
    if (!c->dtileType || c->instType.size() != c->inst.size()) return 0;
    std::vector<int> tt((size_t)c->kg.tilesTotal, -1);
    for (size_t k = 0; k < c->inst.size(); k++)
        for (int t = 0; t < c->inst[k].ntiles; t++) tt[(size_t)c->inst[k].tileBase + t] = c->instType[k];
    HIPCHECK(c, hipMemcpy(c->dtileType, tt.data(), tt.size() * sizeof(int), hipMemcpyHostToDevice));
    c->kg.tileType = c->dtileType;
    return 0;
}

int mpmi::set_instance_types(MpmCtx *c, const std::vector<int> &types) {
    c->instType = types;
    KGrid &g = c->kg;
    g.ntypes = 0;
    g.anyMixed = 0;
    for (int it : c->instType) {
        if (it == 0) { g.anyMixed = 1; continue; }
        if (it < 0) continue;
        bool have = false;
        for (int j = 0; j < g.ntypes; j++) have |= g.types[j] == it;
        if (!have && g.ntypes < 4) g.types[g.ntypes++] = it;
        else if (!have) g.anyMixed = 1;
    }
    if (g.anyMixed)          // a type beyond the four listed (none on this path) takes the all-laws launch
        for (int &it : c->instType) {
            bool listed = false;
            for (int j = 0; j < g.ntypes; j++) listed |= g.types[j] == it;
            if (it > 0 && !listed) it = 0;
        }
    return upload_tile_types(c);
}

// The tile layout of the context for another tile shape: the grid's (or each instance's) tiles,
// a batch's tables, the bin pointers and the loaded particles' per-tile flags. With rebin the
// particles are binned for it, the bins' completed-step count kept, every node visited by the
// next step and the rows sorted into the new tile order.
static int apply_tile_shape(MpmCtx *c, int shape, bool rebin) {
    if (shape != c->kg.tileShape) {
        destroy_graphs(c);
        if (c->inst.empty()) {
            mpmi::set_tiles(c->kg, shape);
            c->kg.tilesTotal = c->kg.ntiles;
        } else {
            int acc = 0;
            for (KGrid &gi : c->inst) {
                mpmi::set_tiles(gi, shape);
                gi.tileBase = acc;
                acc += gi.ntiles;
            }
            for (KGrid &gi : c->inst) gi.tilesTotal = acc;
            KGrid &g = c->kg;
            const KGrid &g0 = c->inst[0];
            g.tileShape = shape;
            g.ntx = g0.ntx; g.nty = g0.nty; g.ntz = g0.ntz; g.ntiles = g0.ntiles;
            g.tilesTotal = acc;
            if (upload_batch_tables(c) || upload_tile_types(c)) return 1;
        }
        // the histogram and cursors of the new layout sit where the old one kept other arrays:
        // every bin array starts from zero, as at creation
        mpmi::assign_bins(c);
        HIPCHECK(c, hipMemsetAsync(c->bbuf, 0, c->bbufBytes, c->stream));
        if (c->kld.nl && build_particle_bcs(c)) return 1;
    }
    if (!rebin || c->kp.n == 0) return 0;
    unsigned long long done = 0;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(&done, c->kb.stop + 2, sizeof(done), hipMemcpyDeviceToHost));
    set_positions(c);
    launch_bin_keys(c->dim, c->kg, c->kp, c->kb, c->stream);
    launch_bin_finish(c->dim, c->kg, c->kp, c->kb, c->stream);
    HIPCHECK(c, hipMemcpyAsync(c->kb.stop + 2, &done, sizeof(done), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, mark_all_needed(c));
    if (mpmi::sort_rows(c)) return 1;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int mpmi::set_tile_shape(MpmCtx *c, int shape) {
    if (shape < 0 || shape >= MPM_TILE_SHAPES) return c->fail("tile shape must be 0 or 1");
    if (c->nextTask != 0) return c->fail("a step is half done through mpmgpu_run_task");
    return apply_tile_shape(c, shape, true);
}

extern "C" {

// Independent instances (each its own grid, time step, materials and particles) stepped together:
// their tiles and nodes are laid out one after another in one context, every kernel looks up the
// instance of its tile or node, and one launch per kernel serves all of them.
int mpmgpu_batch_create(int32_t ninst, const MpmInstanceDesc *desc, const MpmMaterial *mats, int32_t nmat,
                        MpmCtx **out) {
    *out = nullptr;
    if (ninst < 1) return fail_global("a batch needs at least one instance");
    if (nmat < 1 || nmat > MPMGPU_MAX_BATCH_MATERIALS) return fail_global("bad material count");
    if (const char *m = check_materials(mats, nmat)) return fail_global(m);
    int analysis = desc[0].grid.analysis, shape = desc[0].run.shape;
    std::vector<int> matInst(nmat, -1);
    for (int i = 0; i < ninst; i++) {
        const MpmInstanceDesc &d = desc[i];
        if (const char *m = check_run(&d.grid, &d.run)) return fail_global(m);
        if (d.grid.analysis != analysis || d.run.shape != shape)
            return fail_global("all instances of a batch need the same analysis type and shape functions");
        if (d.run.multiMaterial) return fail_global("multimaterial instances cannot be batched");
        if (d.run.xpicOrder > 0) return fail_global("instances with XPIC or FMPM cannot be batched");
        if (d.firstMaterial < 0 || d.numMaterials < 1 || d.firstMaterial + d.numMaterials > nmat)
            return fail_global("instance material range outside the material table");
        for (int m = d.firstMaterial; m < d.firstMaterial + d.numMaterials; m++) {
            if (matInst[m] >= 0) return fail_global("a material belongs to two instances");
            matInst[m] = i;
        }
    }
    MpmCtx *c = new MpmCtx();
    c->dim = analysis == MPMGPU_THREED ? 3 : 2;
    c->gd = desc[0].grid;
    c->run = desc[0].run;
    c->mats.assign(mats, mats + nmat);
    c->matInstH = matInst;
    c->inst.resize(ninst);
    long long nodeAcc = 0;
    int tileAcc = 0;
    for (int i = 0; i < ninst; i++) {
        KGrid &gi = c->inst[i];
        fill_grid(gi, c->dim, &desc[i].grid, &desc[i].run);
        gi.maxLeaving = gi.deleteLeaving = gi.stopOnNan = 0;     // a batch never stops
        gi.tileBase = tileAcc;
        gi.nodeBase = nodeAcc;
        tileAcc += gi.ntiles;
        nodeAcc += (gi.nnodes + 255) / 256 * 256;
    }
    for (auto &gi : c->inst) { gi.nstride = nodeAcc; gi.tilesTotal = tileAcc; }
    KGrid &g = c->kg;
    g = c->inst[0];
    g.nstride = nodeAcc;
    g.tilesTotal = tileAcc;
    for (auto &gi : c->inst) g.damped |= gi.damped;     // the launch takes the damped instance
    for (int m = 0; m < nmat; m++) g.damped |= mats[m].usePDamping;
    if (alloc_context(c, mats, nmat)) { mpmgpu_destroy(c); return 1; }
    // instance tables on the device, with room for the tile count of either tile shape
    for (int sh = 0; sh < MPM_TILE_SHAPES; sh++) c->tileCap = std::max(c->tileCap, (int)mpmi::tiles_total(c, sh));
    if (hipMalloc(&c->dinst, ninst * sizeof(KGrid)) != hipSuccess ||
        hipMalloc(&c->dtables, ((size_t)c->tileCap + nodeAcc / 256 + nmat) * sizeof(int)) != hipSuccess ||
        hipMalloc(&c->dtileType, (size_t)c->tileCap * sizeof(int)) != hipSuccess) {
        mpmgpu_destroy(c);
        return fail_global("out of device memory allocating the batch tables");
    }
    if (upload_batch_tables(c)) {
        std::string m = c->err;
        mpmgpu_destroy(c);
        return fail_global(m);
    }
    *out = c;
    return 0;
}

}  // extern "C"

// ---- particle placement (shared with the solve server) ----------------------------------

int mpmi::place_particles(MpmCtx *c, int64_t n) {
    destroy_graphs(c);
    if (n > 0x7fffffffLL) return c->fail("too many particles for one context");
    if (n > c->ncap) {
        if (c->pbuf) { (void)hipFree(c->pbuf); c->pbuf = nullptr; }
        if (c->ibuf) { (void)hipFree(c->ibuf); c->ibuf = nullptr; }
        if (c->scr) { (void)hipFree(c->scr); c->scr = nullptr; }
        HIPCHECK(c, hipMalloc(&c->pbuf, (size_t)PROWS * n * sizeof(double)));
        HIPCHECK(c, hipMalloc(&c->ibuf, (size_t)IROWS * n * sizeof(int)));
        HIPCHECK(c, hipMalloc(&c->scr, (size_t)n * sizeof(double)));
        if (c->cbuf) { (void)hipFree(c->cbuf); c->cbuf = nullptr; }
        if (c->run.shape == MPMGPU_SHAPE_LCPDI) HIPCHECK(c, hipMalloc(&c->cbuf, (size_t)9 * n * sizeof(double)));
        c->ncap = n;
    }
    for (int a = 0; a < 9; a++) c->kp.cpr[a] = c->cbuf ? c->cbuf + (size_t)a * n : nullptr;
    size_t nn = (size_t)n;
    for (int f = 0; f < MPF_NUM_FIELDS; f++) c->kp.f[f] = c->pbuf + f * nn;
    for (int a = 0; a < 3; a++) {
        c->posA[a] = c->pbuf + (MPF_POS_X + a) * nn;
        c->posB[a] = c->pbuf + (MPF_NUM_FIELDS + a) * nn;
    }
    c->parity = 0;
    c->kp.elem = c->ibuf;
    c->kp.mat = c->ibuf + nn;
    c->kp.id = c->ibuf + 2 * nn;
    c->kp.left = c->ibuf + 3 * nn;
    c->kb.key = c->ibuf + IROW_KEY * nn;
    c->kb.order = c->ibuf + IROW_ORDER * nn;
    c->kb.home = c->ibuf + IROW_HOME * nn;
    c->kb.shome = c->ibuf + IROW_SHOME * nn;
    c->sinceSort = 0;
    c->kp.n = n;
    return 0;
}

// Sort the rows into the order of the current bins, so that a tile's particles sit together in
// memory. Results do not depend on it: inside a cell the particles are summed in upload order.
int mpmi::sort_rows(MpmCtx *c) {
    const long long n = c->kp.n;
    c->sinceSort = 0;
    if (n == 0) return 0;
    hipStream_t s = c->stream;
    for (int r = 0; r < PROWS; r++) {
        double *row = c->pbuf + (size_t)r * n;
        launch_gather_row(row, c->scr, c->kb.order, n, s);
        HIPCHECK(c, hipMemcpyAsync(row, c->scr, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    int *iscr = reinterpret_cast<int *>(c->scr);
    for (int r = 0; r < IROWS; r++) {
        if (r == IROW_ORDER || r == IROW_SHOME) continue;      // per slot, not per particle
        int *row = c->ibuf + (size_t)r * n;
        launch_gather_irow(row, iscr, c->kb.order, n, s);
        HIPCHECK(c, hipMemcpyAsync(row, iscr, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
    launch_iota(c->kb.order, n, 0, s);
    HIPCHECK(c, hipGetLastError());
    return 0;
}

int mpmi::prepare_particles(MpmCtx *c) {
    const size_t nn = (size_t)c->kp.n;
    // the second position buffer starts as a copy: in 2D its z row is never written by a step
    for (int a = 0; a < 3; a++)
        HIPCHECK(c, hipMemcpyAsync(c->posB[a], c->posA[a], nn * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->sbuf, 0, STATS_WORDS * sizeof(unsigned long long), c->stream));
    c->stopped = false;
    set_positions(c);
    launch_bin_keys(c->dim, c->kg, c->kp, c->kb, c->stream);
    launch_bin_finish(c->dim, c->kg, c->kp, c->kb, c->stream);
    // the tile list counts a completed step at every binning; this one is the upload's, not a step's
    HIPCHECK(c, hipMemsetAsync(c->kb.stop + 2, 0, sizeof(unsigned long long), c->stream));
    // the first step visits every node, clearing what an earlier upload left
    HIPCHECK(c, mark_all_needed(c));
    if (mpmi::sort_rows(c)) return 1;
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->nextTask = 0;
    c->steps = 0;
    c->mtime = 0.;
    c->bcTime = 0.;
    c->xpicClock = 0.;
    return put_bc_time(c);
}

extern "C" {

int mpmgpu_upload_particles(MpmCtx *c, int64_t n, const double *dfields, const int32_t *ifields) {
    if (n < 0 || n > 0x7fffffffLL) return c->fail("too many particles for one context");
    size_t nn = (size_t)n;
    // validate materials and elements on the host before anything reaches the device
    int nmat = (int)c->mats.size();
    const bool batch = !c->inst.empty();
    long long nelem = (long long)c->gd.nx * c->gd.ny * (c->dim == 3 ? c->gd.nz : 1);
    for (int64_t i = 0; i < n; i++) {
        int m = ifields[MPI_MAT * nn + i];
        long long e = ifields[MPI_ELEM * nn + i];
        if (m < 0 || m >= nmat) return c->fail("particle with an undefined material");
        if (batch) {
            const KGrid &gi = c->inst[c->matInstH[m]];
            nelem = (long long)gi.nx * gi.ny * gi.nz;
        }
        int t = c->mats[m].type;
        if (t == MPMGPU_MAT_COULOMB || t == MPMGPU_MAT_CONTACT) return c->fail("particle with a contact-law material");
        if (c->kg.nfields > 1 && (c->mats[m].field < 0 || c->mats[m].field >= c->kg.nfields))
            return c->fail("particle material without a velocity field");
        if (e < 0 || e >= nelem) return c->fail("particle element outside the grid");
        if (!(dfields[MPF_MASS * nn + i] > 0.)) return c->fail("particle with a mass that is not positive");
        for (int a = 0; a < c->dim; a++) {
            const double l = dfields[(MPF_LP_X + a) * nn + i];
            if (!(l > 0. && l <= 1.)) return c->fail("particle half size must be in (0, 1] cell half widths");
        }
        if (t == MPMGPU_MAT_ISOPLASTICITY && dfields[MPF_HIST0 * nn + i] == 0.)
            for (int f = MPF_ALT_XX; f <= MPF_ALT_XY; f++)
                if (dfields[f * nn + i] != 0.)
                    return c->fail("IsoPlasticity particle with plastic strain but zero cumulative plastic strain");
    }
    // one material type for all particles lets the strain kernels compile a single law
    c->kg.matType = 0;
    if (n > 0) {
        int t0 = c->mats[ifields[MPI_MAT * nn]].type;
        bool same = true;
        for (int64_t i = 1; i < n && same; i++) same = c->mats[ifields[MPI_MAT * nn + i]].type == t0;
        c->kg.matType = same ? t0 : 0;
    }
    // uniform particle size lets the kernels skip the lp arrays
    c->kg.uniformLp = 1;
    if (n > 0) {
        for (int a = 0; a < 3; a++) c->kg.lp[a] = dfields[(MPF_LP_X + a) * nn];
        for (int64_t i = 0; i < n && c->kg.uniformLp; i++)
            for (int a = 0; a < 3; a++)
                if (dfields[(MPF_LP_X + a) * nn + i] != c->kg.lp[a]) { c->kg.uniformLp = 0; break; }
    }
    if (batch) {
        // the kernels read the instance grids: same size test per instance
        std::vector<int> seen(c->inst.size(), 0);
        for (auto &gi : c->inst) gi.uniformLp = 1;
        for (int64_t i = 0; i < n; i++) {
            KGrid &gi = c->inst[c->matInstH[ifields[MPI_MAT * nn + i]]];
            int k = (int)(&gi - c->inst.data());
            if (!seen[k]) {
                seen[k] = 1;
                for (int a = 0; a < 3; a++) gi.lp[a] = dfields[(MPF_LP_X + a) * nn + i];
            } else if (gi.uniformLp) {
                for (int a = 0; a < 3; a++)
                    if (dfields[(MPF_LP_X + a) * nn + i] != gi.lp[a]) { gi.uniformLp = 0; break; }
            }
        }
        for (auto &gi : c->inst) gi.matType = c->kg.matType;
        // each instance's material type, for the strain updates' per-type launches
        std::vector<int> types(c->inst.size(), -1);
        for (int64_t i = 0; i < n; i++) {
            const int m = ifields[MPI_MAT * nn + i];
            int &it = types[(size_t)c->matInstH[m]];
            const int t = c->mats[m].type;
            it = it == -1 || it == t ? t : 0;
        }
        if (mpmi::set_instance_types(c, types)) return 1;
        HIPCHECK(c, hipMemcpyAsync(c->dinst, c->inst.data(), c->inst.size() * sizeof(KGrid), hipMemcpyHostToDevice,
                                   c->stream));
    }
    // loads and tractions name particles of the previous upload
    if (c->kld.nl) {
        c->loads.clear();
        c->tractions.clear();
        if (build_particle_bcs(c)) return 1;
    }
    if (mpmi::place_particles(c, n)) return 1;
    HIPCHECK(c, hipMemcpyAsync(c->pbuf, dfields, MPF_NUM_FIELDS * nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->ibuf, ifields, MPI_NUM_FIELDS * nn * sizeof(int), hipMemcpyHostToDevice, c->stream));
    launch_iota(c->ibuf + IROW_HOME * nn, n, 0, c->stream);
    return mpmi::prepare_particles(c);
}

int mpmgpu_set_velocity_bcs(MpmCtx *c, int64_t n, const MpmVelocityBC *bcs) {
    if (n < 0 || n > 0x7fffffffLL) return c->fail("bad velocity condition count");
    if (n > 0 && !c->inst.empty()) return c->fail("velocity conditions are for single simulations, not batches");
    if (c->nextTask != 0) return c->fail("a step is half done through mpmgpu_run_task");
    const long long nnodes = c->kg.nnodes;
    for (int64_t i = 0; i < n; i++) {
        if (bcs[i].node < 0 || bcs[i].node >= nnodes) return c->fail("velocity condition on a node outside the grid");
        if (bcs[i].dir < 1 || bcs[i].dir > c->dim) return c->fail("velocity condition direction must be 1 to the dimension");
        if (bcs[i].style < 1 || bcs[i].style > 4)
            return c->fail("velocity condition styles other than 1 to 4 are not on this path");
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    destroy_graphs(c);                    // captured steps hold the old launches
    if (c->bcbuf) { HIPCHECK(c, hipFree(c->bcbuf)); c->bcbuf = nullptr; }
    c->kbc = KBcs{};
    if (n == 0) return 0;
    // group by node, list order kept within a node
    std::vector<int64_t> ord((size_t)n);
    for (int64_t i = 0; i < n; i++) ord[(size_t)i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return bcs[a].node < bcs[b].node; });
    std::vector<long long> node;
    std::vector<int> start, dir, style;
    std::vector<double> value, ftime;
    for (size_t k = 0; k < ord.size(); k++) {
        const MpmVelocityBC &b = bcs[ord[k]];
        if (node.empty() || node.back() != b.node) {
            node.push_back(b.node);
            start.push_back((int)k);
        }
        dir.push_back(b.dir - 1);
        style.push_back(b.style);
        value.push_back(b.value);
        ftime.push_back(b.ftime);
    }
    start.push_back((int)n);
    const size_t nb = node.size(), ne = (size_t)n, nf = (size_t)c->kg.nfields;
    // one buffer: doubles first (value, ftime, cur, save, time), then node indices, then ints
    if (ensure_time_slots(c)) return 1;
    const size_t nd = 3 * ne + 3 * nb * nf;
    const size_t bytes = nd * sizeof(double) + nb * sizeof(long long) + (nb + 1 + 3 * ne) * sizeof(int);
    HIPCHECK(c, hipMalloc(&c->bcbuf, bytes));
    std::vector<char> h(bytes, 0);
    double *hd = (double *)h.data();
    memcpy(hd, value.data(), ne * sizeof(double));
    memcpy(hd + ne, ftime.data(), ne * sizeof(double));
    long long *hn = (long long *)(hd + nd);
    memcpy(hn, node.data(), nb * sizeof(long long));
    int *hi = (int *)(hn + nb);
    memcpy(hi, start.data(), (nb + 1) * sizeof(int));
    memcpy(hi + nb + 1, dir.data(), ne * sizeof(int));
    memcpy(hi + nb + 1 + ne, style.data(), ne * sizeof(int));
    HIPCHECK(c, hipMemcpy(c->bcbuf, h.data(), bytes, hipMemcpyHostToDevice));
    double *dd = (double *)c->bcbuf;
    long long *dn = (long long *)(dd + nd);
    int *di = (int *)(dn + nb);
    KBcs &k = c->kbc;
    k.nb = (int)nb;
    k.value = dd;
    k.ftime = dd + ne;
    k.cur = dd + 2 * ne;
    k.save = dd + 3 * ne;
    k.time = c->tbuf;
    k.node = dn;
    k.start = di;
    k.dir = di + nb + 1;
    k.style = di + nb + 1 + ne;
    k.act = di + nb + 1 + 2 * ne;
    return put_bc_time(c);
}

}  // extern "C"

// The device buffer of the loads and tractions (KLoads) from the context's two host lists. The
// loaded particles (with a load, a traction or both) are numbered in upload order; each keeps its
// loads, and its tractions, in list order.
static int build_particle_bcs(MpmCtx *c) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    destroy_graphs(c);
    if (c->ldbuf) { HIPCHECK(c, hipFree(c->ldbuf)); c->ldbuf = nullptr; }
    c->kld = KLoads{};
    const size_t ne = c->loads.size(), nte = c->tractions.size();
    if (ne == 0 && nte == 0) return 0;
    const long long np = c->kp.n;
    auto by_particle = [](size_t n, auto part) {
        std::vector<size_t> ord(n);
        for (size_t i = 0; i < n; i++) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return part(a) < part(b); });
        return ord;
    };
    const std::vector<size_t> lo = by_particle(ne, [&](size_t i) { return c->loads[i].particle; });
    const std::vector<size_t> to = by_particle(nte, [&](size_t i) { return c->tractions[i].particle; });
    std::vector<int> map((size_t)np, -1), tof((size_t)np, -1);
    for (const MpmParticleLoad &l : c->loads) map[(size_t)l.particle] = 0;
    for (const MpmParticleTraction &t : c->tractions) map[(size_t)t.particle] = 0;
    int nl = 0, nt = 0;
    for (long long i = 0; i < np; i++)
        if (map[(size_t)i] == 0) map[(size_t)i] = nl++;
    for (size_t k = 0; k < nte; k++) {
        const int64_t pi = c->tractions[to[k]].particle;
        if (tof[(size_t)pi] < 0) tof[(size_t)pi] = nt++;
    }
    // loads: entries grouped by loaded particle
    std::vector<int> start((size_t)nl + 1, 0), dir(ne), style(ne);
    std::vector<double> value(ne), ftime(ne);
    for (size_t k = 0; k < ne; k++) {
        const MpmParticleLoad &l = c->loads[lo[k]];
        start[(size_t)map[(size_t)l.particle] + 1]++;
        dir[k] = l.dir - 1;
        style[k] = l.style;
        value[k] = l.value;
        ftime[k] = l.ftime;
    }
    for (int q = 0; q < nl; q++) start[(size_t)q + 1] += start[(size_t)q];
    // tractions: entries grouped by traction table, tables in upload order of their particles
    std::vector<int> tslot((size_t)nl, -1), tstart((size_t)nt + 1, 0), tdir(nte), tface(nte), tstyle(nte);
    std::vector<double> tvalue(nte), tftime(nte), tthick((size_t)nt, 0.);
    for (size_t k = 0; k < nte; k++) {
        const MpmParticleTraction &t = c->tractions[to[k]];
        const int ts = tof[(size_t)t.particle];
        if (tstart[(size_t)ts + 1]++ == 0) tthick[(size_t)ts] = t.thickness;
        tslot[(size_t)map[(size_t)t.particle]] = ts;
        tdir[k] = t.dir;
        tface[k] = t.face;
        tstyle[k] = t.style;
        tvalue[k] = t.value;
        tftime[k] = t.ftime;
    }
    for (int t = 0; t < nt; t++) tstart[(size_t)t + 1] += tstart[(size_t)t];
    if (ensure_time_slots(c)) return 1;
    // doubles: value, ftime, fext, tvalue, tftime, tthick, ttab; then ints: map, start, dir, style,
    // two tile flag slots, tslot, tstart, tdir, tface, tstyle
    const size_t ntall = (size_t)c->kb.ntall;
    const size_t ntab = (size_t)3 * nt * TRAC_NODES;
    const size_t nd = 2 * ne + 3 * (size_t)nl + 2 * nte + (size_t)nt + ntab;
    const size_t ni = (size_t)np + (size_t)nl + 1 + 2 * ne + 2 * ntall + (size_t)nl + (size_t)nt + 1 + 3 * nte;
    const size_t bytes = nd * sizeof(double) + ni * sizeof(int);
    HIPCHECK(c, hipMalloc(&c->ldbuf, bytes));
    std::vector<char> h(bytes, 0);
    double *hd = (double *)h.data();
    int *hi = (int *)(hd + nd);
    auto putd = [&](size_t &at, const std::vector<double> &v) { memcpy(hd + at, v.data(), v.size() * sizeof(double)); at += v.size(); };
    auto puti = [&](size_t &at, const std::vector<int> &v) { memcpy(hi + at, v.data(), v.size() * sizeof(int)); at += v.size(); };
    size_t ad = 0, ai = 0;
    putd(ad, value); putd(ad, ftime);
    const size_t aFext = ad; ad += 3 * (size_t)nl;
    const size_t aTv = ad; putd(ad, tvalue);
    const size_t aTf = ad; putd(ad, tftime);
    const size_t aTt = ad; putd(ad, tthick);
    const size_t aTab = ad;
    puti(ai, map);
    const size_t aStart = ai; puti(ai, start);
    const size_t aDir = ai; puti(ai, dir);
    const size_t aStyle = ai; puti(ai, style);
    const size_t aFlag = ai; ai += 2 * ntall;          // zero from the host buffer
    const size_t aTslot = ai; puti(ai, tslot);
    const size_t aTstart = ai; puti(ai, tstart);
    const size_t aTdir = ai; puti(ai, tdir);
    const size_t aTface = ai; puti(ai, tface);
    const size_t aTstyle = ai; puti(ai, tstyle);
    HIPCHECK(c, hipMemcpy(c->ldbuf, h.data(), bytes, hipMemcpyHostToDevice));
    double *dd = (double *)c->ldbuf;
    int *di = (int *)(dd + nd);
    KLoads &L = c->kld;
    L.nl = nl;
    L.value = dd;
    L.ftime = dd + ne;
    L.fext = dd + aFext;
    L.map = di;
    L.start = di + aStart;
    L.dir = di + aDir;
    L.style = di + aStyle;
    L.tflag[0] = di + aFlag;
    L.tflag[1] = L.tflag[0] + ntall;
    L.time = c->tbuf;
    L.nt = nt;
    if (nt > 0) {
        L.tslot = di + aTslot;
        L.tstart = di + aTstart;
        L.tdir = di + aTdir;
        L.tface = di + aTface;
        L.tstyle = di + aTstyle;
        L.tvalue = dd + aTv;
        L.tftime = dd + aTf;
        L.tthick = dd + aTt;
        L.ttab = dd + aTab;
    }
    L.terr = c->ks.tracErr;
    return put_bc_time(c);
}

extern "C" {

int mpmgpu_set_particle_loads(MpmCtx *c, int64_t n, const MpmParticleLoad *loads) {
    if (n < 0 || n > 0x7fffffffLL) return c->fail("bad particle load count");
    if (n > 0 && !c->inst.empty()) return c->fail("particle loads are for single simulations, not batches");
    if (c->nextTask != 0) return c->fail("a step is half done through mpmgpu_run_task");
    const long long np = c->kp.n;
    for (int64_t i = 0; i < n; i++) {
        if (loads[i].particle < 0 || loads[i].particle >= np) return c->fail("particle load on a particle that was not uploaded");
        if (loads[i].dir < 1 || loads[i].dir > c->dim) return c->fail("particle load direction must be 1 to the dimension");
        if (loads[i].style < 1 || loads[i].style > 4) return c->fail("load styles other than 1 to 4 are not on this path");
    }
    c->loads.assign(loads, loads + n);
    return build_particle_bcs(c);
}

int mpmgpu_set_particle_tractions(MpmCtx *c, int64_t n, const MpmParticleTraction *tr) {
    if (n < 0 || n > 0x7fffffffLL) return c->fail("bad traction count");
    if (n > 0 && !c->inst.empty()) return c->fail("tractions are for single simulations, not batches");
    if (c->nextTask != 0) return c->fail("a step is half done through mpmgpu_run_task");
    const long long np = c->kp.n;
    for (int64_t i = 0; i < n; i++) {
        const MpmParticleTraction &t = tr[i];
        if (t.particle < 0 || t.particle >= np) return c->fail("traction on a particle that was not uploaded");
        if (c->dim == 3 && !((t.dir >= 1 && t.dir <= 3) || t.dir == 11))
            return c->fail("'dir' in <TractionBC> must be 1-3 or 11 for 3D analyses.");
        if (c->dim == 2 && !(t.dir == 1 || t.dir == 2 || t.dir == 11 || t.dir == 12))
            return c->fail("'dir' in <LoadBC> or <ConcFluxBC> must be 1 or 2 and in <TractionBC> must be 1, 2, 11, or 12 for 2D analyses.");
        if (c->dim == 3 && (t.face < 1 || t.face > 6))
            return c->fail("'face' in <TractionBC>, <ConcFluxBC>, or <HeatFluxBC> element must be 1 to 6 for 3D analyses.");
        if (c->dim == 2 && (t.face < 1 || t.face > 4))
            return c->fail("'face' in <TractionBC>, <ConcFluxBC>, or <HeatFluxBC> element must be 1 to 4 for 2D analyses.");
        if (t.style == 5) return c->fail("Tractions boundary conditions cannot use silent style.");
        if (t.style < 1 || t.style > 4) return c->fail("traction styles other than 1 to 4 are not on this path");
        if (c->dim == 2 && !(t.thickness > 0.)) return c->fail("a 2D traction needs the particle's thickness");
    }
    c->tractions.assign(tr, tr + n);
    return build_particle_bcs(c);
}

int mpmgpu_run_task(MpmCtx *c, int32_t task) {
    if (c->stopped) return check_stop(c);
    if (task != c->nextTask)
        return c->fail("tasks must run in the step order (expected task " + std::to_string(c->nextTask) + ")");
    hipStream_t s = c->stream;
    set_positions(c);
    const bool fmpm = c->run.xpicFMPM != 0;
    switch (task) {
        case MPMGPU_TASK_INITIALIZATION:          // natural coordinates are found in each pass
            xpic_schedule(c, c->steps + 1);
            if (c->kld.nl) launch_particle_loads(c->kg, c->kld, c->parity, s);
            break;
        case MPMGPU_TASK_MASS_MOMENTUM:
            launch_p2g_mass(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, s);
            launch_node_mass(c->dim, c->kg, c->kb, c->kn, s);
            break;
        case MPMGPU_TASK_POST_EXTRAPOLATION:          // node pass of the scatter did the rest
            if (c->kg.nfields > 1) {
                launch_p2g_contact(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, false, s);
                launch_node_contact(c->dim, c->kg, c->kb, c->kn, s);
                launch_contact_normals(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, CONTACT_MASS_MOMENTUM, s);
                launch_contact(c->dim, c->kg, c->kn, CONTACT_MASS_MOMENTUM, s);
            }
            if (c->kbc.nb) launch_vel_bcs(c->kg, c->kn, c->kbc, BC_PASS_MASS_MOMENTUM, c->parity, s);
            break;
        case MPMGPU_TASK_UPDATE_STRAINS_FIRST:
            if (fmpm && c->xpicStep > 1)
                launch_xpic(c->dim, c->kg, c->kp, c->kb, c->kn, c->kbc, c->xpicStep, true, false, c->parity, s);
            launch_usf(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, s);
            break;
        case MPMGPU_TASK_GRID_FORCES: launch_p2g_force(c->dim, c->kg, c->kp, c->kb, c->kn, c->kld, c->parity, s); break;
        case MPMGPU_TASK_POST_FORCES: break;          // restore and update in one node pass
        case MPMGPU_TASK_UPDATE_MOMENTA:
            launch_node_force(c->dim, c->kg, c->kb, c->kn, s);
            if (c->kbc.nb) launch_vel_bcs(c->kg, c->kn, c->kbc, BC_PASS_GRID_FORCES, c->parity, s);
            if (c->kg.nfields > 1) {
                launch_contact_normals(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, CONTACT_UPDATE_MOMENTUM, s);
                launch_contact(c->dim, c->kg, c->kn, CONTACT_UPDATE_MOMENTUM, s);
            }
            if (c->kbc.nb && c->xpicStep <= 1)
                launch_vel_bcs(c->kg, c->kn, c->kbc, BC_PASS_UPDATE_MOMENTUM, c->parity, s);
            break;
        case MPMGPU_TASK_UPDATE_PARTICLES: {
            const int xo = c->xpicStep;
            if (xo > 1) launch_xpic(c->dim, c->kg, c->kp, c->kb, c->kn, c->kbc, xo, fmpm, true, c->parity, s);
            launch_g2p_p2g(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, xo == 0 ? 0 : (fmpm ? xo : -xo), s);
            break;
        }
        case MPMGPU_TASK_UPDATE_STRAINS_LAST:
            launch_node_mom2(c->dim, c->kg, c->kb, c->kn, s);
            if (c->kg.nfields > 1) {
                launch_p2g_contact(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, true, s);
                launch_node_contact(c->dim, c->kg, c->kb, c->kn, s);
                launch_contact_normals(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, CONTACT_UPDATE_STRAINS_LAST, s);
                launch_contact(c->dim, c->kg, c->kn, CONTACT_UPDATE_STRAINS_LAST, s);
            }
            if (c->kbc.nb) launch_vel_bcs(c->kg, c->kn, c->kbc, BC_PASS_UPDATE_STRAINS_LAST, c->parity, s);
            if (fmpm && c->xpicStep > 1)
                launch_xpic(c->dim, c->kg, c->kp, c->kb, c->kn, c->kbc, c->xpicStep, true, false, c->parity, s);
            launch_usl_reset(c->dim, c->kg, c->kp, c->kb, c->kn, c->dmats, c->ks, s);
            break;
        case MPMGPU_TASK_RESET_ELEMENTS:
            launch_bin_finish(c->dim, c->kg, c->kp, c->kb, s);
            c->parity ^= 1;
            c->steps++;
            c->mtime += c->run.timestep;
            c->bcTime += c->run.timestep;
            c->xpicClock += c->run.timestep;
            break;
        default: return c->fail("unknown task id");
    }
    HIPCHECK(c, hipGetLastError());
    c->nextTask = (task + 1) % MPMGPU_NUM_TASKS;
    return 0;
}

int mpmgpu_step(MpmCtx *c, int32_t nsteps) {
    if (c->stopped) return check_stop(c);
    if (c->nextTask != 0) return c->fail("a step is half done through mpmgpu_run_task");
    for (int32_t k = 0; k < nsteps; k++) c->bcTime += c->run.timestep;
    if (c->kp.n == 0) {
        c->steps += nsteps;
        c->mtime += nsteps * c->run.timestep;
        return put_bc_time(c);
    }
    HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
    int done = 0;
    const int G = c->graphSteps;
    if (nsteps >= G && !c->profiling && xpic_graphable(c)) {
        int par = c->parity;
        const int phase = xpic_phase(c, c->steps + 1);
        if (c->graph[par] && c->graphPhase[par] != phase) {
            (void)hipGraphExecDestroy(c->graph[par]);
            c->graph[par] = nullptr;
        }
        if (!c->graph[par]) {
            hipGraph_t gr;
            HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < G; k++) {
                c->xpicStep = xpic_order_at(c, c->steps + 1 + k);
                mpmi::enqueue_step(c);
            }
            HIPCHECK(c, hipStreamEndCapture(c->stream, &gr));
            hipError_t e = hipGraphInstantiate(&c->graph[par], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) return c->fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            c->graphPhase[par] = phase;
            c->parity = par;      // capture does not run the work
        }
        for (; done + G <= nsteps; done += G) {
            HIPCHECK(c, hipGraphLaunch(c->graph[par], c->stream));
            for (int k = 1; k <= G; k++) {      // the schedule's state through the graph's steps
                xpic_schedule(c, c->steps + done + k);
                c->xpicClock += c->run.timestep;
            }
            c->sinceSort += G;
            if (c->sortInterval > 0 && c->sinceSort >= c->sortInterval && mpmi::sort_rows(c)) return 1;
        }
    }
    for (; done < nsteps; done++) {
        xpic_schedule(c, c->steps + done + 1);
        mpmi::enqueue_step(c);
        c->xpicClock += c->run.timestep;
        c->sinceSort++;
        if (c->sortInterval > 0 && c->sinceSort >= c->sortInterval && mpmi::sort_rows(c)) return 1;
        if (c->profiling && (!c->ev.err.empty() || c->ev.live() >= MAX_LIVE_EVENTS)) {
            // the group timings recorded so far are folded in, so the events can be recorded again:
            // a long profiled run holds at most this many events; a failed event call ends the call
            HIPCHECK(c, hipStreamSynchronize(c->stream));
            if (collect_marks(c)) return 1;
        }
    }
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
    c->steps += nsteps;
    c->mtime += nsteps * c->run.timestep;
    return 0;
}

int mpmgpu_set_sort_interval(MpmCtx *c, int32_t steps) {
    if (steps < 0) return c->fail("sort interval must be >= 0");
    c->sortInterval = steps;
    return 0;
}

int mpmgpu_set_graph_steps(MpmCtx *c, int32_t steps) {
    if (steps < 2 || steps > 64 || (steps & 1)) return c->fail("graph steps must be even, 2 to 64");
    destroy_graphs(c);
    c->graphSteps = steps;
    return 0;
}

// Tuner: for each tile shape (TileShape: the cells a tile holds, so the bins, the tile partials
// and the workgroups of every tile kernel), times whole steps captured in graphs of 2 to 32 steps
// and keeps the fastest pair. The particle state, bins and counters are saved before the trials and
// restored before each shape's trials and after them, so tuning leaves the simulation where it was
// (in the chosen shape; its particles are binned again when that differs from the shape before).
int mpmgpu_autotune(MpmCtx *c, int32_t trialSteps, int32_t *graphSteps, double *msPerStep) {
    if (c->nextTask != 0) return c->fail("a step is half done through mpmgpu_run_task");
    if (c->kp.n == 0 || c->profiling) return c->fail("nothing to tune");
    size_t n = (size_t)c->ncap;
    size_t pb = PROWS * n * sizeof(double), ib = IROWS * n * sizeof(int);
    const size_t bb = c->bbufBytes, sb = STATS_WORDS * sizeof(unsigned long long);
    char *save = nullptr;
    HIPCHECK(c, hipMalloc(&save, pb + ib + bb + sb));
    hipError_t e0 = hipMemcpyAsync(save, c->pbuf, pb, hipMemcpyDeviceToDevice, c->stream);
    if (e0 == hipSuccess) e0 = hipMemcpyAsync(save + pb, c->ibuf, ib, hipMemcpyDeviceToDevice, c->stream);
    if (e0 == hipSuccess) e0 = hipMemcpyAsync(save + pb + ib, c->bbuf, bb, hipMemcpyDeviceToDevice, c->stream);
    if (e0 == hipSuccess) e0 = hipMemcpyAsync(save + pb + ib + bb, c->sbuf, sb, hipMemcpyDeviceToDevice, c->stream);
    if (e0 != hipSuccess) {       // nothing has changed yet: the state is where it was
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(save);
        return c->fail(std::string("autotune: saving the state: ") + hipGetErrorString(e0));
    }
    const int parity = c->parity, oldG = c->graphSteps, sinceSort = c->sinceSort, shape0 = c->kg.tileShape;
    const int64_t steps = c->steps, xpicNextStep = c->xpicNextStep;
    const double mtime = c->mtime, bcTime = c->bcTime, xpicNextTime = c->xpicNextTime, xpicClock = c->xpicClock;
    // the state from before the trials in tile shape `shape`: the saved bins in the shape the
    // context had, the particles binned again in the other
    auto restore = [&](int shape) -> int {
        HIPCHECK(c, hipMemcpyAsync(c->pbuf, save, pb, hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(c, hipMemcpyAsync(c->ibuf, save + pb, ib, hipMemcpyDeviceToDevice, c->stream));
        // counters and run stop first: a binning of a stopped run would leave the bins empty
        HIPCHECK(c, hipMemcpyAsync(c->sbuf, save + pb + ib + bb, sb, hipMemcpyDeviceToDevice, c->stream));
        c->parity = parity;
        if (shape == shape0) {
            if (apply_tile_shape(c, shape, false)) return 1;
            HIPCHECK(c, hipMemcpyAsync(c->bbuf, save + pb + ib, bb, hipMemcpyDeviceToDevice, c->stream));
            HIPCHECK(c, mark_all_needed(c));      // the next step clears the nodes the trials touched
        } else {
            // binned from the saved rows; every node is visited by the next step
            if (apply_tile_shape(c, shape, true)) return 1;
        }
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        c->sinceSort = sinceSort;
        c->steps = steps;
        c->mtime = mtime;
        c->bcTime = bcTime;
        c->xpicNextStep = xpicNextStep;
        c->xpicNextTime = xpicNextTime;
        c->xpicClock = xpicClock;
        return put_bc_time(c);
    };
    int best = oldG, bestShape = shape0;
    double bestMs = 1e300;
    static const int cand[] = {2, 4, 8, 16, 32};
    bool failed = false;
    for (int sh = 0; sh < MPM_TILE_SHAPES && !failed; sh++) {
        if (restore(sh)) { failed = true; break; }
        for (int G : cand) {
            int k = trialSteps < G ? G : (trialSteps / G) * G;
            mpmgpu_set_graph_steps(c, G);
            if (mpmgpu_step(c, G) || hipStreamSynchronize(c->stream) != hipSuccess) { failed = true; break; }
            if (mpmgpu_step(c, k) || hipStreamSynchronize(c->stream) != hipSuccess) { failed = true; break; }
            float ms = 0.f;
            hipError_t e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
            if (e != hipSuccess) {
                c->fail(std::string("hipEventElapsedTime: ") + hipGetErrorString(e));
                failed = true;
                break;
            }
            double per = ms / k;
            if (per < bestMs) { bestMs = per; best = G; bestShape = sh; }
        }
    }
    if (failed) { best = oldG; bestShape = shape0; }
    std::string trialErr = failed ? c->err : std::string();
    const int rc = restore(bestShape);
    (void)hipFree(save);
    if (rc) return 1;
    mpmgpu_set_graph_steps(c, best);
    if (failed) return c->fail("autotune trial failed: " + trialErr);
    if (graphSteps) *graphSteps = best;
    if (msPerStep) *msPerStep = bestMs;
    return 0;
}

int mpmgpu_set_tile_shape(MpmCtx *c, int32_t shape) { return mpmi::set_tile_shape(c, shape); }

int32_t mpmgpu_tile_shape(const MpmCtx *c) { return c ? c->kg.tileShape : -1; }

// lCPDI corner faults are fatal, as the reference's exception (MatPoint3D.cpp:610-615); so is a
// contact node seen by more particles than a regression list holds
static int check_cpdi(MpmCtx *c) {
    if (c->kn.cnorm) {
        unsigned long long o = 0;
        HIPCHECK(c, hipMemcpy(&o, c->ks.lrOverflow, sizeof(o), hipMemcpyDeviceToHost));
        if (o) return c->fail("a contact node is seen by more than 1024 particles, more than the regression normals' "
                              "list holds (" + std::to_string(o) + " node passes)");
    }
    if (c->kld.nt) {
        unsigned long long t = 0;
        HIPCHECK(c, hipMemcpy(&t, c->ks.tracErr, sizeof(t), hipMemcpyDeviceToHost));
        if (t) return c->fail("A Traction edge node has left the grid, or an lCPDI traction corner lies beyond the "
                              "particle's neighbouring elements (" + std::to_string(t) + " corners)");
    }
    if (c->run.shape != MPMGPU_SHAPE_LCPDI) return 0;
    unsigned long long e = 0;
    HIPCHECK(c, hipMemcpy(&e, c->ks.cpdiErr, sizeof(e), hipMemcpyDeviceToHost));
    if (e) return c->fail("A CPDI particle domain node has left the grid, or a particle domain reaches "
                          "beyond the neighbouring elements (" + std::to_string(e) + " particle passes)");
    return 0;
}

int mpmgpu_sync(MpmCtx *c) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (check_cpdi(c)) return 1;
    if (check_stop(c)) return 1;
    if (collect_marks(c)) return 1;
    float ms = 0.f;
    // before any timed step (task by task) the events hold nothing; that error is not kept
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) c->lastStepMs = ms;
    else (void)hipGetLastError();
    return 0;
}

int mpmgpu_write_archive(MpmCtx *c, const char *path, const char *mpmOrder, double timeMs, const double *thickness,
                         double defaultThickness) {
    if (!c->inst.empty()) return c->fail("archives are written per simulation, not for a batch");
    const size_t n = (size_t)c->kp.n;
    std::vector<double> d((size_t)MPF_NUM_FIELDS * n);
    std::vector<int32_t> iv((size_t)MPI_NUM_FIELDS * n);
    if (n && mpmgpu_download_particles(c, d.data(), iv.data())) return 1;
    if (mpmgpu_archive_write(path, c->dim, mpmOrder, timeMs, (int64_t)n, d.data(), iv.data(), c->mats.data(),
                             (int32_t)c->mats.size(), thickness, defaultThickness))
        return c->fail(mpmgpu_archive_error());
    return 0;
}

int mpmgpu_download_particles(MpmCtx *c, double *dfields, int32_t *ifields) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (check_cpdi(c)) return 1;
    unsigned long long why = 0;
    if (read_stop(c, &why)) return 1;      // a stopped run's particles are still downloaded
    size_t nn = (size_t)c->kp.n;
    if (nn == 0) return 0;
    set_positions(c);
    // rows back to upload order through the scratch row
    hipStream_t s = c->stream;
    for (int f = 0; f < MPF_NUM_FIELDS; f++) {
        launch_scatter_row(c->kp.f[f], c->scr, c->kb.home, (long long)nn, s);
        HIPCHECK(c, hipMemcpyAsync(dfields + f * nn, c->scr, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    if (ifields) {
        int *iscr = reinterpret_cast<int *>(c->scr);
        for (int q = 0; q < MPI_NUM_FIELDS; q++) {
            launch_scatter_irow(c->ibuf + q * nn, iscr, c->kb.home, (long long)nn, s);
            HIPCHECK(c, hipMemcpyAsync(ifields + q * nn, iscr, nn * sizeof(int), hipMemcpyDeviceToHost, s));
        }
    }
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(s));
    return 0;
}

int mpmgpu_download_contact(MpmCtx *c, double *vol, double *disp, double *grad) {
    if (c->kg.nfields < 2) return c->fail("no contact extrapolations in single-material mode");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    size_t nn = (size_t)c->kg.nstride * c->kg.nfields;     // every instance of a batch
    if (vol) HIPCHECK(c, hipMemcpy(vol, c->kn.cvol, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (disp) HIPCHECK(c, hipMemcpy(disp, c->kn.cdisp, 3 * nn * sizeof(double), hipMemcpyDeviceToHost));
    if (grad) HIPCHECK(c, hipMemcpy(grad, c->kn.cgrad, 3 * nn * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int mpmgpu_download_grid(MpmCtx *c, double *mass, double *pk, double *ftot) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    size_t nn = (size_t)c->kg.nstride * c->kg.nfields;     // every instance of a batch
    if (mass) HIPCHECK(c, hipMemcpy(mass, c->kn.mass, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (pk) HIPCHECK(c, hipMemcpy(pk, c->kn.pk, 3 * nn * sizeof(double), hipMemcpyDeviceToHost));
    if (ftot) HIPCHECK(c, hipMemcpy(ftot, c->kn.ftot, 3 * nn * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int mpmgpu_query(MpmCtx *c, MpmStats *st) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (check_cpdi(c)) return 1;
    unsigned long long h[2] = {0, 0}, why = 0;
    HIPCHECK(c, hipMemcpy(h, c->sbuf, sizeof(h), hipMemcpyDeviceToHost));
    if (read_stop(c, &why)) return 1;
    st->stopped = (int32_t)why;
    st->pad = 0;
    st->steps = c->steps;
    st->nparticles = c->kp.n;
    st->leftGrid = (int64_t)h[0];
    st->nanParticles = (int64_t)h[1];
    st->time = c->mtime;
    st->lastStepMs = c->lastStepMs;
    return 0;
}

int mpmgpu_set_profiling(MpmCtx *c, int32_t on) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    const int r = collect_marks(c);
    c->profiling = on != 0;
    c->ev.reset_totals();
    return r;
}

int mpmgpu_kernel_times(MpmCtx *c, double *totalMs, int64_t *launches) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (collect_marks(c)) return 1;
    for (int g = 0; g < MPMGPU_NUM_KERNEL_GROUPS; g++) {
        if (totalMs) totalMs[g] = c->ev.totalMs[g];
        if (launches) launches[g] = c->ev.launches[g];
    }
    return 0;
}

const char *mpmgpu_kernel_group_name(int32_t g) {
    static const char *names[MPMGPU_NUM_KERNEL_GROUPS] = {
        "k_p2g_mass", "k_node_mass", "k_usf", "k_p2g_force", "k_node_force",
        "k_g2p_p2g", "k_node_mom2", "k_usl_reset", "binning", "contact"};
    return g >= 0 && g < MPMGPU_NUM_KERNEL_GROUPS ? names[g] : "";
}

void mpmgpu_destroy(MpmCtx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->ev.destroy();
    destroy_graphs(c);
    if (c->pbuf) (void)hipFree(c->pbuf);
    if (c->ibuf) (void)hipFree(c->ibuf);
    if (c->scr) (void)hipFree(c->scr);
    if (c->cbuf) (void)hipFree(c->cbuf);
    if (c->dinst) (void)hipFree(c->dinst);
    if (c->dtables) (void)hipFree(c->dtables);
    if (c->dtileType) (void)hipFree(c->dtileType);
    if (c->bbuf) (void)hipFree(c->bbuf);
    if (c->kn.mass) (void)hipFree(c->kn.mass);
    if (c->kn.cnt) (void)hipFree(c->kn.cnt);
    if (c->kn.part) (void)hipFree(c->kn.part);
    if (c->dmats) (void)hipFree(c->dmats);
    if (c->sbuf) (void)hipFree(c->sbuf);
    if (c->bcbuf) (void)hipFree(c->bcbuf);
    if (c->ldbuf) (void)hipFree(c->ldbuf);
    if (c->tbuf) (void)hipFree(c->tbuf);
    if (c->lrbuf) (void)hipFree(c->lrbuf);
    if (c->lrlist) (void)hipFree(c->lrlist);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
