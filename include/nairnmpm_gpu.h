/*
 * nairnmpm_gpu.h -- C ABI of the MI355X explicit MPM time step.
 *
 * This is the drop-in boundary for NairnMPM's explicit step. The reference runs the step as a
 * linked list of MPMTask objects whose `virtual bool Execute(int)` works on process globals
 * (NairnMPM/src/NairnMPM_Class/MPMTask.hpp:20-47, Common/System/CommonTask.hpp:26,
 * NairnMPM/src/NairnMPM_Class/NairnMPM.cpp:284-335). Every entry point below replaces one piece
 * of that interface and takes plain pointers and sizes only: no C++ types, no exceptions.
 *
 * Status convention: 0 = success; non-zero = failure, with a message from mpmgpu_last_error()
 * (per context) or mpmgpu_global_error() (calls that have no context yet). This replaces the
 * reference's CommonException throw/catch (NairnMPM.cpp:202-225).
 *
 * Units are the reference's Legacy units after input scaling (mm, g, s; stress in Pa, stored
 * "specific", i.e. divided by rho0, as NairnMPM does).
 */
#ifndef NAIRNMPM_GPU_H
#define NAIRNMPM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* analysis types: same numbers as the reference (NairnMPM/src/System/MPMPrefix.hpp:114-115) */
enum { MPMGPU_PLANE_STRAIN = 10, MPMGPU_PLANE_STRESS = 11, MPMGPU_THREED = 12 };

/* shape functions: same numbers as ElementBase::useGimp (MPMPrefix.hpp:127-137) */
enum { MPMGPU_SHAPE_LINEAR = 0, MPMGPU_SHAPE_UGIMP = 1, MPMGPU_SHAPE_LCPDI = 10 };

/* material types: the reference's XML Type= numbers (Read_MPM/MaterialController.cpp:105-190) */
enum {
    MPMGPU_MAT_ISOTROPIC = 1,
    MPMGPU_MAT_MOONEY = 8,
    MPMGPU_MAT_ISOPLASTICITY = 9,
    MPMGPU_MAT_NEOHOOKEAN = 28,
    MPMGPU_MAT_CONTACT = 60,
    MPMGPU_MAT_COULOMB = 61
};

/* per-particle FP64 fields of the structure-of-arrays layout; a host buffer holds
   MPF_NUM_FIELDS consecutive arrays of n doubles. Tensor order is the reference's
   Tensor (xx,yy,zz,yz,xz,xy) and TensorAntisym (yz,xz,xy) (Common/System/DataTypes.hpp:41-56). */
enum {
    MPF_POS_X = 0, MPF_POS_Y, MPF_POS_Z,
    MPF_VEL_X, MPF_VEL_Y, MPF_VEL_Z,
    MPF_MASS,
    MPF_LP_X, MPF_LP_Y, MPF_LP_Z,                    /* dimensionless half size (mpm_lp) */
    MPF_EP_XX, MPF_EP_YY, MPF_EP_ZZ, MPF_EP_YZ, MPF_EP_XZ, MPF_EP_XY,   /* ep: encodes F */
    MPF_WR_YZ, MPF_WR_XZ, MPF_WR_XY,                 /* wrot: encodes F */
    MPF_SP_XX, MPF_SP_YY, MPF_SP_ZZ, MPF_SP_YZ, MPF_SP_XZ, MPF_SP_XY,   /* specific stress */
    MPF_PRESSURE,
    MPF_ALT_XX, MPF_ALT_YY, MPF_ALT_ZZ, MPF_ALT_YZ, MPF_ALT_XZ, MPF_ALT_XY, /* eplast or B */
    MPF_HIST0, MPF_HIST1,                            /* J2: alpha; hyperelastic: J, Jres */
    MPF_ORIG_X, MPF_ORIG_Y, MPF_ORIG_Z,              /* origpos */
    MPF_NUM_FIELDS
};

/* per-particle int32 fields */
enum {
    MPI_ELEM = 0,      /* zero-based element (reference inElem-1) */
    MPI_MAT,           /* zero-based material (reference matnum-1) */
    MPI_ID,            /* creation index (reference mpm[] order) */
    MPI_LEFT,          /* 1 once the particle has left the grid (HasLeftTheGridBefore) */
    MPI_NUM_FIELDS
};

/* Regular structured grid, after the one-cell GIMP pad per side
   (Read_MPM/Generators.cpp:1760-1833). Node i on x is at xmin + i*dx. */
typedef struct MpmGridDesc {
    int32_t analysis;          /* MPMGPU_PLANE_STRAIN or MPMGPU_THREED */
    int32_t nx, ny, nz;        /* cells per axis (nz = 0 in 2D) */
    double xmin, ymin, zmin;   /* padded origin (mm) */
    double dx, dy, dz;         /* (max-min)/n per axis (mm) */
    double thickness;          /* 2D out-of-plane thickness (mm) */
} MpmGridDesc;

#define MPMGPU_MAX_MATERIALS 16
#define MPMGPU_MAX_XPIC_ORDER 32

/* Material constants after VerifyAndLoadProperties, in the reference's specific units. */
typedef struct MpmMaterial {
    int32_t type;              /* MPMGPU_MAT_* */
    int32_t field;             /* velocity field (multimaterial mode), 0 otherwise */
    double rho;                /* g/mm^3 */
    double E, nu, G, K;        /* Pa (as entered, after unit scaling and completion) */
    double C[6][6];            /* ElasticProperties::C, specific (Common/Materials/Elastic.cpp:43-150) */
    double Gred, Kred;         /* IsoPlasticity reduced moduli (IsoPlasticity.cpp:60-66) */
    double yield, Ep;          /* Pa */
    double yldred, yldredMin, Epred, alphaMax;   /* LinearHardening (LinearHardening.cpp:56-76) */
    double Gsp, Lamesp, Ksp;   /* Neohookean specific (Neohookean.cpp:135-138) */
    double G1sp, G2sp;         /* Mooney specific */
    double waveSpeed;          /* mm/s, used for the CFL time step */
    double friction;           /* CoulombFriction frictionCoeff (contact laws only; <0 stick) */
    double frictionStatic;     /* CoulombFriction frictionCoeffStatic (ignored if <= 0) */
    int32_t ujOption;          /* HyperElastic UofJOption */
    int32_t usePDamping;       /* <PDamping> in <Material>: the material's particles take pdamping in
                                  place of the run's particle damping (MaterialBase::GetMaterialDamping,
                                  MaterialBaseMPM.cpp:731-744; UpdateParticlesTask.cpp:118) */
    double pdamping;           /* 1/s */
} MpmMaterial;

/* Step parameters (NairnMPM.hpp:23 globals timestep, strainTimestepFirst/Last) */
typedef struct MpmRunParams {
    double timestep;           /* s */
    double strainFirst;        /* USAVG first half: fractionUSF*timestep */
    double strainLast;         /* timestep - strainFirst */
    int32_t shape;             /* MPMGPU_SHAPE_* */
    int32_t mpmMethod;         /* 2 = USAVG (only method of the hot path) */
    int32_t multiMaterial;     /* 0 single-material mode */
    int32_t contactLaw;        /* material index of the default contact law, -1 none */
    int32_t normalsMethod;     /* MultiMaterialMode Normals= (0 MAXG, 1 MAXV, 2 AVGG, 3 OWNG,
                                  4 specified, 5 linear regression, 6 logistic regression, the
                                  reference's default; MeshInfo.hpp:31-32) */
    int32_t deleteLeaving;     /* fmobj->deleteLeavingParticles: a particle that leaves the grid
                                  would be deleted into the reservoir; not supported, so the
                                  first one stops the run with an error */
    int32_t maxLeaving;        /* pushed-back particles after which the run stops with "Too many
                                  particles have left the grid" (0 = never); the XML reader
                                  resolves <LeaveLimit> into these two as NairnMPM::CreateWarnings
                                  does (NairnMPM.cpp:814-834). Batches ignore both. */
    int32_t numFields;         /* material velocity fields (maxMaterialFields); 1 unless multimaterial */
    double gridDamping;        /* <Damping>, 1/s: BodyForce::GetGridDamping (BodyForce.cpp:167-184) */
    double particleDamping;    /* <PDamping>, 1/s: BodyForce::GetParticleDamping (:187-204) */
    double gravity[3];         /* <Gravity> BodyX/Y/ZForce, mm/s^2: BodyForce gforce (BodyForce.cpp:76-95) */
    double contactNormal[3];   /* normals method 4: the unit normal from Polar/Azimuth
                                  (MeshInfo::SetContactNormal, MeshInfo.cpp:1211-1227) */
    /* The PeriodicXPIC custom task (Custom_Tasks/PeriodicXPIC.cpp): in the steps it selects the
       particle update is XPIC(k) (xpicFMPM = 0: XPICExtrapolationTask.cpp, MatPoint3D.cpp:152-170)
       or FMPM(k) (xpicFMPM = 1, which also takes the strain updates' grid velocities from the
       extrapolation, UpdateStrainsFirstTask.cpp:106-114), k = xpicOrder; 0 = FLIP only. The task
       selects every xpicPeriodSteps-th step when that is > 0, else steps by time, one each
       xpicPeriodTime seconds (periodicCFL already multiplied by the CFL time step), as
       PeriodicXPIC::PrepareForStep does (PeriodicXPIC.cpp:283-312). Single material; not in
       batches, the solve server or mpmgpu_solve_static; order at most MPMGPU_MAX_XPIC_ORDER. */
    int32_t xpicOrder;
    int32_t xpicFMPM;
    int32_t xpicPeriodSteps;
    int32_t xpicReserved;      /* 0 */
    double xpicPeriodTime;
} MpmRunParams;

/* Counters reported back from the device (ResetElementsTask.cpp:71-151 warnings) */
typedef struct MpmStats {
    int64_t steps;             /* steps run so far */
    int64_t nparticles;
    int64_t leftGrid;          /* particles that left the grid (first time) */
    int64_t nanParticles;      /* particles with a NaN position */
    double time;               /* simulation time (s) */
    double lastStepMs;         /* device time of the last timed step (ms) */
    int32_t stopped;           /* why the run stopped (ResetElementsTask's exceptions,
                                  ResetElementsTask.cpp:71-151): 0 running, 1 too many particles
                                  left the grid, 2 a NaN position, 3 a leaving particle would be
                                  deleted; mpmgpu_sync and mpmgpu_step then fail with the message,
                                  mpmgpu_download_particles gives the state the stopping step
                                  left, and a new upload starts again */
    int32_t pad;
} MpmStats;

/* Task ids for mpmgpu_run_task: one per hot-path MPMTask (NairnMPM.cpp:952-1108) */
enum {
    MPMGPU_TASK_INITIALIZATION = 0,     /* InitializationTask.cpp:43 */
    MPMGPU_TASK_MASS_MOMENTUM,          /* MassAndMomentumTask.cpp:48 */
    MPMGPU_TASK_POST_EXTRAPOLATION,     /* PostExtrapolationTask.cpp:46 */
    MPMGPU_TASK_UPDATE_STRAINS_FIRST,   /* UpdateStrainsFirstTask.cpp:52 */
    MPMGPU_TASK_GRID_FORCES,            /* GridForcesTask.cpp:40 */
    MPMGPU_TASK_POST_FORCES,            /* PostForcesTask.cpp:43 */
    MPMGPU_TASK_UPDATE_MOMENTA,         /* UpdateMomentaTask.cpp:44 */
    MPMGPU_TASK_UPDATE_PARTICLES,       /* UpdateParticlesTask.cpp:44 */
    MPMGPU_TASK_UPDATE_STRAINS_LAST,    /* UpdateStrainsLastContactTask.cpp:60 */
    MPMGPU_TASK_RESET_ELEMENTS,         /* ResetElementsTask.cpp:44 */
    MPMGPU_NUM_TASKS
};

/* kernel groups of one step, for mpmgpu_kernel_times */
enum {
    MPMGPU_KG_P2G_MASS = 0,   /* MassAndMomentumTask scatter */
    MPMGPU_KG_NODE_MASS,      /* PostExtrapolationTask + grid velocity */
    MPMGPU_KG_USF,            /* UpdateStrainsFirstTask */
    MPMGPU_KG_P2G_FORCE,      /* GridForcesTask */
    MPMGPU_KG_NODE_FORCE,     /* PostForcesTask + UpdateMomentaTask */
    MPMGPU_KG_G2P_P2G,        /* UpdateParticlesTask + USL momentum re-extrapolation */
    MPMGPU_KG_NODE_MOM2,      /* USL grid velocity */
    MPMGPU_KG_USL_RESET,      /* USL strain update + ResetElementsTask */
    MPMGPU_KG_BINNING,        /* cell sort of the particles for the next step */
    MPMGPU_KG_CONTACT,        /* multimaterial contact extrapolations and contact passes */
    MPMGPU_NUM_KERNEL_GROUPS
};

typedef struct MpmProblem MpmProblem;   /* host-side problem (input + generated particles) */
typedef struct MpmCtx MpmCtx;           /* device context: all device buffers of one simulation */

/* ---- input boundary: the reference's XML input (MPMReadHandler) ---------------------- */

/* Build a problem in code, element by element, as the XML reader does:
   <Header><Analysis> (MPMGPU_PLANE_STRAIN or MPMGPU_THREED) ...
   Replaces MPMReadHandler::myStartElement for the hot-path elements
   (Read_MPM/MPMReadHandler.cpp:83-700). */
int mpmgpu_problem_create(int32_t analysis, MpmProblem **out);
/* <Grid xmin= xmax= ymin= ymax= zmin= zmax=> with <Horiz cellsize=> etc.; cell3[a] <= 0 means
   the count n3[a] is used instead. Pads one cell per side (Read_MPM/Generators.cpp:1675-1833). */
int mpmgpu_problem_set_grid(MpmProblem *prob, const double *ext6, const double *cell3,
                            const int32_t *n3, double thickness);
/* <MPMHeader> settings: "MatlPtsPerElement", "CFL", "TimeStep" (s), "MaxTime" (s),
   "GIMP" (MPMGPU_SHAPE_*), "LeaveLimit" */
int mpmgpu_problem_set(MpmProblem *prob, const char *key, double value);
/* <Material Type= Name=> (Read_MPM/MaterialController.cpp:105-190); id is 1-based */
int mpmgpu_problem_add_material(MpmProblem *prob, int32_t type, const char *name, int32_t *id);
/* one property element of a material (<E>, <nu>, <G>, <K>, <rho>, <yield>, <Ep>, ...), XML units */
int mpmgpu_problem_material_property(MpmProblem *prob, int32_t id, const char *prop, double value);
/* <Body mat= vx= vy= vz= thick=> holding one shape: "Box", "Cylinder", "Rect", "Oval", "Sphere"
   (MPMReadHandler::MPMPts, Read_MPM/Generators.cpp:1423-1545) */
int mpmgpu_problem_add_body(MpmProblem *prob, int32_t mat, const char *shape, const double *ext6,
                            int32_t axis, const double *vel3, double thick);
/* explicit <mp matl= lpx= lpy= lpz=><pt/><vel/><mass/></mp> (MPMReadHandler.cpp:1159-1290);
   mass < 0 means rho*volume */
int mpmgpu_problem_add_point(MpmProblem *prob, int32_t mat, const double *pos3, const double *vel3,
                             const double *lp3, double thick, double mass);

/* Parse an XML input file's text (the NairnMPM.dtd subset of the hot path) and generate the
   grid and particles, then do PreliminaryParticleCalcs and CFLTimeStep.
   Replaces CommonAnalysis::ReadFile (Common/System/CommonAnalysis.cpp:184-273) +
   MPMReadHandler::grid/MPMPts (Read_MPM/Generators.cpp:1423,1675) +
   NairnMPM::PreliminaryParticleCalcs/CFLTimeStep (NairnMPM.cpp:589-730,1207-1240). */
int mpmgpu_problem_from_xml(const char *xml_text, MpmProblem **out);
/* Grid, materials and step parameters of a problem */
int mpmgpu_problem_describe(const MpmProblem *prob, MpmGridDesc *grid, MpmMaterial *mats,
                            int32_t *nmat, MpmRunParams *run, int64_t *nparticles,
                            double *maxTime);
/* Initial particle state in mpm[] order: dfields is MPF_NUM_FIELDS*n doubles, ifields is
   MPI_NUM_FIELDS*n int32 */
int mpmgpu_problem_particles(const MpmProblem *prob, double *dfields, int32_t *ifields);
/* Archive settings of a problem: the MPM archive order of <MPMArchiveOrder> and <MPMArchive>
   (up to 25 characters and a NUL; "iY" when not given), <ArchiveTime> in s (< 0 when not given)
   and each particle's thickness (2D). Replaces ArchiveData's XML settings
   (Read_MPM/MPMReadHandler.cpp:227-240,654-800). */
int mpmgpu_problem_archive(const MpmProblem *prob, char *order, double *archiveTime, double *thickness);

/* One grid velocity condition on one node (NodalVelBC, Boundary_Conditions/NodalVelBC.hpp:17-72):
   node index (0-based, x fastest), axis dir (1 x, 2 y, 3 z), style (BoundaryCondition styles:
   1 constant, 2 linear, 3 sine, 4 cosine), value in mm/s (mm/s^2 for linear) and ftime in s
   (styles 1-2: start time) or 1/s (styles 3-4: angular frequency). The value at step start time
   t is value (t >= ftime), value*(t-ftime) (t >= ftime), value*sin(t*ftime) or
   value*cos(t*ftime); styles 1-2 leave the node free before ftime
   (BoundaryCondition::BCValue and GetNodeNum, BoundaryCondition.cpp:72-140). */
typedef struct MpmVelocityBC {
    int64_t node;
    int32_t dir;
    int32_t style;
    double value;
    double ftime;
} MpmVelocityBC;
/* The velocity conditions of <GridBCs> (<BCLine> in 2D, <BCBox> in 3D, each holding <DisBC>),
   in the reference's list order: BC by BC in document order, nodes in node order
   (MPMReadHandler BCLine/BCBox/DisBC, Read_MPM/Generators.cpp:831-1010). Returns the count in
   *n; copies min(*n on entry, count) entries when bcs is not null. */
int mpmgpu_problem_velocity_bcs(const MpmProblem *prob, MpmVelocityBC *bcs, int64_t *n);
/* One external load on one particle (MatPtLoadBC, Boundary_Conditions/MatPtLoadBC.hpp): particle
   index in mpm[] order (0-based), axis dir (1 x, 2 y, 3 z), style 1-4 as MpmVelocityBC, value in
   internal force units (g mm/s^2 = uN; uN/s for linear; a <net/> load already divided by the
   number of particles) and ftime as MpmVelocityBC. Each step the particle's pFext is zeroed, then
   every load adds its value at the step's start time (MatPtLoadBC::SetParticleFext,
   MatPtLoadBC.cpp:210-221), and the force pass adds N * pFext to the nodes
   (MatPoint3D::GetFintPlusFext, MatPoint3D.cpp:248-252). */
typedef struct MpmParticleLoad {
    int64_t particle;
    int32_t dir;
    int32_t style;
    double value;
    double ftime;
} MpmParticleLoad;
/* The <LoadBC> loads of <ParticleBCs> (<BCLine> in 2D, <BCBox> in 3D; <net/> and <perParticle/>),
   in list order: load by load in document order, particles in mpm[] order
   (Read_MPM/Generators.cpp:1012-1023,1105-1188; MatPtLoadBC::SetBCValue, MatPtLoadBC.cpp:202-206).
   Count and copy as mpmgpu_problem_velocity_bcs. */
int mpmgpu_problem_particle_loads(const MpmProblem *prob, MpmParticleLoad *loads, int64_t *n);
/* One traction on one face of one particle (MatPtTractionBC, Boundary_Conditions/MatPtTractionBC.hpp):
   particle index in mpm[] order (0-based); dir 1 x, 2 y, 3 z (3D), 11 the face's outward normal,
   12 the face's first edge direction (2D); face 1-4 (2D: bottom, right, top, left) or 1-6 (3D:
   -y, +x, +y, -x, -z, +z); style 1-4 as MpmVelocityBC; value in internal stress units
   (g/(mm s^2) = Pa, Pa/s for linear); ftime as MpmVelocityBC; thickness the particle's thickness
   (MatPoint2D::thickness, used in 2D only). Each step, after the grid forces and RestoreMomenta
   and before gravity, every traction adds its value at the step's start time times the face
   weight of GetSurfaceInfo (MatPoint2D.cpp:1138-1300, MatPoint3D.cpp:628-851: hat n times half
   the face length times the thickness in 2D, a quarter of the face area in 3D) to the nodes of
   the elements holding the face's corners, weighted by the elements' linear shape functions at
   the corners, on the particle's material velocity field where that field has points
   (MatPtTractionBC::AddMPFluxBC, MatPtTractionBC.cpp:64-227, without cracks or ExactTractions).
   uGIMP and classic particles take the corners of the undeformed particle and the area of the
   deformed face; lCPDI particles the deformed corners. */
typedef struct MpmParticleTraction {
    int64_t particle;
    int32_t dir;
    int32_t face;
    int32_t style;
    int32_t reserved;            /* 0 */
    double value;
    double ftime;
    double thickness;
} MpmParticleTraction;
/* The <TractionBC> conditions of <ParticleBCs> (<BCLine> in 2D, <BCBox> in 3D), in list order:
   condition by condition in document order, particles in mpm[] order (Read_MPM/Generators.cpp:
   1105-1202; MatPtLoadBC::SetBCValue, MatPtLoadBC.cpp:202-206: MPa to Pa). Count and copy as
   mpmgpu_problem_velocity_bcs. */
int mpmgpu_problem_particle_tractions(const MpmProblem *prob, MpmParticleTraction *tractions, int64_t *n);
void mpmgpu_problem_destroy(MpmProblem *prob);

/* ---- archive files (ArchiveData, System/ArchiveData.cpp), host only ----------------- */
/* Normalises an MPM archive order as ArchiveData::CalcArchiveSize does (ArchiveData.cpp:266-283)
   into out[26]; fails for items this implementation does not carry (energies, temperature,
   concentration, element crossings, damage normal, spin, histories past the fourth, size). */
int mpmgpu_archive_order(const char *mpmOrder, char *out);
/* Record bytes of an order (CalcArchiveSize, :266-424); -1 if the order is not carried */
int32_t mpmgpu_archive_record_size(const char *mpmOrder, int32_t dim);
/* Writes a version-6 archive (header SetArchiveHeader :443-470, records ArchiveResults :721-1000)
   of n particles given in the MPF_ and MPI_ row layout in mpm[] order; little endian. */
int mpmgpu_archive_write(const char *path, int32_t dim, const char *mpmOrder, double timeMs, int64_t n,
                         const double *dfields, const int32_t *ifields, const MpmMaterial *mats,
                         int32_t nmat, const double *thickness, double defaultThickness);
/* Message of the last failed archive call */
const char *mpmgpu_archive_error(void);

/* ---- device context (NairnMPM::CMPreparations/CMAnalysis state, NairnMPM.cpp:135-280) -- */

/* Create a device context for one simulation on the current HIP device. */
int mpmgpu_create(const MpmGridDesc *grid, const MpmMaterial *mats, int32_t nmat,
                  const MpmRunParams *run, MpmCtx **out);
/* ---- batches of independent simulations ----------------------------------------------- */
/* No counterpart in the reference (one analysis per process): many independent instances, each
   with its own grid, time step, materials and particles, advance together in one context. Every
   step launches each kernel once for all instances. Instance results equal those of the same
   instance run alone, bit for bit (each keeps its own coordinates and sums in the same order). */
#define MPMGPU_MAX_BATCH_MATERIALS 65536
typedef struct MpmInstanceDesc {
    MpmGridDesc grid;          /* the instance's own grid (all instances: same analysis type) */
    MpmRunParams run;          /* its time step; same shape functions for all; no multimaterial */
    int32_t firstMaterial;     /* its materials: mats[firstMaterial .. +numMaterials) */
    int32_t numMaterials;
} MpmInstanceDesc;
int mpmgpu_batch_create(int32_t ninst, const MpmInstanceDesc *inst, const MpmMaterial *mats, int32_t nmat,
                        MpmCtx **ctx);
/* A batch takes its particles through mpmgpu_upload_particles: MPI_MAT holds batch material
   indices (which name the instance), MPI_ELEM the element in the instance's own grid, and
   positions are in the instance's coordinates. */

/* Upload particles (layout as mpmgpu_problem_particles); replaces the mpm[] globals
   (MPM_Classes/MPMBase.hpp:269-270). */
int mpmgpu_upload_particles(MpmCtx *ctx, int64_t n, const double *dfields, const int32_t *ifields);
/* Grid velocity conditions of a single simulation (not a batch); n = 0 removes them. Each step
   imposes them as NodalVelBC::GridVelocityConditions does (NodalVelBC.cpp:321-400): after the
   mass-and-momentum contact (the strain update's velocities; the force pass starts again from the
   momenta before it, MatVelocityField::RestoreMomenta), on the forces before the momentum update,
   after the momentum update and after the second extrapolation, each pass zeroing the controlled
   component of every field with points, then adding mass * value. Values are taken at the step's
   start time, counted from the time of this call's step count (mpmgpu_query's time). */
int mpmgpu_set_velocity_bcs(MpmCtx *ctx, int64_t n, const MpmVelocityBC *bcs);
/* External particle loads of a single simulation (not a batch); n = 0 removes them. Particle
   indices are upload indices (the order of mpmgpu_upload_particles); upload the particles first.
   Values are taken at the step's start time, as for the velocity conditions. */
int mpmgpu_set_particle_loads(MpmCtx *ctx, int64_t n, const MpmParticleLoad *loads);
/* Traction conditions of a single simulation (not a batch); n = 0 removes them. Particle indices
   are upload indices; values are taken at the step's start time, as for the loads. A corner that
   leaves the grid stops the run with the reference's error ("A Traction edge node has left the
   grid"); so does an lCPDI corner beyond the particle's neighbour cells, which this
   implementation does not reach. */
int mpmgpu_set_particle_tractions(MpmCtx *ctx, int64_t n, const MpmParticleTraction *tractions);
/* Run one hot-path MPMTask's Execute() on the device (MPMGPU_TASK_*), in reference order. */
int mpmgpu_run_task(MpmCtx *ctx, int32_t task);
/* Run nsteps whole steps (NairnMPM::MPMStep, NairnMPM.cpp:284-335), fused and graph-replayed. */
int mpmgpu_step(MpmCtx *ctx, int32_t nsteps);
/* Steps captured per graph launch (even, 2..64; default 2). No counterpart in the reference. */
int mpmgpu_set_graph_steps(MpmCtx *ctx, int32_t steps);
/* Steps between sorts of the particle rows into tile order in device memory (default 50; 0 =
   only at upload). A memory-layout choice with no counterpart in the reference: results are the
   same for any interval, since particles inside a cell are summed in upload order. */
int mpmgpu_set_sort_interval(MpmCtx *ctx, int32_t steps);
/* Tuner (no counterpart in the reference): for each tile shape (the cells one bin tile holds:
   shape 0 is 4x4x4 cells in 3D and 8x8 in 2D, shape 1 is 8x4x4 and 16x8) times graph lengths
   2..32 over trialSteps steps each, keeps the fastest pair, and restores the simulation state it
   started from in the chosen shape (mpmgpu_tile_shape reports it). Results of one shape are
   deterministic; the two shapes sum grid values in different orders, so they agree to rounding. */
int mpmgpu_autotune(MpmCtx *ctx, int32_t trialSteps, int32_t *graphSteps, double *msPerStep);
/* Set the tile shape (0 or 1, as above; default 0) between whole steps: the particles are binned
   again. No counterpart in the reference. */
int mpmgpu_set_tile_shape(MpmCtx *ctx, int32_t shape);
/* The context's tile shape (-1 for a null context). */
int32_t mpmgpu_tile_shape(const MpmCtx *ctx);
/* Wait for all queued device work of this context. */
int mpmgpu_sync(MpmCtx *ctx);
/* Download particles in creation (mpm[]) order. */
int mpmgpu_download_particles(MpmCtx *ctx, double *dfields, int32_t *ifields);
/* Archive of the context's particles at timeMs (ms), NairnMPM's archive file format version 6
   (ArchiveData::ArchiveResults, System/ArchiveData.cpp:721-1000); thickness per particle (2D) or
   null for defaultThickness. Not for batches. */
int mpmgpu_write_archive(MpmCtx *ctx, const char *path, const char *mpmOrder, double timeMs,
                         const double *thickness, double defaultThickness);
/* Counters and timings (MPMTask::TrackTimes, MPMTask.cpp:51-72; warnings). */
int mpmgpu_query(MpmCtx *ctx, MpmStats *stats);
/* Grid node values (mass, momentum, force) after the last task, nodes numbered x fastest; one
   block per material velocity field (numFields of them, field-major): mass[f*nn + n],
   pk/ftot[(3*f + c)*nn + n]. In a batch context nn covers every instance: instance i's nodes
   start at the sum over earlier instances of their node counts rounded up to 256. */
int mpmgpu_download_grid(MpmCtx *ctx, double *mass, double *pk, double *ftot);
/* Multimaterial contact extrapolations of every field (same layout): volume, mass-weighted
   displacement, volume gradient (MatVelocityField::contactInfo, MatVelocityField.hpp:40-124).
   Fails in single-material mode. */
int mpmgpu_download_contact(MpmCtx *ctx, double *volume, double *disp, double *grad);
/* Per-kernel-group HIP event timing on the context's stream (MPMTask::TrackTimes,
   MPMTask.cpp:51-72). Turning it on or off clears the totals; mpmgpu_step then launches
   eagerly (no graph replay) with an event between kernel groups. */
int mpmgpu_set_profiling(MpmCtx *ctx, int32_t on);
/* totals (ms) and launch counts per MPMGPU_KG_* group since profiling was turned on */
int mpmgpu_kernel_times(MpmCtx *ctx, double *totalMs, int64_t *launches);
const char *mpmgpu_kernel_group_name(int32_t group);
const char *mpmgpu_last_error(const MpmCtx *ctx);
const char *mpmgpu_global_error(void);
void mpmgpu_destroy(MpmCtx *ctx);

/* ---- a solve that allocates nothing ---------------------------------------------------- */
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): advances n particles
   held in the caller's device memory by nsteps explicit steps, as mpmgpu_create +
   mpmgpu_upload_particles + mpmgpu_step + mpmgpu_download_particles would, and returns with the
   stream waited for. dfields/ifields are device pointers to the row layout of
   mpmgpu_problem_particles (MPF_NUM_FIELDS rows of n doubles, MPI_NUM_FIELDS rows of n int32),
   read as the initial state and overwritten with the final one. The grid, tile partials, bins,
   counters and second position buffer come from one static device array of
   MPMGPU_STATIC_SCRATCH_BYTES per device, held until the process ends; a solve needing more is
   refused with a message giving its size. No velocity conditions or particle loads; no row
   sorting (results do not depend on it). stream: a hipStream_t, or null for the null stream.
   Every call on a device shares its array: calls that overlap on one device race. Errors come
   from mpmgpu_global_error(); a run stopped on the device (ResetElementsTask's exceptions)
   returns its message with the stopping step's state in the rows. */
#define MPMGPU_STATIC_SCRATCH_BYTES (256ull << 20)
int mpmgpu_solve_static(const MpmGridDesc *grid, const MpmMaterial *mats, int32_t nmat, const MpmRunParams *run,
                        int64_t n, double *dfields, int32_t *ifields, int32_t nsteps, void *stream);

/* ---- solve server: continuous batching of independent solves -------------------------- */
/* No counterpart in the reference (NairnMPM runs one analysis per process). A request is a whole
   solve: grid, materials, run parameters, initial particles and a number of steps. Requests wait
   in arrival order until their device footprint fits the server's memory budget; admitted
   requests stay resident on the device and advance together as one batch (one step of every
   resident solve per batch step, one launch per kernel). When solves finish or new ones are
   admitted, the batch is laid out again and the resident particle state moves to the new layout
   on the device. Every result equals the same solve run alone (mpmgpu_create + mpmgpu_step) bit
   for bit. All requests of a server share its analysis type and shape functions and are
   single-material-mode problems. */
typedef struct MpmServer MpmServer;

typedef struct MpmServerParams {
    int64_t memoryBudget;      /* device bytes for resident solves, moves between layouts,
                                  uploads staged on the device and results not yet collected
                                  (0: 90% of the free device memory at creation); a solve is
                                  admitted alone when nothing is resident even if staged rows of
                                  others would exceed it */
    int32_t maxResident;       /* most solves resident at once (0: no limit) */
    int32_t graphSteps;        /* batch steps per captured graph and per admission check (even,
                                  2..64; 0: 8) */
    int32_t analysis;          /* MPMGPU_PLANE_STRAIN or MPMGPU_THREED */
    int32_t shape;             /* MPMGPU_SHAPE_UGIMP or MPMGPU_SHAPE_LCPDI */
} MpmServerParams;

enum { MPMGPU_REQ_QUEUED = 0, MPMGPU_REQ_RESIDENT = 1, MPMGPU_REQ_DONE = 2, MPMGPU_REQ_COLLECTED = 3 };

typedef struct MpmRequestInfo {
    int32_t state;             /* MPMGPU_REQ_* */
    int32_t stepsTotal;
    int32_t stepsDone;
    int32_t pad;
    int64_t nparticles;
    int64_t footprint;         /* device bytes while resident (mpmgpu_footprint) */
    double submitted;          /* seconds on the server clock (mpmgpu_server_clock) */
    double admitted;           /* -1 until admitted */
    double finished;           /* -1 until done */
} MpmRequestInfo;

/* Device bytes a single-material solve holds: particle rows, grid, tile partials and bins. */
int64_t mpmgpu_footprint(const MpmGridDesc *grid, const MpmRunParams *run, int64_t nparticles);
int mpmgpu_server_create(const MpmServerParams *params, MpmServer **srv);
/* Queue a solve. Its rows are copied to a device staging buffer on the calling thread when the
   budget has room (else to host memory, uploaded at admission). Fails for a request that can
   never fit. */
int mpmgpu_server_submit(MpmServer *srv, const MpmGridDesc *grid, const MpmMaterial *mats, int32_t nmat,
                         const MpmRunParams *run, int64_t n, const double *dfields, const int32_t *ifields,
                         int32_t steps, int64_t *id);
/* One round of the request loop on the caller's thread: admit what fits, lay out the batch again
   if its membership changed, run up to maxSteps batch steps (fewer when a solve finishes sooner)
   and retire finished solves; *completed = number retired. */
int mpmgpu_server_advance(MpmServer *srv, int32_t maxSteps, int32_t *completed);
/* Run the request loop on a thread of the server until mpmgpu_server_stop (advance is then not
   called by the caller). */
int mpmgpu_server_start(MpmServer *srv);
int mpmgpu_server_stop(MpmServer *srv);
/* Wait until request id is done (or timeoutSec passes: returns 0 with *done = 0). */
int mpmgpu_server_wait(MpmServer *srv, int64_t id, double timeoutSec, int32_t *done);
int mpmgpu_server_info(MpmServer *srv, int64_t id, MpmRequestInfo *info);
/* Final particles of a done request, in its own numbering (materials from 0, mpm[] order); the
   result rows wait on the device until this call. */
int mpmgpu_server_collect(MpmServer *srv, int64_t id, double *dfields, int32_t *ifields);
/* Queue and residency now; layouts built so far; seconds the request loop spent building layouts
   and running batch steps. */
int mpmgpu_server_counts(MpmServer *srv, int32_t *queued, int32_t *resident, int64_t *residentBytes,
                         int64_t *layouts, double *layoutSeconds, double *stepSeconds);
double mpmgpu_server_clock(const MpmServer *srv);
const char *mpmgpu_server_error(const MpmServer *srv);
void mpmgpu_server_destroy(MpmServer *srv);

#ifdef __cplusplus
}
#endif
#endif
