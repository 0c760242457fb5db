// mpm_problem.h -- host-side problem description shared by the XML reader, the generators and
// the device context. It holds what the reference keeps in process globals after reading the
// input file (mpm[], theMaterials[], mpmgrid, timestep; NairnMPM.hpp, MPMBase.hpp:269-270).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nairnmpm_gpu.h"

// Raw material input in XML units (MPa, g/cm^3), before VerifyAndLoadProperties.
struct MaterialInput {
    std::string name;
    int type = 0;
    double rho = 1.;                 // MaterialBase default (Common/Materials/MaterialBase.cpp:49)
    double E = -1., nu = -1.e30, G = -1., K = -1., Lame = -1.;
    double yield = 1.e50, yieldMin = 0., Ep = -1., Khard = 0.;
    double G2 = 0.;                  // Mooney::Mooney default (neo-Hookean limit)
    double friction = 0.;            // CoulombFriction::CoulombFriction defaults
    double frictionStatic = -1.;
    int ujOption = 0;
    double pdamping = -1.e12;        // <PDamping> in <Material> (MPMReadHandler.cpp:319-324): none below -1e12
    bool readE = false, readNu = false, readG = false;
};

struct MpmProblem {
    int analysis = MPMGPU_PLANE_STRAIN;
    int dim = 2;
    // <Grid> input
    bool haveGrid = false;
    double gxmin = 0, gxmax = 0, gymin = 0, gymax = 0, gzmin = 0, gzmax = 0;
    double cellx = -1, celly = -1, cellz = -1;
    int nhoriz = 0, nvert = 0, ndepth = 0;
    double thickness = 1.;           // MeshInfo default thickness for 2D
    MpmGridDesc grid{};
    double meshLim[6] = {0, 0, 0, 0, 0, 0};   // grid limits before the extra GIMP/CPDI elements
    bool gridBuilt = false;
    // materials
    std::vector<MaterialInput> matIn;
    std::vector<MpmMaterial> mats;
    // MPMHeader
    MpmRunParams run{};
    double cfl = 0.5;                // FractCellTime (NairnMPM.cpp:106)
    double timestepIn = 1.;          // <TimeStep> (global timestep default)
    double maxTime = 1.;             // maxtime (NairnMPM.cpp:104)
    int ptsPerSide = 2;              // MatlPtsPerElement (NairnMPM.cpp:98-99)
    int leaveLimit = 0;              // <LeaveLimit> (warnParticleLeftGrid input, NairnMPM.cpp:110)
    int deleteLimit = -1;            // <DeleteLimit> (warnParticleDeleted input, NairnMPM.cpp:112)
    bool ptsSet = false;
    std::string archiveRoot;
    std::string archiveOrder = "iY";   // <MPMArchiveOrder>, <MPMArchive> items (ArchiveData defaults)
    double archiveTime = -1.;          // <ArchiveTime> (s), < 0 if not given
    // <MultiMaterialMode> contact settings
    bool mmRequested = false;
    bool normalsGiven = false;
    double polarAngle = 0., azimuthAngle = 0.;    // <MultiMaterialMode Polar= Azimuth=> (degrees)
    std::string contactLawName;      // <Friction lawname=>
    int contactLawNum = 0;           // <Friction law=> (1-based), <0: automatic law -n-1, 0: none
    std::vector<MaterialInput> autoLaws;
    // <GridBCs> velocity conditions in the reference's list order
    std::vector<MpmVelocityBC> velBCs;
    // <ParticleBCs> loads in the reference's list order
    std::vector<MpmParticleLoad> loads;
    // <ParticleBCs> tractions in the reference's list order
    std::vector<MpmParticleTraction> tractions;
    // <CustomTasks><Schedule name="PeriodicXPIC"> (PeriodicXPIC.cpp:23-55 defaults)
    bool xpicTask = false;
    int xpicOrderIn = 0, fmpmOrderIn = 0, xpicStepsIn = -1;
    double xpicTimeIn = -1., xpicCFLIn = -1.;     // s, CFL factor
    // particles, in mpm[] order
    std::vector<double> d[MPF_NUM_FIELDS];
    std::vector<int32_t> iv[MPI_NUM_FIELDS];
    std::vector<double> pthick;      // MatPoint2D thickness
    std::vector<uint32_t> filled;    // ElementBase::filled flags used by <Body> generation
    bool finalized = false;
    std::string err;

    int64_t nparticles() const { return (int64_t)d[0].size(); }
};

// generators and preparations (mpm_setup.cpp)
namespace mpmsetup {
bool build_grid(MpmProblem &p);
bool add_material(MpmProblem &p, const MaterialInput &m);
bool add_body(MpmProblem &p, int mat1, int shapeKind, const double *ext, int axis,
              const double *vel, double thick);
bool add_point(MpmProblem &p, int mat1, const double *pos, const double *vel, const double *lp,
               double thick, double mass);
bool finalize(MpmProblem &p);
bool material_verify(int analysis, const MaterialInput &in, MpmMaterial &out, std::string &err);
}  // namespace mpmsetup

// XML input (mpm_xml.cpp)
bool mpm_read_xml(const char *text, MpmProblem &p);
