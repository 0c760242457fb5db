// NairnMPM archive files, format version 6 (ArchiveData, System/ArchiveData.cpp): a 64-byte
// header (SetArchiveHeader, :443-470) and one fixed-size record per particle (ArchiveResults,
// :721-1000, record size from CalcArchiveSize, :266-424), written from a particle state in the
// structure-of-arrays layout of the C ABI. Records are little endian (byte-order mark 'i', the
// reference's setting on x86 without its -r flag). Host code only.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nairnmpm_gpu.h"

namespace {

// MPM archive items (ArchiveData.hpp:22-32)
enum {
    A_BYTE_ORDER, A_DEFAULTS, A_VELOCITY, A_STRESS, A_STRAIN, A_PLASTIC_STRAIN, A_OLD_ORIG_POSITION,
    A_WORK_ENERGY, A_DELTA_TEMP, A_PLASTIC_ENERGY, A_VER2_EMPTY, A_SHEAR_COMPONENTS, A_STRAIN_ENERGY,
    A_HISTORY, A_CONCENTRATION, A_HEAT_ENERGY, A_ELEMENT_CROSSINGS, A_ROT_STRAIN, A_DAMAGE_NORMAL,
    A_SPIN_MOMENTUM, A_SPIN_VELOCITY, A_HISTORY59, A_HISTORY1014, A_HISTORY1519, A_SIZE, A_NUM_MPM_ITEMS
};
constexpr int NUM_CRACK_ITEMS = 7;       // ARCH_MAXCRACKITEMS (ArchiveData.hpp:35-36)
constexpr int HEADER_LENGTH = 64;
const double PI_CONSTANT = 3.141592653589793;

std::string g_archErr;

int fail(const std::string &m) {
    g_archErr = m;
    return 1;
}

// the item is written ('Y', or history bits)
bool on(const std::string &o, int i) { return o[i] != 'N'; }

// history values selected by an ARCH_History setting: 'Y' is the first, other characters ('0' | bits,
// MPMReadHandler.cpp:769-793) carry bit flags for histories 1-4 (ArchiveData.cpp:1005-1026)
std::vector<int> history_numbers(char c) {
    if (c == 'N') return {};
    if (c == 'Y') return {1};
    std::vector<int> h;
    for (int b = 0; b < 4; b++)
        if (c & (1 << b)) h.push_back(b + 1);
    return h;
}

// the record-size rules of CalcArchiveSize for the items this writer supports
int mpm_record_size(const std::string &o, bool threeD) {
    const int vec = threeD ? 24 : 16, ten = threeD ? 48 : 32;
    int s = 4 + 3 * 8 + 2 * vec + 2 + 2 + (threeD ? 8 : 0);      // defaults
    if (on(o, A_VELOCITY)) s += vec;
    if (on(o, A_STRESS)) s += ten;
    if (on(o, A_STRAIN)) s += ten;
    if (on(o, A_PLASTIC_STRAIN)) s += ten;
    if (on(o, A_SHEAR_COMPONENTS)) s += 16;
    s += 8 * (int)history_numbers(o[A_HISTORY]).size();
    if (on(o, A_ROT_STRAIN)) s += threeD ? 24 : 8;
    return s;
}

int crack_record_size(bool threeD) { return 4 + 8 + 2 + 2 + 2 * 4 + (threeD ? 12 : 8) * 8; }

// specific stress times the current density: rho0 / (relative volume), with the pressure folded
// into the full stress for materials that keep it apart (MPMBase::ReadStressTensor,
// MaterialBase::GetStress / GetStressPandDev, MaterialBaseMPM.cpp:1624-1640;
// GetCurrentRelativeVolume: 1 by default, J for Neohookean.cpp:374 and Mooney.cpp:408)
void full_stress(const MpmMaterial &m, const double *f, double s[6], double &rho) {
    double relV = 1.;
    bool pAndDev = false;
    if (m.type == MPMGPU_MAT_NEOHOOKEAN || m.type == MPMGPU_MAT_MOONEY) { relV = f[MPF_HIST0]; pAndDev = true; }
    else if (m.type == MPMGPU_MAT_ISOPLASTICITY) pAndDev = true;
    rho = m.rho / relV;
    const double P = pAndDev ? f[MPF_PRESSURE] : 0.;
    s[0] = f[MPF_SP_XX] - P; s[1] = f[MPF_SP_YY] - P; s[2] = f[MPF_SP_ZZ] - P;
    s[3] = f[MPF_SP_YZ]; s[4] = f[MPF_SP_XZ]; s[5] = f[MPF_SP_XY];
}

// Material::GetHistory(num) (MaterialBaseMPM.cpp:810-817, HardeningLawBase.cpp:80-87): the
// material's num-th history double, zero past the ones it has (J2: alpha; hyperelastic: J, Jres)
double history(const MpmMaterial &m, const double *f, int num) {
    int have = 0;
    if (m.type == MPMGPU_MAT_ISOPLASTICITY) have = 1;
    else if (m.type == MPMGPU_MAT_NEOHOOKEAN || m.type == MPMGPU_MAT_MOONEY) have = 2;
    if (num < 1 || num > have) return 0.;
    return f[MPF_HIST0 + num - 1];
}

struct Rec {
    std::vector<char> b;
    size_t at = 0;
    template <class T> void put(T v) { std::memcpy(&b[at], &v, sizeof(T)); at += sizeof(T); }
};

}  // namespace

extern "C" {

const char *mpmgpu_archive_error(void) { return g_archErr.c_str(); }

// Normalises an archive order as CalcArchiveSize does (pad with 'N', defaults 'Y', the retired
// items 'N', rotation angles with the strain, byte-order mark 'i') into out[26]; fails for items
// this implementation does not carry (energies, temperature, concentration, element crossings,
// damage normals, spin, histories past the fourth, particle size).
int mpmgpu_archive_order(const char *mpmOrder, char *out) {
    std::string o = mpmOrder ? mpmOrder : "";
    if (o.size() < 2) o = "iY";
    if ((int)o.size() > A_NUM_MPM_ITEMS) o.resize(A_NUM_MPM_ITEMS);
    while ((int)o.size() < A_NUM_MPM_ITEMS) o.push_back('N');
    o[A_BYTE_ORDER] = 'i';
    o[A_DEFAULTS] = 'Y';
    o[A_OLD_ORIG_POSITION] = 'N';
    o[A_VER2_EMPTY] = 'N';
    o[A_ROT_STRAIN] = o[A_STRAIN];
    static const struct { int item; const char *name; } missing[] = {
        {A_WORK_ENERGY, "work energy"}, {A_DELTA_TEMP, "temperature"}, {A_PLASTIC_ENERGY, "plastic energy"},
        {A_STRAIN_ENERGY, "strain energy"}, {A_CONCENTRATION, "concentration"}, {A_HEAT_ENERGY, "heat energy"},
        {A_ELEMENT_CROSSINGS, "element crossings"}, {A_DAMAGE_NORMAL, "damage normal"},
        {A_SPIN_MOMENTUM, "spin momentum"}, {A_SPIN_VELOCITY, "spin velocity"}, {A_HISTORY59, "history 5-9"},
        {A_HISTORY1014, "history 10-14"}, {A_HISTORY1519, "history 15-19"}, {A_SIZE, "particle size"}};
    for (const auto &m : missing)
        if (o[m.item] != 'N')
            return fail(std::string("archive item ") + std::to_string(m.item) + " (" + m.name +
                        ") is not carried by this implementation; set it to 'N'");
    std::memcpy(out, o.c_str(), A_NUM_MPM_ITEMS + 1);
    return 0;
}

// Bytes of one record for a normalised order: the larger of the particle and crack records
int32_t mpmgpu_archive_record_size(const char *order, int32_t dim) {
    char o[A_NUM_MPM_ITEMS + 1];
    if (mpmgpu_archive_order(order, o)) return -1;
    const bool threeD = dim == 3;
    const int m = mpm_record_size(o, threeD), c = crack_record_size(threeD);
    return m > c ? m : c;
}

// Writes an archive of n particles (MPF_*/MPI_* rows of n entries, in the reference's mpm[] order)
// at time timeMs (ms, the header's float). thickness: per particle (2D), or null for
// defaultThickness. mats: the materials the MPI_MAT indices name.
int mpmgpu_archive_write(const char *path, int32_t dim, const char *mpmOrder, double timeMs, int64_t n,
                         const double *dfields, const int32_t *ifields, const MpmMaterial *mats, int32_t nmat,
                         const double *thickness, double defaultThickness) {
    if (dim != 2 && dim != 3) return fail("archive dimension must be 2 or 3");
    char order[A_NUM_MPM_ITEMS + 1];
    if (mpmgpu_archive_order(mpmOrder, order)) return 1;
    const std::string o = order;
    const bool threeD = dim == 3;

    // header (SetArchiveHeader): "ver6", length + mpmOrder, length + crackOrder, '3' or '2', '1'
    // (a regular <Grid> is structured), 4-byte float time in ms, zeros
    char header[HEADER_LENGTH];
    std::memset(header, 0, sizeof(header));
    std::string crack = "iY";
    while ((int)crack.size() < NUM_CRACK_ITEMS) crack.push_back('N');
    size_t k = 0;
    std::memcpy(header, "ver6", 4); k = 4;
    header[k++] = (char)A_NUM_MPM_ITEMS;
    std::memcpy(header + k, o.c_str(), A_NUM_MPM_ITEMS); k += A_NUM_MPM_ITEMS;
    header[k++] = (char)NUM_CRACK_ITEMS;
    std::memcpy(header + k, crack.c_str(), NUM_CRACK_ITEMS); k += NUM_CRACK_ITEMS;
    header[k++] = threeD ? '3' : '2';
    header[k++] = '1';
    const float stamp = (float)timeMs;
    std::memcpy(header + k, &stamp, sizeof(float));

    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(std::string("cannot open archive file ") + path);
    if (std::fwrite(header, 1, HEADER_LENGTH, f) != (size_t)HEADER_LENGTH) {
        std::fclose(f);
        return fail(std::string("error writing archive header to ") + path);
    }
    const int mrec = mpm_record_size(o, threeD), crec = crack_record_size(threeD);
    const int recSize = mrec > crec ? mrec : crec;
    const std::vector<int> hist = history_numbers(o[A_HISTORY]);
    Rec r;
    r.b.assign(recSize, 0);
    double fp[MPF_NUM_FIELDS];
    for (int64_t p = 0; p < n; p++) {
        for (int q = 0; q < MPF_NUM_FIELDS; q++) fp[q] = dfields[(size_t)q * n + p];
        const int mat = ifields[(size_t)MPI_MAT * n + p];
        if (mat < 0 || mat >= nmat) {
            std::fclose(f);
            return fail("particle material index out of range");
        }
        const MpmMaterial &m = mats[mat];
        std::fill(r.b.begin(), r.b.end(), 0);
        r.at = 0;
        // defaults: element (one based), mass, material (one based, short + 2 pad bytes)
        r.put<int32_t>(ifields[(size_t)MPI_ELEM * n + p] + 1);
        r.put<double>(fp[MPF_MASS]);
        r.put<int16_t>((int16_t)(mat + 1));
        r.put<char>(0); r.put<char>(0);
        // rotation angles in degrees from wrot with zero initial angles (MPMBase.cpp:601-622)
        if (threeD) {
            r.put<double>(180. * (0. - 0.5 * fp[MPF_WR_XY]) / PI_CONSTANT);
            r.put<double>(180. * (0. + 0.5 * fp[MPF_WR_XZ]) / PI_CONSTANT);
            r.put<double>(180. * (0. - 0.5 * fp[MPF_WR_YZ]) / PI_CONSTANT);
        } else {
            r.put<double>(180. * (0. - 0.5 * fp[MPF_WR_XY]) / PI_CONSTANT);
            r.put<double>(thickness ? thickness[p] : defaultThickness);
        }
        r.put<double>(fp[MPF_POS_X]); r.put<double>(fp[MPF_POS_Y]);
        if (threeD) r.put<double>(fp[MPF_POS_Z]);
        r.put<double>(fp[MPF_ORIG_X]); r.put<double>(fp[MPF_ORIG_Y]);
        if (threeD) r.put<double>(fp[MPF_ORIG_Z]);
        if (on(o, A_VELOCITY)) {
            r.put<double>(fp[MPF_VEL_X]); r.put<double>(fp[MPF_VEL_Y]);
            if (threeD) r.put<double>(fp[MPF_VEL_Z]);
        }
        if (on(o, A_STRESS)) {            // xx, yy, zz, xy (, xz, yz) in Pa
            double s[6], rho;
            full_stress(m, fp, s, rho);
            r.put<double>(rho * s[0]); r.put<double>(rho * s[1]); r.put<double>(rho * s[2]);
            r.put<double>(rho * s[5]);
            if (threeD) { r.put<double>(rho * s[4]); r.put<double>(rho * s[3]); }
        }
        if (on(o, A_STRAIN)) {            // ep (absolute)
            r.put<double>(fp[MPF_EP_XX]); r.put<double>(fp[MPF_EP_YY]); r.put<double>(fp[MPF_EP_ZZ]);
            r.put<double>(fp[MPF_EP_XY]);
            if (threeD) { r.put<double>(fp[MPF_EP_XZ]); r.put<double>(fp[MPF_EP_YZ]); }
        }
        if (on(o, A_PLASTIC_STRAIN)) {    // eplast or B (absolute)
            r.put<double>(fp[MPF_ALT_XX]); r.put<double>(fp[MPF_ALT_YY]); r.put<double>(fp[MPF_ALT_ZZ]);
            r.put<double>(fp[MPF_ALT_XY]);
            if (threeD) { r.put<double>(fp[MPF_ALT_XZ]); r.put<double>(fp[MPF_ALT_YZ]); }
        }
        if (on(o, A_SHEAR_COMPONENTS)) {  // gradU(0,1), gradU(1,0) with gradU = F - I
            r.put<double>(0.5 * (fp[MPF_EP_XY] - fp[MPF_WR_XY]));
            r.put<double>(0.5 * (fp[MPF_EP_XY] + fp[MPF_WR_XY]));
        }
        for (int h : hist) r.put<double>(history(m, fp, h));
        if (on(o, A_ROT_STRAIN)) {        // initial angles (degrees): zero
            r.put<double>(0.);
            if (threeD) { r.put<double>(0.); r.put<double>(0.); }
        }
        if (std::fwrite(r.b.data(), 1, recSize, f) != (size_t)recSize) {
            std::fclose(f);
            return fail(std::string("error writing archive records to ") + path);
        }
    }
    if (std::fclose(f) != 0) return fail(std::string("error closing archive file ") + path);
    return 0;
}

}  // extern "C"
