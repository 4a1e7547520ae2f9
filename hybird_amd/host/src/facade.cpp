// C++ host layer: the reference's LB / DEM / IO classes over the C ABI of
// libhybird_amd.so. Parameter parsing reproduces LB.cpp:85-188,
// DEM.cpp:13-183 and hybird.cpp's flags; all solver work happens on the GPU.
#include <cerrno>
#include <cstdio>
#include <fstream>
#include <algorithm>
#include <iomanip>
#include <sstream>

#include "DEM.h"
#include "IO.h"
#include "LB.h"

void hybird_check(int rc, const char* what, const char* file, int line) {
    if (rc == HYB_OK) return;
    cout << file << ":" << line << " ERROR: " << what << " failed (" << rc << "): " << hyb_last_error() << std::endl;
    exit(1);
}

namespace {

// Optional extension keys: read when present, without the missing-parameter
// warning that the reference's keys produce. "-name value" overrides them too.
template <class T>
T extension_key(GetPot& cfg, GetPot& command_line, const char* name, T def) {
    T v = cfg(name, def);
    const std::string flag = std::string("-") + name;
    if (command_line.search(flag.c_str())) v = command_line.next(v);
    return v;
}

std::string upper(std::string s) {
    for (char& c : s) c = (char)toupper((unsigned char)c);
    return s;
}

// Writes a host view to <dir>/<name>.bin (raw little-endian).
template <class T>
void dump(const std::string& dir, const char* name, const std::vector<T>& v) {
    FILE* f = std::fopen((dir + "/" + name + ".bin").c_str(), "wb");
    if (!f) return;
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

template <class T, class H, class G>
bool view(H* h, G get, const char* name, std::vector<T>& out) {
    size_t need = 0;
    if (get(h, name, nullptr, 0, &need) != HYB_OK) return false;
    out.resize(need / sizeof(T));
    if (need && get(h, name, out.data(), need, &need) != HYB_OK) return false;
    return true;
}

std::vector<double> saved_part(hyb_dem* h, int32_t ev, const char* part) {
    std::vector<double> out;
    size_t need = 0;
    HYB_CALL(hyb_dem_saved_get(h, ev, part, nullptr, 0, &need));
    out.resize(need / sizeof(double));
    if (need) HYB_CALL(hyb_dem_saved_get(h, ev, part, out.data(), need, &need));
    return out;
}

// One .fstat contact line (DEM.cpp:2479-2565): time, the two indices, contact
// centre, overlap, tangential spring length (always 0), |Fn|, |Ft|, en, et.
void stat_line(ofstream& f, double time, long i, long j, const tVect& centre, const double* r) {
    f << setprecision(3) << time << " " << i << " " << j;
    const double v[13] = {centre.x, centre.y, centre.z, r[2], 0.0, std::abs(r[3]), std::abs(r[4]),
                          r[5], r[6], r[7], r[8], r[9], r[10]};
    for (double x : v) f << " " << setprecision(15) << x;
    f << endl;
}

}  // namespace

// ======================================================================== LB
LB::LB() {
    hyb_config_defaults(&cfg_);
    boundary.assign(6, 1);
    lbSize.assign(3, 1);
}

LB::~LB() {
    if (h_ && inited_ && !saveStateFile_.empty()) {  // extension: checkpoint at the end of the run
        if (hyb_state_save(h_, dem_, saveStateFile_.c_str()) != HYB_OK)
            cout << "WARNING: state file not written: " << hyb_last_error() << "\n";
        else
            cout << "state written to " << saveStateFile_ << "\n";
    }
    const char* dir = std::getenv("HYBIRD_AMD_DUMP_DIR");
    if (h_ && inited_ && dir && *dir) {
        std::vector<double> d;
        std::vector<unsigned char> t;
        if (view(h_, hyb_lb_get, "types", t)) dump(dir, "lb_types", t);
        for (const char* name : {"f", "n", "u", "hydroForce", "mass", "fluidMass"})
            if (view(h_, hyb_lb_get, name, d)) dump(dir, (std::string("lb_") + name).c_str(), d);
    }
    if (h_) hyb_lb_destroy(h_);
}

void LB::latticeBoltzmannGet(GetPot& lbmCfgFile, GetPot& command_line) {  // LB.cpp:85-188
    PARSE_CLASS_MEMBER(lbmCfgFile, unit.Length, "unitLength", 1.0);
    ASSERT(unit.Length > 0);
    PARSE_CLASS_MEMBER(lbmCfgFile, unit.Time, "unitTime", 1.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, unit.Density, "unitDensity", 1.0);
    ASSERT(unit.Density > 0);
    double lbSizeX = 0.0, lbSizeY = 0.0, lbSizeZ = 0.0;
    PARSE_CLASS_MEMBER(lbmCfgFile, lbSizeX, "lbSizeX", 0.0);
    ASSERT(lbSizeX > 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, lbSizeY, "lbSizeY", 0.0);
    ASSERT(lbSizeY > 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, lbSizeZ, "lbSizeZ", 0.0);
    ASSERT(lbSizeZ > 0.0);
    lbSize.assign(3, 0);
    lbSize[0] = int(std::floor(lbSizeX / unit.Length + 0.5));
    lbSize[1] = int(std::floor(lbSizeY / unit.Length + 0.5));
    lbSize[2] = int(std::floor(lbSizeZ / unit.Length + 0.5));
    PARSE_CLASS_MEMBER(lbmCfgFile, plasticVisc, "plasticVisc", 1.0);
    ASSERT(plasticVisc >= 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, yieldStress, "yieldStress", 1.0);
    ASSERT(yieldStress >= 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, initDynVisc, "initVisc", 1.0);
    ASSERT(initDynVisc > 0.0);
    cfg_.unitLength = unit.Length;
    cfg_.unitTime = unit.Time;
    cfg_.unitDensity = unit.Density;
    cfg_.lbSize[0] = lbSizeX; cfg_.lbSize[1] = lbSizeY; cfg_.lbSize[2] = lbSizeZ;
    cfg_.plasticVisc = plasticVisc;
    cfg_.yieldStress = yieldStress;
    cfg_.initVisc = initDynVisc;
    if (unit.Time == 0.0) {  // time step for tau = 1 (LB.cpp:124-133)
        const double tauOpt = 1.0;
        unit.Time = nonNewtonian ? unit.Density * unit.Length * unit.Length / plasticVisc * (1.0 / 3.0) * (tauOpt - 0.5)
                                 : unit.Density * unit.Length * unit.Length / initDynVisc * (1.0 / 3.0) * (tauOpt - 0.5);
    }
    unit.setComposite();
    initDynVisc /= unit.DynVisc;
    plasticVisc /= unit.DynVisc;
    yieldStress /= unit.Stress;
    initDensity = 1.0;
    double initVelocityX = 0.0, initVelocityY = 0.0, initVelocityZ = 0.0;
    PARSE_CLASS_MEMBER(lbmCfgFile, initVelocityX, "initVelocityX", 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, initVelocityY, "initVelocityY", 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, initVelocityZ, "initVelocityZ", 0.0);
    cfg_.initVelocity[0] = initVelocityX; cfg_.initVelocity[1] = initVelocityY; cfg_.initVelocity[2] = initVelocityZ;
    initVelocity = tVect(initVelocityX, initVelocityY, initVelocityZ) / unit.Speed;
    double lbFX = 0.0, lbFY = 0.0, lbFZ = 0.0;
    PARSE_CLASS_MEMBER(lbmCfgFile, lbFX, "lbFX", 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, lbFY, "lbFY", 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, lbFZ, "lbFZ", 0.0);
    cfg_.lbF[0] = lbFX; cfg_.lbF[1] = lbFY; cfg_.lbF[2] = lbFZ;
    lbF = tVect(lbFX, lbFY, lbFZ) / unit.Accel;
    boundary.assign(6, 1);
    PARSE_CLASS_MEMBER(lbmCfgFile, boundary[0], "boundary0", 1);
    PARSE_CLASS_MEMBER(lbmCfgFile, boundary[1], "boundary1", 1);
    PARSE_CLASS_MEMBER(lbmCfgFile, boundary[2], "boundary2", 1);
    PARSE_CLASS_MEMBER(lbmCfgFile, boundary[3], "boundary3", 1);
    PARSE_CLASS_MEMBER(lbmCfgFile, boundary[4], "boundary4", 1);
    PARSE_CLASS_MEMBER(lbmCfgFile, boundary[5], "boundary5", 1);
    for (int k = 0; k < 6; ++k) cfg_.boundary[k] = (int32_t)boundary[k];
    PARSE_CLASS_MEMBER(lbmCfgFile, slipCoefficient, "slipCoefficient", 0.0);
    ASSERT(slipCoefficient >= 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, turbConst, "turbConst", 0.0);
    ASSERT(turbConst >= 0.0);
    cfg_.slipCoefficient = slipCoefficient;
    cfg_.turbConst = turbConst;
    // extensions: collision model, storage precision, generic liquid box
    const std::string coll = upper(extension_key<std::string>(lbmCfgFile, command_line, "collisionModel", "BGK"));
    ASSERT(coll == "BGK" || coll == "TRT");
    cfg_.collision = coll == "TRT" ? HYB_COLLISION_TRT : HYB_COLLISION_BGK;
    cfg_.trtMagic = extension_key<double>(lbmCfgFile, command_line, "trtMagic", 0.25);
    const std::string prec = upper(extension_key<std::string>(lbmCfgFile, command_line, "precision", "FP64"));
    ASSERT(prec == "FP64" || prec == "FP32");
    cfg_.precision = prec == "FP32" ? HYB_FP32_SHIFTED : HYB_FP64;
    const std::string rheo = upper(extension_key<std::string>(lbmCfgFile, command_line, "rheology", "BINGHAM"));
    ASSERT(rheo == "BINGHAM" || rheo == "FRICTIONAL");
    cfg_.rheology = rheo == "FRICTIONAL" ? HYB_RHEOLOGY_FRICTIONAL : HYB_RHEOLOGY_BINGHAM;
    cfg_.frictionCoefFluid = extension_key<double>(lbmCfgFile, command_line, "frictionCoefFluid", 0.0);
    ASSERT(cfg_.frictionCoefFluid >= 0.0);
    // checkpoint and restart (the reference cannot resume a run)
    restartFile_ = extension_key<std::string>(lbmCfgFile, command_line, "restartFile", "");
    // fluid files and the status line need u of every collision (kept on the device in
    // free-surface mode)
    fluidExport_ = extension_key<double>(lbmCfgFile, command_line, "fluidExpTime", 0.0) > 0.0 ||
                   extension_key<double>(lbmCfgFile, command_line, "screenExpTime", 0.0) > 0.0;
    saveStateFile_ = extension_key<std::string>(lbmCfgFile, command_line, "saveStateFile", "");
    const char* boxKeys[6] = {"liquidBoxXMin", "liquidBoxXMax", "liquidBoxYMin", "liquidBoxYMax", "liquidBoxZMin",
                              "liquidBoxZMax"};
    bool box = false;
    for (int k = 0; k < 6; ++k) box |= lbmCfgFile.have_variable(boxKeys[k]);
    if (box) {
        const double inf = 1.0e300;
        for (int k = 0; k < 6; ++k)
            cfg_.liquidBox[k] = extension_key<double>(lbmCfgFile, command_line, boxKeys[k], (k % 2) ? inf : -inf);
        cfg_.problem = HYB_PROBLEM_LIQUID_BOX;
    } else {
        cfg_.problem = -1;  // taken from problemName in latticeDefinition
    }
}

void LB::latticeDefinition() {  // LB.cpp:42-83 (the lattice itself lives on the device)
    if (cfg_.problem != HYB_PROBLEM_LIQUID_BOX) cfg_.problem = (int32_t)problemName;
    cfg_.avalanchePosit = avalanchePosit;  // parsed by the driver for NET and BARRIER (hybird.cpp:216-219)
    cfg_.fluidMass = fluidMass;            // parsed by the driver for DRUM (hybird.cpp:199-205)
    cfg_.freeSurface = freeSurface;
    cfg_.forceField = forceField;
    cfg_.nonNewtonian = nonNewtonian;
    cfg_.turbulence = turbulenceOn;
    cfg_.lbSolve = 1;
    HYB_CALL(hyb_lb_create(&cfg_, &h_));
    totNodes = lbSize[0] * lbSize[1] * lbSize[2];
}

void LB::LBShow() const {  // LB.cpp:2290-2312
    cout << "LATTICE CHARACTERISTICS (lattice units):\n";
    cout << "directions=19;\n";
    cout << "time unit = " << unit.Time << "; length unit = " << unit.Length << "; density unit = " << unit.Density
         << ";\n";
    cout << "lbSize: " << lbSize[0] << ", " << lbSize[1] << ", " << lbSize[2] << ";\n";
    cout << "Init kin viscosity=" << initDynVisc << ", Init tau=" << (0.5 + 3.0 * initDynVisc) << ";\n";
    cout << "Init density=" << initDensity << ";\n";
    cout << "Init velocity=";
    initVelocity.show();
    cout << ";\nForce field=";
    lbF.show();
    cout << ";\nCollision=" << (cfg_.collision == HYB_COLLISION_TRT ? "TRT" : "BGK")
         << "; populations=" << (cfg_.precision == HYB_FP32_SHIFTED ? "FP32 (f - w)" : "FP64") << ";\n";
}

void LB::latticeBolzmannInit(cylinderList& cylinders, wallList& walls, particleList& particles,
                             objectList& objects) {  // LB.cpp:190-219
    (void)walls;
    (void)objects;  // the DEM handle carries them (hyb_dem_set_objects)
    if (!cylinders.empty()) {
        cout << INFO << "ERROR: cylinders are not supported by the GPU lattice" << std::endl;
        exit(1);
    }
    DEM* dem = particles.owner;
    dem_ = dem ? dem->handle() : nullptr;
    HYB_CALL(hyb_lb_init(h_, dem_));
    inited_ = true;
    if (freeSurface && fluidExport_) HYB_CALL(hyb_lb_set_capture(h_, 1));
    if (!restartFile_.empty()) {  // extension: continue from a state file
        HYB_CALL(hyb_state_load(h_, dem_, restartFile_.c_str()));
        cout << "state restored from " << restartFile_ << "\n";
    }
    std::vector<double> tm;
    if (view(h_, hyb_lb_get, "totalMass", tm) && !tm.empty()) totalMass = tm[0];
}

void LB::latticeBolzmannStep(elmtList& elmts, particleList& particles, wallList& walls) {  // LB.cpp:221-233
    (void)particles;
    (void)walls;
    DEM* dem = elmts.owner;
    HYB_CALL(hyb_lb_step(h_, dem ? dem->handle() : nullptr));
}

void LB::latticeBoltzmannCouplingStep(bool& newNeighborList, elmtList& elmts,
                                      particleList& particles) {  // LB.cpp:247-280
    (void)particles;
    DEM* dem = elmts.owner;
    int32_t flag = newNeighborList ? 1 : 0;
    HYB_CALL(hyb_lb_couple(h_, dem ? dem->handle() : nullptr, &flag));
    newNeighborList = flag != 0;
}

void LB::latticeBoltzmannFreeSurfaceStep() { HYB_CALL(hyb_lb_fs_step(h_)); }  // LB.cpp:235-245

std::vector<double> LB::field(const char* name) const {
    std::vector<double> v;
    if (!view(h_, hyb_lb_get, name, v)) hybird_check(HYB_E_ARG, name, __FILE__, __LINE__);
    return v;
}

std::vector<unsigned char> LB::nodeFlags(const char* name) const {
    std::vector<unsigned char> v;
    if (!view(h_, hyb_lb_get, name, v)) hybird_check(HYB_E_ARG, name, __FILE__, __LINE__);
    return v;
}

std::vector<unsigned int> LB::solidIndices() const {
    std::vector<unsigned int> v;
    if (!view(h_, hyb_lb_get, "solidIndex", v)) hybird_check(HYB_E_ARG, "solidIndex", __FILE__, __LINE__);
    return v;
}

std::vector<unsigned char> LB::nodeTypes() const {
    std::vector<unsigned char> t;
    if (!view(h_, hyb_lb_get, "types", t)) hybird_check(HYB_E_ARG, "types", __FILE__, __LINE__);
    return t;
}

double LB::totalFluidMass() const {
    const std::vector<double> m = field("fluidMass");
    return m.empty() ? 0.0 : m[0];
}

std::vector<double> LB::statusValues() const { return field("status"); }

// ======================================================================== DEM
DEM::DEM() {
    hyb_config_defaults(&cfg_);
    walls.owner = this;
    cylinders.owner = this;
    pbcs.owner = this;
    elmts.owner = this;
    particles.owner = this;
    objects.owner = this;
}

DEM::~DEM() {
    const char* dir = std::getenv("HYBIRD_AMD_DUMP_DIR");
    if (h_ && dir && *dir) {
        std::vector<double> d;
        std::vector<uint32_t> u;
        for (const char* name : {"e_x", "e_q", "e_w", "e_forces", "dem_scalars"})
            if (view(h_, hyb_dem_get, name, d)) dump(dir, (std::string("dem_") + name).c_str(), d);
        for (const char* name : {"neighborTable", "nearWallTable"})
            if (view(h_, hyb_dem_get, name, u)) dump(dir, (std::string("dem_") + name).c_str(), u);
    }
    if (h_) hyb_dem_destroy(h_);
}

void DEM::discreteElementGet(GetPot& lbmCfgFile, GetPot& command_line) {  // DEM.cpp:13-183
    double density = 1.0;
    PARSE_CLASS_MEMBER(lbmCfgFile, density, "density", 1.0);
    ASSERT(density > 0.0);
    string contactModelString;
    PARSE_CLASS_MEMBER(lbmCfgFile, contactModelString, "contactModel", "none");
    if (contactModelString == "HERTZIAN") cfg_.contactModel = HYB_CONTACT_HERTZIAN;
    else if (contactModelString == "LINEAR") cfg_.contactModel = HYB_CONTACT_LINEAR;
    double youngMod = 1.0, poisson = 1.0, linearStiff = 1.0, restitution = 1.0;
    PARSE_CLASS_MEMBER(lbmCfgFile, youngMod, "youngMod", 1.0);
    ASSERT(youngMod > 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, poisson, "poisson", 1.0);
    ASSERT(poisson >= 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, linearStiff, "linearStiff", 1.0);
    ASSERT(linearStiff >= 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, restitution, "restitution", 1.0);
    ASSERT(restitution > 0.0);
    ASSERT(restitution <= 1.0);
    double viscTang = 1.0, frictionCoefPart = 1.0, frictionCoefWall = 1.0;
    PARSE_CLASS_MEMBER(lbmCfgFile, viscTang, "viscTang", 1.0);
    ASSERT(viscTang >= 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, frictionCoefPart, "frictionCoefPart", 1.0);
    ASSERT(frictionCoefPart >= 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, frictionCoefWall, "frictionCoefWall", 1.0);
    ASSERT(frictionCoefWall >= 0.0);
    cfg_.density = density;
    cfg_.youngMod = youngMod;
    cfg_.poisson = poisson;
    cfg_.linearStiff = linearStiff;
    cfg_.restitution = restitution;
    cfg_.viscTang = viscTang;
    cfg_.frictionCoefPart = frictionCoefPart;
    cfg_.frictionCoefWall = frictionCoefWall;
    {  // extension: tangential contact law
        const std::string tm = upper(extension_key<std::string>(lbmCfgFile, command_line, "tangentialModel", "VISCOUS"));
        ASSERT(tm == "VISCOUS" || tm == "SPRING");
        cfg_.tangentialModel = tm == "SPRING" ? HYB_TANGENTIAL_SPRING : HYB_TANGENTIAL_VISCOUS;
    }
    {  // extension: contacts with every wall in range instead of the first one
        const std::string wc = upper(extension_key<std::string>(lbmCfgFile, command_line, "wallContacts", "FIRST"));
        ASSERT(wc == "FIRST" || wc == "ALL");
        cfg_.wallContacts = wc == "ALL" ? HYB_WALL_CONTACTS_ALL : HYB_WALL_CONTACTS_FIRST;
    }
    // particle initial state (DEM.cpp:63-116)
    string particleFile;
    PARSE_CLASS_MEMBER(lbmCfgFile, particleFile, "particleFile", "particles.dat");
    double translateX = 0.0, translateY = 0.0, translateZ = 0.0;
    PARSE_CLASS_MEMBER(lbmCfgFile, translateX, "translateX", 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, translateY, "translateY", 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, translateZ, "translateZ", 0.0);
    double scale = 1.0;
    PARSE_CLASS_MEMBER(lbmCfgFile, scale, "scale", 1.0);
    cfg_.translate[0] = translateX; cfg_.translate[1] = translateY; cfg_.translate[2] = translateZ;
    cfg_.scale = scale;
    ifstream particleFileID(particleFile.c_str(), ios::in);
    ASSERT(particleFileID.is_open());
    unsigned int totElmt = 0;
    particleFileID >> totElmt;
    records_.assign((size_t)totElmt * HYB_ELEMENT_RECORD, 0.0);
    elmts.clear();
    for (unsigned int n = 0; n < totElmt; ++n) {
        double* r = &records_[(size_t)n * HYB_ELEMENT_RECORD];
        for (int k = 0; k < HYB_ELEMENT_RECORD; ++k) particleFileID >> r[k];
        elmt e;
        e.index = (unsigned)r[0];
        e.size = (unsigned)r[1];
        e.radius = r[2] * scale;
        e.x0 = tVect(r[3] * scale, r[4] * scale, r[5] * scale) + tVect(translateX, translateY, translateZ);
        e.x1 = tVect(r[6], r[7], r[8]);
        e.w0 = tVect(r[9], r[10], r[11]);
        e.q0 = tQuat(r[12], r[13], r[14], r[15]);
        e.q1 = tQuat(r[16], r[17], r[18], r[19]);
        elmts.push_back(e);
    }
    // objects (DEM.cpp:131-173): static spheres; the file must exist. Positions
    // are taken as written (no scale or translation); the rotation values are skipped.
    string objectFile;
    PARSE_CLASS_MEMBER(lbmCfgFile, objectFile, "objectFile", "objects.dat");
    ifstream objectFileID(objectFile.c_str(), ios::in);
    ASSERT(objectFileID.is_open());
    unsigned int totObjects = 0;
    objectFileID >> totObjects;
    stdObjects = totObjects;
    objects.clear();
    objectRecords_.clear();
    for (unsigned int n = 0; n < totObjects; ++n) {
        double v[20];
        for (int k = 0; k < 20; ++k) objectFileID >> v[k];
        object o;
        o.index = o.originalIndex = (unsigned int)v[0];
        o.r = v[2];
        o.x0 = tVect(v[3], v[4], v[5]);
        o.x1 = tVect(v[6], v[7], v[8]);
        objects.push_back(o);
        for (int k = 0; k < HYB_OBJECT_RECORD; ++k) objectRecords_.push_back(v[k]);
    }
    double numVisc = 0.0;
    PARSE_CLASS_MEMBER(lbmCfgFile, numVisc, "numVisc", 0.0);
    PARSE_CLASS_MEMBER(lbmCfgFile, multiStep, "multiStep", 1);
    PARSE_CLASS_MEMBER(lbmCfgFile, criticalRatio, "criticalRatio", 0.1);
    cfg_.numVisc = numVisc;
}

void DEM::discreteElementInit(unsIntList& boundary, unsIntList& lbSize, measureUnits& unit,
                              tVect& lbF) {  // DEM.cpp:185-297
    cfg_.multiStep = (int32_t)multiStep;
    // the problem decides the cylinders (DEM.cpp:870-935)
    cfg_.problem = (int32_t)problemName;
    cfg_.drumSpeed = drumSpeed;  // parsed by the driver for DRUM (hybird.cpp:199-205)
    cfg_.criticalRatio = criticalRatio;
    cfg_.demInitialRepeat = demInitialRepeat;
    cfg_.demSolve = 1;
    const int32_t P = (int32_t)(records_.size() / HYB_ELEMENT_RECORD);
    HYB_CALL(hyb_dem_create(&cfg_, P, P ? records_.data() : nullptr, &h_));
    if (!objectRecords_.empty())
        HYB_CALL(hyb_dem_set_objects(h_, (int32_t)(objectRecords_.size() / HYB_OBJECT_RECORD), objectRecords_.data()));
    int32_t b[6], s[3];
    for (int k = 0; k < 6; ++k) b[k] = (int32_t)boundary[k];
    for (int k = 0; k < 3; ++k) s[k] = (int32_t)lbSize[k];
    const double F[3] = {lbF.x, lbF.y, lbF.z};
    HYB_CALL(hyb_dem_init_lattice(h_, b, s, unit.Length, unit.Time, unit.Density, F));
    std::vector<double> sc;
    if (view(h_, hyb_dem_get, "dem_scalars", sc) && sc.size() >= 12) {
        deltat = sc[10];
        multiStep = (unsigned)sc[11];
    }
    demF = lbF * unit.Accel;
    // box walls as host records (DEM.cpp:435-638)
    walls.clear();
    const tVect ns[6] = {tVect(1, 0, 0), tVect(-1, 0, 0), tVect(0, 1, 0), tVect(0, -1, 0), tVect(0, 0, 1), tVect(0, 0, -1)};
    for (int k = 0; k < 6; ++k) {
        const unsigned bc = boundary[k];
        if (bc < 5 || bc > 8) continue;
        wall w;
        w.index = (unsigned)walls.size();
        w.n = ns[k];
        const double pos = (k % 2 == 0) ? 0.5 : double(lbSize[k / 2]) - 1.5;
        w.p = tVect(k / 2 == 0 ? pos : 0.0, k / 2 == 1 ? pos : 0.0, k / 2 == 2 ? pos : 0.0) * unit.Length;
        w.moving = bc == 6 || bc == 8;
        w.slip = bc == 5 || bc == 6;
        walls.push_back(w);
    }
    // periodic pairs as host records (DEM.cpp:937-988): planes at 0.5 with period N - 2
    pbcs.clear();
    for (int k = 0; k < 3; ++k) {
        if (boundary[2 * k] != 4) continue;
        pbc b;
        const double span = (double(lbSize[k]) - 2.0) * unit.Length;
        b.p1 = tVect(k == 0 ? 0.5 * unit.Length : 0.0, k == 1 ? 0.5 * unit.Length : 0.0, k == 2 ? 0.5 * unit.Length : 0.0);
        b.v = tVect(k == 0 ? span : 0.0, k == 1 ? span : 0.0, k == 2 ? span : 0.0);
        b.p2 = b.p1 + b.v;
        pbcs.push_back(b);
    }
    // the spheres of every element, consecutive and in prototype order
    // (elmt::generateParticles, elmt.cpp:122-137); positions from the device
    particles.clear();
    for (elmt& e : elmts) {
        e.components.clear();
        for (unsigned i = 0; i < e.size; ++i) {
            particle p;
            p.particleIndex = (unsigned)particles.size();
            p.clusterIndex = e.index;
            p.protoIndex = i;
            p.r = e.radius;
            p.x0 = e.x0;
            p.x1 = e.x1;
            e.components.push_back((int)p.particleIndex);
            particles.push_back(p);
        }
    }
    stdParticles = (unsigned)particles.size();
    if (stdParticles != elmts.size()) syncHost();
}

void DEM::discreteElementStep(IO& io) {  // DEM.cpp:331-376
    if (io.saveCount != saveCountSet_) {
        HYB_CALL(hyb_dem_set_save_count(h_, io.saveCount));
        saveCountSet_ = io.saveCount;
    }
    HYB_CALL(hyb_dem_step(h_));
    // the same time bookkeeping as the device side, without a device round trip
    for (unsigned it = 0; it < multiStep; ++it) {
        demTimeStep++;
        demTime += deltat;
    }
    if (io.saveCount != 0) exportSaved(io);
}

// The MercuryDPM files of each saving sub-step of the last device step, in
// sub-step order: exportDataFile (DEM.cpp:2280-2333, appended), exportRestartFile
// (2337-2434, rewritten), printHeaderStatFile and the contact lines of the force
// evaluation (1131-1146, 2447-2565, appended): particle-particle contacts in
// neighbour-table order without ghosts, then wall contacts, then object contacts,
// each only with tangential motion. Particle state is the predicted one.
void DEM::exportSaved(IO& io) {
    int32_t n = 0;
    HYB_CALL(hyb_dem_saved(h_, &n));
    if (n == 0) return;
    const std::vector<double> misc = field("e_misc");  // [E][19]: m, radius, I, ...
    const size_t E = elmts.size();
    const size_t P = stdParticles;  // the captures hold standard particles
    const auto cl = [&](size_t i) { return (size_t)particles[i].clusterIndex; };
    const double zero = 0.0;
    for (int32_t ev = 0; ev < n; ++ev) {
        const std::vector<double> info = saved_part(h_, ev, "info");
        const std::vector<double> el = saved_part(h_, ev, "elements");
        const double time = info[0];
        const unsigned int step = (unsigned int)info[1];
        const double dampCoeff = info[5], dt = info[6];
        const double size[3] = {info[7], info[8], info[9]};
        auto pos = [&](size_t e) { return tVect(el[9 * e], el[9 * e + 1], el[9 * e + 2]); };
        auto vel = [&](size_t e) { return tVect(el[9 * e + 3], el[9 * e + 4], el[9 * e + 5]); };
        auto wg = [&](size_t e) { return tVect(el[9 * e + 6], el[9 * e + 7], el[9 * e + 8]); };
        auto put3 = [](ofstream& f, const tVect& v) { f << v.x << " " << v.y << " " << v.z << " "; };
        // ---- .data: precision 5 for the time, then 6 + 5 (the stream's default + 5)
        {
            ofstream f(io.dataFileName.c_str(), ios::app);
            f << stdParticles + stdObjects << " " << setprecision(5) << time << " " << setprecision(11) << zero << " "
              << setprecision(11) << zero << " " << setprecision(11) << zero << " " << setprecision(11) << size[0]
              << " " << setprecision(11) << size[1] << " " << setprecision(11) << size[2] << "\n";
            for (unsigned o = 0; o < stdObjects && o < objects.size(); ++o) {
                put3(f, objects[o].x0);
                put3(f, objects[o].x1);
                f << objects[o].r << " " << zero << " " << zero << " " << zero << " " << zero << " " << zero << " "
                  << zero << " " << zero << " " << "\n";
            }
            for (size_t i = 0; i < P; ++i) {
                put3(f, pos(i));
                put3(f, vel(i));
                f << elmts[cl(i)].radius << " " << zero << " " << zero << " " << zero << " ";
                put3(f, wg(i));
                f << zero << " " << "\n";
            }
        }
        // ---- .restart (rewritten; the first one starts at the default precision)
        {
            ofstream f(io.restartFileName.c_str());
            f << setprecision(io.restartPrecision);
            const double counter = (double)(step / (unsigned int)io.saveCount);
            f << "restart_version " << 1 << " name" << " " << io.problemNameString << "\n";
            const char* files[5] = {"dataFile ", "fStatFile ", "eneFile ", "restartFile ", "statFile "};
            for (const char* nm : files)
                f << nm << "fileType " << "ONE_FILE " << "saveCount " << io.saveCount << " " << "counter " << counter
                  << " nextSavedTimeStep " << 0 << "\n";
            f << "xMin " << 0 << " xMax " << size[0] << " yMin " << 0 << " yMax " << size[1] << " zMin " << 0
              << " zMax " << size[2] << "\n"
              << "timeStep " << setprecision(10) << dt << " time " << time << " ntimeSteps " << step << " timeMax "
              << time << "\n"
              << "systemDimensions " << 3 << " particleDimensions " << 3 << " gravity " << demF.x << " " << demF.y
              << " " << demF.z << "\n"
              << "Species " << 1 << "\n"
              << "LinearViscoelasticSlidingFrictionSpecies id " << 0 << " density " << cfg_.density << " stiffness "
              << cfg_.linearStiff << " dissipation " << dampCoeff << " slidingStiffness " << 0
              << " slidingDissipation " << 0 << " slidingFrictionCoefficient " << cfg_.frictionCoefPart
              << " slidingFrictionCoefficientStatic " << 0 << "\n"
              << "Walls " << 0 << "\n"
              << "Boundaries " << 0 << "\n"
              << "Particles " << stdParticles + stdObjects << "\n";
            io.restartPrecision = 10;
            for (unsigned o = 0; o < stdObjects && o < objects.size(); ++o) {
                f << "BaseParticle id " << o << " " << "indSpecies " << 0 << " " << "position ";
                put3(f, objects[o].x0);
                f << "orientation " << 0 << " " << 0 << " " << 0 << " " << 0 << " " << "velocity ";
                put3(f, objects[o].x1);
                f << "angularVelocity " << zero << " " << zero << " " << zero << " " << 0 << " " << "force " << 0 << " "
                  << 0 << " " << 0 << " " << "torque " << 0 << " " << 0 << " " << 0 << " " << "radius "
                  << objects[o].r << " " << "invMass " << zero << " " << "invInertia " << zero << " " << "\n";
            }
            for (size_t i = 0; i < P; ++i) {
                f << "BaseParticle id " << i + stdObjects << " " << "indSpecies " << 0 << " " << "position ";
                put3(f, pos(i));
                f << "orientation " << 0 << " " << 0 << " " << 0 << " " << 0 << " " << "velocity ";
                put3(f, vel(i));
                f << "angularVelocity ";
                put3(f, wg(i));
                f << 0 << " " << "force " << 0 << " " << 0 << " " << 0 << " " << "torque " << 0 << " " << 0 << " " << 0
                  << " " << "radius " << elmts[cl(i)].radius << " " << "invMass " << 1 / misc[19 * cl(i)] << " "
                  << "invInertia " << 1 / misc[19 * cl(i) + 2] << " " << "\n";
            }
            f << "Interactions " << 0;
            f << "\n";
        }
        // ---- .fstat: header, then the contact lines of this sub-step's force evaluation
        {
            ofstream f(io.statisticFileName.c_str(), ios::app);
            const double r0 = E ? elmts[0].radius : 0.0;  // the reference reads particles[0]
            f << "#" << " " << setprecision(5) << time << " " << zero << endl;
            f << "#" << " " << zero << " " << zero << " " << zero << " " << size[0] << " " << size[1] << " " << size[2]
              << endl;
            f << "#" << " " << r0 << " " << r0 << " " << zero << " " << zero << " " << zero << " " << zero << endl;
            const std::vector<double> pr = saved_part(h_, ev, "pairs");
            for (size_t k = 0; k + 12 <= pr.size(); k += 12) {
                const double* r = &pr[k];
                const size_t a = (size_t)r[0], b = (size_t)r[1];
                if (r[11] == 0.0 || a >= P || b >= P) continue;
                stat_line(f, time, (long)(a + stdObjects), (long)(b + stdObjects), 0.5 * (pos(a) + pos(b)), r);
            }
            const std::vector<double> wl = saved_part(h_, ev, "walls");
            for (size_t k = 0; k + 12 <= wl.size(); k += 12) {
                const double* r = &wl[k];
                if (r[11] == 0.0) continue;
                const size_t e = (size_t)r[1];
                const tVect en(r[5], r[6], r[7]);
                const tVect centre = pos(e) - en * (elmts[cl(e)].radius - r[2]);
                stat_line(f, time, (long)(e + stdObjects), -(long)(r[0] + 1), centre, r);
            }
            const std::vector<double> ob = saved_part(h_, ev, "objects");
            const std::vector<double> orow = ob.empty() ? std::vector<double>() : field("objects");  // [O][8]
            for (size_t k = 0; k + 12 <= ob.size(); k += 12) {
                const double* r = &ob[k];
                if (r[11] == 0.0) continue;
                const size_t e = (size_t)r[1], o = (size_t)r[0];
                const tVect ox(orow[8 * o + 2], orow[8 * o + 3], orow[8 * o + 4]);
                stat_line(f, time, (long)orow[8 * o], (long)(e + stdObjects), 0.5 * (ox + pos(e)), r);
            }
        }
    }
}

void DEM::syncHost() {
    std::vector<double> x, w, fr, misc, en;
    if (!view(h_, hyb_dem_get, "e_x", x) || !view(h_, hyb_dem_get, "e_w", w) ||
        !view(h_, hyb_dem_get, "e_forces", fr) || !view(h_, hyb_dem_get, "e_misc", misc) ||
        !view(h_, hyb_dem_get, "energy", en) || en.size() != 6)
        hybird_check(HYB_E_STATE, "DEM::syncHost", __FILE__, __LINE__);
    particleEnergy.kin = en[0];
    particleEnergy.rotKin = en[1];
    particleEnergy.trKin = en[2];
    particleEnergy.grav = en[3];
    particleEnergy.elastic = en[4];
    particleEnergy.total = en[5];
    for (size_t e = 0; e < elmts.size(); ++e) {
        elmt& el = elmts[e];
        const double* X = &x[e * 18];
        el.x0 = tVect(X[0], X[1], X[2]);
        el.x1 = tVect(X[3], X[4], X[5]);
        el.w0 = tVect(w[e * 18], w[e * 18 + 1], w[e * 18 + 2]);
        const double* F = &fr[e * 21];
        el.FHydro = tVect(F[0], F[1], F[2]);
        el.MHydro = tVect(F[3], F[4], F[5]);
        el.FParticle = tVect(F[6], F[7], F[8]);
        el.MParticle = tVect(F[9], F[10], F[11]);
        el.FWall = tVect(F[12], F[13], F[14]);
        el.MWall = tVect(F[15], F[16], F[17]);
        el.FGrav = tVect(F[18], F[19], F[20]);
        el.m = misc[e * 19];
        el.fluidVolume = misc[e * 19 + 17];
        if (particles.size() == elmts.size()) {  // single spheres: the particle is its element
            particles[e].x0 = el.x0;
            particles[e].x1 = el.x1;
        }
    }
    if (particles.size() != elmts.size()) {  // spheres of multi-sphere elements: [S][10] x0, x1, radiusVec, r
        std::vector<double> pd;
        if (!view(h_, hyb_dem_get, "p_data", pd) || pd.size() < 10 * particles.size())
            hybird_check(HYB_E_STATE, "DEM::syncHost", __FILE__, __LINE__);
        for (size_t i = 0; i < particles.size(); ++i) {
            const double* r = &pd[10 * i];
            particles[i].x0 = tVect(r[0], r[1], r[2]);
            particles[i].x1 = tVect(r[3], r[4], r[5]);
            particles[i].radiusVec = tVect(r[6], r[7], r[8]);
        }
    }
}

std::vector<unsigned int> DEM::indexField(const char* name) const {
    std::vector<unsigned int> v;
    if (!view(h_, hyb_dem_get, name, v)) hybird_check(HYB_E_ARG, name, __FILE__, __LINE__);
    return v;
}

std::vector<double> DEM::field(const char* name) const {
    std::vector<double> v;
    if (!view(h_, hyb_dem_get, name, v)) hybird_check(HYB_E_ARG, name, __FILE__, __LINE__);
    return v;
}

// ======================================================================== IO
void IO::initialize() {  // IO.cpp:8-130 (status file and the fluid / particle data folders)
    if (!workDirectory.empty()) {
        const int a = mkdir(workDirectory.c_str(), 0777);
        if (a != 0 && errno != EEXIST) cout << "WARNING: cannot create " << workDirectory << "\n";
    }
    const std::string base = workDirectory.empty() ? std::string(".") : workDirectory;
    partDirectory = base + "/particleData";
    fluidDirectory = base + "/fluidData";
    if (demSolve) mkdir(partDirectory.c_str(), 0777);
    if (lbmSolve) mkdir(fluidDirectory.c_str(), 0777);
    fluidFileFormat = fluidDirectory + "/fluid%010u.vti";
    partFileFormat = partDirectory + "/part%010u.vtu";
    objectFileFormat = partDirectory + "/object%010u.vtu";
    exportFileName = base + "/export.dat";
    // energy.dat header; maxVel.dat, force.dat and plasticity.dat headers (IO.cpp:23-39, 105-123)
    energyFileName = base + "/energy.dat";
    {
        ofstream f(energyFileName.c_str(), ios::app);
        const int width = (int)f.precision() + 5;
        f << "\t" << " time";
        for (const char* h : {"ene_grav", "ene_kin", "ene_rot", "ene_ela", "X_com", "Y_com", "Z_com"})
            f << "\t" << setw(width) << h;
        f << "\n";
    }
    maxSpeedFileName = base + "/maxVel.dat";
    forceFileName = base + "/force.dat";
    plasticityFileName = base + "/plasticity.dat";
    { ofstream f(maxSpeedFileName.c_str(), ios::app); f << "time maxFluidSpeed\n"; }
    { ofstream f(forceFileName.c_str(), ios::app); f << "time collisionForces hydraulicForces\n"; }
    { ofstream f(plasticityFileName.c_str(), ios::app); f << "time percPlastic\n"; }
    restartFileName = base + "/" + problemNameString + ".restart";
    dataFileName = base + "/" + problemNameString + ".data";
    statisticFileName = base + "/" + problemNameString + ".fstat";
    lastScreenExp = lastFluidExp = lastPartExp = lastObjectExp = lastOutputExp = lastRecycleExp = 0;
    realTime = 0.0;
}

// IO.cpp:698-831 (exportParaviewFluidOld, the writer createParaviewFiles calls):
// ASCII ImageData with the node type (1 inside particles), velocity and
// pressure in physical units, viscosity (non-Newtonian), mass (free surface)
// and solidIndex (with DEM). Nodes without a node object in the reference
// (inactive ones) print zeros, as do the wall nodes it keeps (density rho0,
// zero velocity and mass). solidIndex: particle index on particle nodes, on wall
// nodes the last wall whose plane holds the node (LB.cpp:497-533), 0 elsewhere
// (the reference keeps the index of the last particle that covered a node).
void IO::exportParaviewFluid(const LB& lb, const DEM& dem, const string& fluidFile) const {
    const size_t N = lb.totNodes;
    const std::vector<unsigned char> t = lb.nodeTypes();
    const std::vector<unsigned char> pf = lb.nodeFlags("pflag");
    const std::vector<double> n = lb.field("n"), u = lb.field("u");
    std::vector<double> visc, mass;
    if (lb.nonNewtonian) visc = lb.field("visc");
    if (lb.freeSurface) mass = lb.field("mass");
    std::vector<unsigned int> sid;
    std::vector<double> obj;  // index, r, x0, x1 per object (with periodic copies)
    if (demSolve) {
        sid = lb.solidIndices();
        obj = dem.field("objects");
    }
    const auto active = [&](size_t i) { return t[i] == 0 || t[i] == 3; };
    ofstream f(fluidFile.c_str());
    const double zero = 0.0;
    f << "<VTKFile type=\"ImageData\" version=\"0.1\">\n";
    f << " <ImageData WholeExtent=\"0 " << lb.lbSize[0] - 1 << " 0 " << lb.lbSize[1] - 1 << " 0 " << lb.lbSize[2] - 1
      << "\" " << "Origin=\"0.0 0.0 0.0\" Spacing=\"" << lb.unit.Length << " " << lb.unit.Length << " "
      << lb.unit.Length << "\">\n";
    f << "  <Piece Extent=\"0 " << lb.lbSize[0] - 1 << " 0 " << lb.lbSize[1] - 1 << " 0 " << lb.lbSize[2] - 1 << "\">\n";
    f << "   <PointData>\n";
    f << "    <DataArray type=\"Int8\" Name=\"type\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"8\">\n";
    for (size_t i = 0; i < N; ++i) f << ((active(i) && pf[i]) ? 1u : (unsigned int)t[i]) << " ";
    f << "    </DataArray>\n";
    f << "    <DataArray type=\"Float64\" Name=\"v\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (size_t i = 0; i < N; ++i) {
        if (!active(i)) {
            f << zero << " " << zero << " " << zero << " ";
        } else {
            const tVect uPhysic = tVect(u[3 * i], u[3 * i + 1], u[3 * i + 2]) * lb.unit.Speed;
            f << uPhysic.x << " " << uPhysic.y << " " << uPhysic.z << " ";
        }
    }
    f << "    </DataArray>\n";
    f << "    <DataArray type=\"Float64\" Name=\"pressure\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"2\">\n";
    for (size_t i = 0; i < N; ++i) {
        if (!active(i)) f << zero << " ";
        else if (n[i] != 0.0) f << 0.3333333 * (n[i] - 1.0) * lb.unit.Pressure << " ";
        else f << n[i] << " ";
    }
    f << "    </DataArray>\n";
    if (lb.nonNewtonian) {
        f << "    <DataArray type=\"Float64\" Name=\"dynVisc\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"2\">\n";
        for (size_t i = 0; i < N; ++i) f << (active(i) ? visc[i] * lb.unit.DynVisc : zero) << " ";
        f << "    </DataArray>\n";
    }
    if (lb.freeSurface) {
        f << "    <DataArray type=\"Float64\" Name=\"AAAmass\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"2\">\n";
        for (size_t i = 0; i < N; ++i) f << (active(i) ? mass[i] * lb.unit.Density : zero) << " ";
        f << "    </DataArray>\n";
    }
    if (demSolve) {
        f << "    <DataArray type=\"Int16\" Name=\"solidIndex\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"8\">\n";
        for (size_t i = 0; i < N; ++i) {
            unsigned int s = (active(i) && pf[i]) ? sid[i] : 0u;
            if (t[i] >= 5 && t[i] <= 8) {
                const tVect pos(lb.getX(i), lb.getY(i), lb.getZ(i));
                for (const wall& w : dem.walls) {
                    const tVect d = pos - w.p / lb.unit.Length;
                    if (w.n.x * d.x + w.n.y * d.y + w.n.z * d.z < 0.0) s = w.index;
                }
                for (size_t o = 0; o + 8 <= obj.size(); o += 8) {  // LB.cpp:535-550
                    const tVect d = pos - tVect(obj[o + 2], obj[o + 3], obj[o + 4]) / lb.unit.Length;
                    const double rr = obj[o + 1] / lb.unit.Length;
                    if (d.x * d.x + d.y * d.y + d.z * d.z < rr * rr) s = (unsigned int)obj[o];
                }
            }
            f << s << " ";
        }
        f << "    </DataArray>\n";
    }
    f << "   </PointData>\n";
    f << "   <CellData>\n";
    f << "   </CellData>\n";
    f << "  </Piece>\n";
    f << " </ImageData>\n";
    f << "</VTKFile>\n";
}

// IO.cpp:289-419: one point per particle (ghosts included) with its element's
// mass, velocities and forces.
// IO.cpp:422-480: one vertex per object (periodic copies included) with radius,
// index, velocity and the force of the last sub-step's contacts (FParticle).
void IO::exportParaviewObjects(const DEM& dem, const string& objectFile) const {
    const std::vector<double> ob = dem.field("objects");       // [O][8]: index, r, x0, x1
    const std::vector<double> fo = dem.field("objectForces");  // [O][3]
    const size_t O = ob.size() / 8;
    auto line = [](ofstream& f, const double* v) { f << v[0] << " " << v[1] << " " << v[2] << "\n"; };
    ofstream f(objectFile.c_str());
    f << "<?xml version=\"1.0\"?>\n";
    f << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
         "compressor=\"vtkZLibDataCompressor\">\n";
    f << " <UnstructuredGrid GhostLevel=\"0\">\n";
    f << "  <Piece NumberOfPoints=\"" << O << "\" NumberOfCells=\"" << O << "\">\n";
    f << "   <PointData>\n";
    f << "    <DataArray type=\"Float64\" Name=\"radius\"/>\n";
    for (size_t i = 0; i < O; ++i) f << ob[8 * i + 1] << "\n";
    f << "    <DataArray type=\"Float64\" Name=\"objectIndex\"/>\n";
    for (size_t i = 0; i < O; ++i) f << (unsigned int)ob[8 * i] << "\n";
    f << "    <DataArray type=\"Float64\" Name=\"v\" NumberOfComponents=\"3\"/>\n";
    for (size_t i = 0; i < O; ++i) line(f, &ob[8 * i + 5]);
    f << "    <DataArray type=\"Float64\" Name=\"Force\" NumberOfComponents=\"3\"/>\n";
    for (size_t i = 0; i < O; ++i) line(f, &fo[3 * i]);
    f << "   </PointData>\n";
    f << "   <CellData>\n";
    f << "   </CellData>\n";
    f << "   <Points>\n";
    f << "    <DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\"/>\n";
    for (size_t i = 0; i < O; ++i) line(f, &ob[8 * i + 2]);
    f << "   </Points>\n";
    f << "   <Cells>\n";
    f << "    <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
    for (size_t i = 1; i < O + 1; ++i) f << i - 1 << "\n";
    f << "    </DataArray>\n";
    f << "    <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
    for (size_t i = 1; i < O + 1; ++i) f << i << "\n";
    f << "    </DataArray>\n";
    f << "    <DataArray type=\"Int64\" Name=\"types\" format=\"ascii\">\n";
    for (size_t i = 0; i < O; ++i) f << 1 << "\n";
    f << "    </DataArray>\n";
    f << "   </Cells>\n";
    f << "  </Piece>\n";
    f << " </UnstructuredGrid>\n";
    f << "</VTKFile>";
}

void IO::exportParaviewParticles(const DEM& dem, const string& particleFile) const {
    const std::vector<double> pd = dem.field("p_data");       // [S][10]: x0, x1, radiusVec, r
    const std::vector<unsigned int> pi = dem.indexField("p_idx");  // [S][4]: particle, cluster, proto, cell
    const std::vector<double> misc = dem.field("e_misc");     // [E][19]: m, radius, I, wGlobal, ...
    const std::vector<double> fr = dem.field("e_forces");     // [E][7][3]
    const size_t S = pd.size() / 10;
    ofstream f(particleFile.c_str());
    const auto vec = [&](const double* v) { f << v[0] << " " << v[1] << " " << v[2] << "\n"; };
    const auto cl = [&](size_t i) { return (size_t)pi[4 * i + 1]; };
    f << "<?xml version=\"1.0\"?>\n";
    f << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" compressor=\"vtkZLibDataCompressor\">\n";
    f << " <UnstructuredGrid GhostLevel=\"0\">\n";
    f << "  <Piece NumberOfPoints=\"" << S << "\" NumberOfCells=\"" << S << "\">\n";
    f << "   <PointData>\n";
    f << "    <DataArray type=\"Float64\" Name=\"radius\"/>\n";
    for (size_t i = 0; i < S; ++i) f << pd[10 * i + 9] << "\n";
    f << "    <DataArray type=\"Float64\" Name=\"mass\"/>\n";
    for (size_t i = 0; i < S; ++i) f << misc[19 * cl(i)] << "\n";
    f << "    <DataArray type=\"Float64\" Name=\"particleIndex\"/>\n";
    for (size_t i = 0; i < S; ++i) f << pi[4 * i] << "\n";
    f << "    <DataArray type=\"Float64\" Name=\"clusterIndex\"/>\n";
    for (size_t i = 0; i < S; ++i) f << pi[4 * i + 1] << "\n";
    f << "    <DataArray type=\"Float64\" Name=\"v\" NumberOfComponents=\"3\"/>\n";
    for (size_t i = 0; i < S; ++i) vec(&pd[10 * i + 3]);
    f << "    <DataArray type=\"Float64\" Name=\"w\" NumberOfComponents=\"3\"/>\n";
    for (size_t i = 0; i < S; ++i) vec(&misc[19 * cl(i) + 5]);
    // e_forces order: FHydro, MHydro, FParticle, MParticle, FWall, MWall, FGrav
    const int order[6] = {2, 6, 3, 4, 5, 0};
    const char* names[6] = {"FParticle", "FGrav", "MParticle", "FWall", "MWall", "FHydro"};
    for (int k = 0; k < 6; ++k) {
        f << "    <DataArray type=\"Float64\" Name=\"" << names[k] << "\" NumberOfComponents=\"3\"/>\n";
        for (size_t i = 0; i < S; ++i) vec(&fr[21 * cl(i) + 3 * order[k]]);
    }
    f << "   </PointData>\n";
    f << "   <CellData>\n";
    f << "   </CellData>\n";
    f << "   <Points>\n";
    f << "    <DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\"/>\n";
    for (size_t i = 0; i < S; ++i) vec(&pd[10 * i]);
    f << "   </Points>\n";
    f << "   <Cells>\n";
    f << "    <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
    for (size_t i = 1; i < S + 1; ++i) f << i - 1 << "\n";
    f << "    </DataArray>\n";
    f << "    <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
    for (size_t i = 1; i < S + 1; ++i) f << i << "\n";
    f << "    </DataArray>\n";
    f << "    <DataArray type=\"Int64\" Name=\"types\" format=\"ascii\">\n";
    for (size_t i = 0; i < S; ++i) f << 1 << "\n";
    f << "    </DataArray>\n";
    f << "   </Cells>\n";
    f << "  </Piece>\n";
    f << " </UnstructuredGrid>\n";
    f << "</VTKFile>";
}

// IO.cpp:486-535: the element file format the code reads back (DEM.cpp:78-127),
// positions wrapped into the periodic box.
void IO::exportRecycleParticles(const DEM& dem) const {
    std::stringstream ss;
    ss << currentTimeStep;
    const string path = partDirectory + "/recycle_" + ss.str() + ".dat";
    const std::vector<double> x = dem.field("e_x"), w = dem.field("e_w"), q = dem.field("e_q"),
                              misc = dem.field("e_misc");
    const size_t E = dem.elmts.size();
    ofstream f(path.c_str());
    f << E << "\n";
    const auto v3 = [&](const double* v) { f << v[0] << " " << v[1] << " " << v[2] << " "; };
    for (size_t i = 0; i < E; ++i) {
        f << dem.elmts[i].index << " ";
        f << dem.elmts[i].size << " ";
        f << misc[19 * i + 1] << " ";
        const tVect x0(x[18 * i], x[18 * i + 1], x[18 * i + 2]);
        tVect printPosition = x0;
        for (size_t b = 0; b < dem.pbcs.size(); ++b) {
            const pbc& pb = dem.pbcs[b];
            const tVect n1 = pb.v / pb.v.norm();
            const tVect d1 = x0 - pb.p1, d2 = x0 - pb.p2;
            const double leftDist = n1.x * d1.x + n1.y * d1.y + n1.z * d1.z;
            const tVect n2 = n1 * -1.0;
            const double rightDist = n2.x * d2.x + n2.y * d2.y + n2.z * d2.z;
            if (leftDist < 0.0) printPosition = printPosition + pb.v;
            if (rightDist < 0.0) printPosition = printPosition - pb.v;
        }
        f << printPosition.x << " " << printPosition.y << " " << printPosition.z << " ";
        v3(&x[18 * i + 3]);
        v3(&w[18 * i]);
        const double* q0 = &q[24 * i];
        const double* q1 = &q[24 * i + 4];
        f << q0[0] << " " << q0[1] << " " << q0[2] << " " << q0[3] << " ";
        f << q1[0] << " " << q1[1] << " " << q1[2] << " " << q1[3] << "\n";
    }
}

void IO::outputStep(const double& demTime, const LB& lb, const DEM& dem) {  // IO.cpp:132-198
    realTime = demTime;
    // createParaviewFiles (IO.cpp:250-287)
    char filePathBuffer[1024];
    if (lbmSolve) {
        const unsigned int c = (fluidExpTime > 0 ? static_cast<unsigned int>(realTime / fluidExpTime) + 1 : 0);
        if (c > lastFluidExp) {
            lastFluidExp = c;
            std::snprintf(filePathBuffer, sizeof(filePathBuffer), fluidFileFormat.c_str(), currentTimeStep);
            exportParaviewFluid(lb, dem, filePathBuffer);
        }
    }
    if (demSolve) {
        const unsigned int c = (partExpTime > 0 ? static_cast<unsigned int>(realTime / partExpTime) + 1 : 0);
        if (c > lastPartExp) {
            lastPartExp = c;
            std::snprintf(filePathBuffer, sizeof(filePathBuffer), partFileFormat.c_str(), currentTimeStep);
            exportParaviewParticles(dem, filePathBuffer);
        }
        const unsigned int r = (recycleExpTime > 0 ? static_cast<unsigned int>(realTime / recycleExpTime) + 1 : 0);
        if (r > lastRecycleExp) {
            lastRecycleExp = r;
            exportRecycleParticles(dem);
        }
        const unsigned int o = (objectExpTime > 0 ? static_cast<unsigned int>(realTime / objectExpTime) + 1 : 0);
        if (o > lastObjectExp) {
            lastObjectExp = o;
            std::snprintf(filePathBuffer, sizeof(filePathBuffer), objectFileFormat.c_str(), currentTimeStep);
            exportParaviewObjects(dem, filePathBuffer);
        }
    }
    const unsigned int screenExpCounter =
        (screenExpTime > 0 ? static_cast<unsigned int>(realTime / screenExpTime) + 1 : 0);
    if (screenExpCounter > lastScreenExp) {
        lastScreenExp = screenExpCounter;
        exportFile.open(exportFileName.c_str(), ios::app);
        cout << currentTimeStep << "; time=" << realTime << "\t";
        exportFile << currentTimeStep << "; time=" << realTime << "\t";
        std::vector<double> st;
        if (lbmSolve) {
            st = lb.statusValues();
            // exportMaxSpeedFluid (IO.cpp:835-850)
            const double maxFluidSpeed = sqrt(st[0]);
            cout << "MaxFSpeed= " << maxFluidSpeed * lb.unit.Speed << "("
                 << int(1.0 * 1.0 * maxFluidSpeed * 100.0 / 0.01) << "%)\t";
            exportFile << "MaxFSpeed= " << maxFluidSpeed * lb.unit.Speed << "("
                       << int(1.0 * 1.0 * maxFluidSpeed * 100.0 / 0.01) << "%)\t";
            {
                ofstream f(maxSpeedFileName.c_str(), ios::app);
                f << realTime << " " << maxFluidSpeed * lb.unit.Speed << "\n";
            }
            // exportTotalMass (IO.cpp:878-883)
            const double massTot = lb.totalFluidMass();
            cout << "Volume=" << massTot * lb.unit.Volume << "; Mass = " << massTot * lb.unit.Mass << "\t";
            exportFile << "Volume=" << massTot * lb.unit.Volume << "; Mass = " << massTot * lb.unit.Mass << "\t";
        }
        if (lbmSolve & lb.nonNewtonian) {  // exportPlasticity (IO.cpp:885-895, 969-985)
            const double perc = 100.0 * st[2] / st[1];
            cout << "Plastic =" << int(perc) << "%;\t";
            exportFile << "Plastic =" << int(perc) << "%;\t";
            ofstream f(plasticityFileName.c_str(), ios::app);
            f << realTime << " " << perc << "\n";
        }
        if (dem.elmts.size()) {
            // exportMaxSpeedParticles (IO.cpp:852-876) and exportForces (IO.cpp:930-938)
            const std::vector<double> x = dem.field("e_x"), misc = dem.field("e_misc"), fr = dem.field("e_forces");
            const size_t E = dem.elmts.size();
            double maxPartSpeed = 0.0, maxPartSpin = 0.0, collisionForce = 0.0, hydraulicForce = 0.0;
            for (size_t e = 0; e < E; ++e)
                maxPartSpeed = std::max(maxPartSpeed, tVect(x[18 * e + 3], x[18 * e + 4], x[18 * e + 5]).norm2());
            maxPartSpeed = sqrt(maxPartSpeed);
            cout << "MaxPSpeed= " << maxPartSpeed << "\t";
            exportFile << "MaxPSpeed= " << maxPartSpeed << "\t";
            for (size_t e = 0; e < E; ++e)
                maxPartSpin = std::max(maxPartSpin, tVect(misc[19 * e + 5], misc[19 * e + 6], misc[19 * e + 7]).norm2());
            maxPartSpin = sqrt(maxPartSpin);
            cout << "MaxPSpin= " << maxPartSpin << "\t";
            exportFile << "MaxPSpin= " << maxPartSpin << "\t";
            cout << "\t";
            for (size_t e = 0; e < E; ++e) {
                const double* f = &fr[21 * e];
                collisionForce += tVect(f[6], f[7], f[8]).norm();
                collisionForce += tVect(f[12], f[13], f[14]).norm();
            }
            for (size_t e = 0; e < E; ++e) hydraulicForce += tVect(fr[21 * e], fr[21 * e + 1], fr[21 * e + 2]).norm();
            ofstream f(forceFileName.c_str(), ios::app);
            f << realTime << " " << collisionForce << " " << hydraulicForce << "\n";
        }
        if (dem.elmts.size() && !lbmSolve) {  // exportEnergy (IO.cpp:919-927)
            const std::vector<double> en = dem.field("energy");
            cout << "eKin = " << en[0] << "\t";
            exportFile << "eKin = " << en[0] << "\t";
            cout << "eGrav = " << en[3] << "\t";
            exportFile << "eGrav = " << en[3] << "\t";
            cout << "eTot = " << en[5] << "\t";
            exportFile << "eTot = " << en[5] << "\t";
        }
        cout << "\n";
        exportFile << "\n";
        exportFile.close();
        cout.flush();
    }
}

void IO::outputFinal() { cout << "Simulation finished at step " << currentTimeStep << "\n"; }
