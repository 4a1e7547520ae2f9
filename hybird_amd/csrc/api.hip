// C ABI of libhybird_amd.so (declared in include/hybird_amd.h).
//
// The host side of each entry point restates the corresponding reference
// method's set-up logic (parsing conversions, wall and periodic planes, element
// mass and inertia, neighbour-cell sizing) and then drives the HIP kernels of
// lb.hip and dem.hip on one stream per device. Nothing here synchronises with
// the device except the host views, hyb_sync and the timing queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "adj_kernels.h"
#include "dem_kernels.h"
#include "fs_kernels.h"
#include "hybird_amd.h"
#include "lb_kernels.h"

using namespace hyb;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail(HYB_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Value of an exact fixed-point sum (four 32-bit limbs summed in 64-bit words,
// LSB 2^-96), the host form of the device's exact_value.
double exact_value_host(const unsigned long long* acc) {
    long long a[4];
    for (int i = 0; i < 4; ++i) a[i] = (long long)acc[i];
    for (int i = 0; i < 3; ++i) {
        const long long carry = a[i] >> 32;
        a[i] -= carry * 4294967296LL;
        a[i + 1] += carry;
    }
    return (((double)a[0] * 0x1p-96 + (double)a[1] * 0x1p-64) + (double)a[2] * 0x1p-32) + (double)a[3];
}

hipStream_t shared_stream() {
    static hipStream_t s = nullptr;
    if (!s) (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    return s;
}

Units make_units(double L, double T, double D) {  // node.cpp:476-488
    Units u;
    u.Length = L;
    u.Time = T;
    u.Density = D;
    u.Volume = L * L * L;
    u.Speed = L / T;
    u.Accel = L / T / T;
    u.AngVel = 1.0 / T;
    u.KinVisc = L * L / T;
    u.DynVisc = D * L * L / T;
    u.Force = D * L * L * L * L / T / T;
    u.Torque = D * L * L * L * L * L / T / T;
    u.Mass = D * L * L * L;
    u.Stress = D * L * L / T / T;
    return u;
}

template <class T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return HYB_OK;
    HIPCHK(hipMalloc((void**)p, count * sizeof(T)));
    return HYB_OK;
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// A device buffer that belongs to a handle's persistent state (written by
// hyb_state_save, restored by hyb_state_load).
struct Blk {
    std::string name;
    void* ptr;
    size_t bytes;
};

// dalloc that also records the buffer in the handle's state list.
template <class T>
int ralloc(std::vector<Blk>& reg, const char* name, T** p, size_t count) {
    const int rc = dalloc(p, count);
    if (rc == HYB_OK && *p) reg.push_back(Blk{name, (void*)*p, count * sizeof(T)});
    return rc;
}

// Error word of a handle: pinned host memory mapped into the device, so that a
// kernel that finds a failure can store its code where the host reads it
// without synchronising (see the device-side failures in hybird_amd.h).
struct ErrWord {
    int* host = nullptr;  // host address
    int* dev = nullptr;   // device address of the same word
};

int errword_alloc(ErrWord& w) {
    if (w.host) return HYB_OK;
    void* h = nullptr;
    HIPCHK(hipHostMalloc(&h, 64, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(h, 0, 64);
    void* dp = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&dp, h, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(h);
        return fail(HYB_E_HIP, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
    }
    w.host = (int*)h;
    w.dev = (int*)dp;
    return HYB_OK;
}

void errword_free(ErrWord& w) {
    if (w.host) (void)hipHostFree(w.host);
    w.host = w.dev = nullptr;
}

int errword_code(const ErrWord& w) { return w.host ? __atomic_load_n(w.host, __ATOMIC_ACQUIRE) : 0; }

void errword_clear(ErrWord& w) {
    if (w.host) __atomic_store_n(w.host, 0, __ATOMIC_RELEASE);
}

void unregister(std::vector<Blk>& reg, const void* ptr) {
    for (size_t i = 0; i < reg.size(); ++i)
        if (reg[i].ptr == ptr) {
            reg.erase(reg.begin() + (long)i);
            return;
        }
}

}  // namespace

// ------------------------------------------------------------------ handles
struct hyb_dem {
    hyb_config cfg;
    int E = 0;  // elements (particle-file records)
    int P = 0;  // standard particles: the spheres of all elements
    std::vector<double> rec;
    // material (DEM.cpp:15-61)
    double knConst = 0, ksConst = 0, dampCoeff = 0;
    Units unit;
    double demTime = 0.0, deltat = 1.0, demInitialRepeat = 0.0;
    uint32_t demTimeStep = 0;
    int multiStep = 1;
    double demSize[3] = {1, 1, 1};
    std::vector<WallDev> walls;
    std::vector<PbcDev> pbcs;
    std::vector<ObjDev> objects;  // object file entries, then periodic copies (set at init)
    std::vector<CylDev> cylinders;  // problem cylinders (DEM.cpp:870-935)
    int stdObjects = 0;
    DEMParams p;
    DEMBuffers b;
    bool inited = false;
    hipStream_t stream = nullptr;
    int64_t substeps = 0;
    ErrWord err;             // device-side failures of the DEM kernels
    std::vector<Blk> saved;  // persistent device state
    // MercuryDPM-format saving (DEM.cpp:355-360, 1131-1146): captures of the
    // saving sub-steps of the last hyb_dem_step, one slot each
    int saveCount = 0;
    int saveSlots = 0;
    double* saveElems = nullptr;    // [slot][P][9]
    double* saveInfo = nullptr;     // [slot]: pairs of the sub-step
    StatRec* savePairs = nullptr;   // [slot][pairCap]
    StatRec* saveWalls = nullptr;   // [slot][P]
    StatRec* saveObjs = nullptr;    // [slot][P][NEAR_OBJ] (objects only)
    std::vector<std::pair<double, uint32_t>> saveEvents;  // demTime, demTimeStep of each capture
};

struct hyb_lb {
    hyb_config cfg;
    Units unit;
    int X = 1, Y = 1, Z = 1;
    size_t N = 1;
    double initDynVisc = 0, plasticVisc = 0, yieldStress = 0, turbConst = 0, slipCoefficient = 0, initDensity = 1.0;
    double initVelocity[3] = {0, 0, 0}, lbF[3] = {0, 0, 0};
    int boundary[6] = {1, 1, 1, 1, 1, 1};
    bool freeSurface = false, forceField = false, nonNewtonian = false, turbulenceOn = false;
    bool frictional = false;      // frictional apparent viscosity (extension)
    double frictionFluid = 0.0;
    Prec prec = Prec::F64;
    hipStream_t stream = nullptr;
    // device state
    void* f[2] = {nullptr, nullptr};
    int cur = 0;
    uint8_t* type = nullptr;
    uint32_t* link = nullptr;
    uint32_t* slipw = nullptr;  // slip-wall directions per node (only when a box face is a slip wall)
    // cylinders (curved walls): curved-link directions per node, n and u of the last
    // two collisions ([4][N] each, ping-pong: macBuf[macCur] holds the last one)
    uint32_t* curvw = nullptr;
    double* macBuf[2] = {nullptr, nullptr};
    int macCur = 0;
    CylDev cyl;
    uint32_t* solid = nullptr;
    // per-slot hydrodynamic run sums appended by the collision (segCap per slot)
    HydroSeg* segs = nullptr;
    uint32_t* segCount = nullptr;
    uint32_t segCap = 0;
    double* visc = nullptr;
    double* mac = nullptr;
    void* scratch = nullptr;
    int* maxXYZ = nullptr;
    unsigned long long* dcount = nullptr;
    CoupleRec* cp = nullptr;
    CoupleRecF* cpf = nullptr;
    NodeBox* boxes = nullptr;
    int* nslots = nullptr;  // [0] slots this step, [1] boxes marked last step
    int slotCap = 0;
    uint32_t ctag = 0;       // coupling-step tag of the node ownership words
    hyb_dem* dem = nullptr;  // DEM handle coupled at init (read by the host views)
    int64_t steps = 0;
    bool inited = false;
    bool macValid = false;
    bool havePart = false;
    unsigned long long activeCount = 0;
    double totalMass = 0.0;
    uint32_t time = 0;
    // profiling of the fused kernel
    // free surface (LB.cpp:1492-1804)
    FSBuffers fsb;
    int fsCur = 0;         // current interface list
    int rhoCur = 0;        // density buffer of the last collision
    int fluidMassBuf = 0;  // density buffer that is the fluid-node mass (set by the free-surface step)
    uint32_t fsTag = 0;
    uint32_t actTag = 0;   // tag of a free-surface pass not yet followed by a collision
    int64_t fsSteps = 0;
    bool capture = false;  // keep n, u, hydroForce of every collision for the host views
    float* ffield = nullptr;  // optional per-node body force [3][N] (lattice units)
    // adjoint: checkpoints of the post-collision state every adjEvery steps, the
    // reverse pass's buffers (states of one interval, two adjoint population
    // buffers, dL/dF, tau partials) and the device time of both loops
    int adjSteps = -1, adjEvery = 0;
    std::vector<void*> adjCk;
    std::vector<void*> adjSeg;
    double* adjLam[2] = {nullptr, nullptr};
    double* adjGF = nullptr;
    double* adjTP = nullptr;
    double adjFwdMs = 0.0, adjBwdMs = 0.0;
    hipEvent_t adjEv[2] = {nullptr, nullptr};
    bool profiling = false;
    std::vector<hipEvent_t> evs;
    size_t evUsed = 0;
    double profMs = 0.0;
    int64_t profLaunches = 0;
    ErrWord err;             // device-side failures of the LB kernels
    std::vector<Blk> saved;  // persistent device state

    bool varVisc() const { return nonNewtonian || turbulenceOn || frictional; }
    bool hasSlip() const {
        for (int k = 0; k < 6; ++k)
            if (boundary[k] == 5 || boundary[k] == 6) return true;
        return false;
    }
    LBKParams params() const {
        LBKParams p;
        p.X = X; p.Y = Y; p.Z = Z;
        p.PS = planeStride();
        p.per[0] = boundary[0] == 4; p.per[1] = boundary[2] == 4; p.per[2] = boundary[4] == 4;
        // collision() zeroes the force permanently when forceFieldSolve = 0 (LB.cpp:1074-1076)
        for (int k = 0; k < 3; ++k) p.F[k] = forceField ? lbF[k] : 0.0;
        p.visc0 = initDynVisc;
        p.plasticVisc = plasticVisc; p.yieldStress = yieldStress; p.turbConst = turbConst;
        p.nonNewtonian = nonNewtonian; p.turbulence = turbulenceOn;
        p.frictional = frictional; p.frictionFluid = frictionFluid;
        p.collision = cfg.collision;
        p.trtMagic = cfg.trtMagic;
        p.unitLength = unit.Length; p.unitSpeed = unit.Speed; p.unitAngVel = unit.AngVel;
        p.wallSides = 0;
        for (int k = 0; k < 6; ++k)
            if (boundary[k] >= 5 && boundary[k] <= 8) p.wallSides |= 1 << k;
        p.initDensity = initDensity;
        p.ctag = ctag;
        p.ffield = ffield;
        p.slipS1 = slipCoefficient;  // LB.cpp:1360-1361
        p.slipS2 = 1.0 - slipCoefficient;
        p.slipw = slipw;
        p.linkw = link;
        p.actPrev = fsb.actPrev;
        p.actTag = actTag;
        // curved walls: the pull reads the last collision's n and u, the collision
        // writes its own into the other buffer
        p.curvw = curvw;
        p.macPrev = macBuf[macCur];
        p.macNext = macBuf[macCur ^ 1];
        p.typew = type;
        p.viscw = varVisc() ? visc : nullptr;
        p.cyl = cyl;
        p.herr = err.dev;
        return p;
    }
    // planes are padded so that successive planes do not start on the same
    // memory-channel boundary: N rounded up to 256 elements plus planePad
    // (256 elements: 1 KiB in FP32, 2 KiB in FP64; measured on the 256^3 dense
    // suspension, both fused kernels run about 3% faster than with unpadded planes)
    size_t planePad = 256;
    size_t planeStride() const { return ((N + 255) / 256) * 256 + planePad; }
    size_t popBytes() const { return (prec == Prec::F64 ? 8 : 4) * 19 * planeStride(); }
};

extern "C" {

const char* hyb_last_error(void) { return g_err.c_str(); }

// The pending device-side failure of a handle, if any, as a return code.
static int dem_pending(const hyb_dem* d) {
    const int e = d ? errword_code(d->err) : 0;
    if (!e) return HYB_OK;
    if (e == DEM_ERR_NUMERIC)
        return fail(HYB_E_NUMERIC, "DEM: non-finite element velocity; the DEM stopped at that sub-step");
    return fail(HYB_E_CAPACITY, "DEM device error " + std::to_string(e) +
                                    " (1/4: ghost or pair capacity, 2: particle outside the cell grid, 3: stencil outside"
                                    " the cell table, 5: more than 16 objects near one particle); the DEM stopped at"
                                    " that sub-step");
}

static int lb_pending(const hyb_lb* l) {
    const int e = l ? errword_code(l->err) : 0;
    if (!e) return HYB_OK;
    if (e == LB_ERR_SEGMENTS)
        return fail(HYB_E_CAPACITY, "LB: a particle slot's hydrodynamic run list overflowed its capacity");
    return fail(HYB_E_NUMERIC, "LB: non-finite density in a collision");
}

int hyb_device_count(int* count) {
    if (!count) return fail(HYB_E_ARG, "null count");
    HIPCHK(hipGetDeviceCount(count));
    return HYB_OK;
}

void hyb_config_defaults(hyb_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->problem = HYB_PROBLEM_NONE;
    c->unitLength = c->unitTime = c->unitDensity = 1.0;
    c->lbSize[0] = c->lbSize[1] = c->lbSize[2] = 16;
    c->plasticVisc = 1.0;
    c->yieldStress = 1.0;
    c->initVisc = 1.0;
    for (int k = 0; k < 6; ++k) c->boundary[k] = 1;
    c->density = 1.0;
    c->contactModel = HYB_CONTACT_HERTZIAN;
    c->youngMod = 1.0;
    c->poisson = 1.0;
    c->linearStiff = 1.0;
    c->restitution = 1.0;
    c->viscTang = 1.0;
    c->frictionCoefPart = 1.0;
    c->frictionCoefWall = 1.0;
    c->multiStep = 1;
    c->criticalRatio = 0.1;
    c->scale = 1.0;
    c->collision = HYB_COLLISION_BGK;
    c->trtMagic = 0.25;
    c->precision = HYB_FP64;
}

// ------------------------------------------------------------------ DEM
int hyb_dem_create(const hyb_config* cfg, int32_t n, const double* rec, hyb_dem** out) {
    if (!cfg || !out || n < 0 || (n > 0 && !rec)) return fail(HYB_E_ARG, "hyb_dem_create: bad arguments");
    hyb_dem* d = new hyb_dem();
    d->cfg = *cfg;
    // DEM.cpp:15-61 material constants; the HERTZIAN case falls through to the
    // LINEAR damping formula (with +pi), which is what the reference uses.
    if (!(cfg->density > 0.0) || !(cfg->restitution > 0.0) || cfg->restitution > 1.0 || !(cfg->youngMod > 0.0)) {
        delete d;
        return fail(HYB_E_ARG, "invalid material parameters (density, restitution, youngMod)");
    }
    d->knConst = 2.0 / 3.0 * cfg->youngMod / (1 - cfg->poisson * cfg->poisson);
    d->ksConst = 2.0 * cfg->youngMod / (2.0 - cfg->poisson) / (1.0 + cfg->poisson);
    const double le = std::log(cfg->restitution);
    d->dampCoeff = -1.0 * std::sqrt(2.0) * le / std::sqrt((le * le + M_PI));
    if (!(d->dampCoeff < 1.0)) { delete d; return fail(HYB_E_ARG, "dampCoeff >= 1"); }
    d->E = n;
    d->P = 0;
    d->rec.assign(rec, rec + (size_t)n * HYB_ELEMENT_RECORD);
    for (int e = 0; e < n; ++e) {
        const double* r = rec + (size_t)e * HYB_ELEMENT_RECORD;
        if ((int)r[0] != e) { delete d; return fail(HYB_E_ARG, "element indices must be 0..n-1 in file order"); }
        // the clump prototypes exist for 1 to 4 spheres (DEM.cpp:404-433)
        if (!(r[1] >= 1.0 && r[1] <= 4.0) || r[1] != (double)(int)r[1]) {
            delete d;
            return fail(HYB_E_ARG, "element size must be 1, 2, 3 or 4 spheres");
        }
        d->P += (int)r[1];
    }
    d->multiStep = cfg->multiStep;
    d->demInitialRepeat = cfg->demInitialRepeat;
    d->stream = shared_stream();
    *out = d;
    return HYB_OK;
}

int hyb_dem_set_objects(hyb_dem* d, int32_t n, const double* rec) {  // DEM.cpp:131-173
    if (!d || n < 0 || (n > 0 && !rec)) return fail(HYB_E_ARG, "hyb_dem_set_objects: bad arguments");
    if (d->inited) return fail(HYB_E_STATE, "hyb_dem_set_objects after init");
    d->objects.clear();
    for (int o = 0; o < n; ++o) {
        const double* r = rec + (size_t)o * HYB_OBJECT_RECORD;
        // periodic copies index the list with the object's index (DEM.cpp:1049-1050)
        if ((int)r[0] != o) return fail(HYB_E_ARG, "object indices must be 0..n-1 in file order");
        if (!(r[2] > 0.0)) return fail(HYB_E_ARG, "object radius must be > 0");
        ObjDev ob;
        ob.index = o;
        ob.r = r[2];
        ob.x0 = V3(r[3], r[4], r[5]);
        ob.x1 = V3(r[6], r[7], r[8]);
        d->objects.push_back(ob);
    }
    d->stdObjects = n;
    return HYB_OK;
}

// What DEM::discreteElementInit receives from the lattice (DEM.cpp:185-297):
// boundary codes, node counts, units and the body force in lattice units.
struct LatticeInfo {
    int boundary[6];
    int X, Y, Z;
    Units unit;
    double lbF[3];
};

static int dem_init(hyb_dem* d, const LatticeInfo& lb);

int hyb_dem_init(hyb_dem* d, const hyb_lb* lb) {
    if (!d || !lb) return fail(HYB_E_ARG, "hyb_dem_init: null handle");
    LatticeInfo li;
    for (int k = 0; k < 6; ++k) li.boundary[k] = lb->boundary[k];
    li.X = lb->X; li.Y = lb->Y; li.Z = lb->Z;
    li.unit = lb->unit;
    for (int k = 0; k < 3; ++k) li.lbF[k] = lb->lbF[k];
    return dem_init(d, li);
}

int hyb_dem_init_lattice(hyb_dem* d, const int32_t boundary[6], const int32_t lbSize[3], double unitLength,
                         double unitTime, double unitDensity, const double lbF[3]) {
    if (!d || !boundary || !lbSize || !lbF) return fail(HYB_E_ARG, "hyb_dem_init_lattice: null argument");
    if (!(unitLength > 0) || !(unitTime > 0) || !(unitDensity > 0)) return fail(HYB_E_ARG, "units must be > 0");
    LatticeInfo li;
    for (int k = 0; k < 6; ++k) li.boundary[k] = boundary[k];
    li.X = lbSize[0]; li.Y = lbSize[1]; li.Z = lbSize[2];
    if (li.X < 3 || li.Y < 3 || li.Z < 3) return fail(HYB_E_ARG, "lattice must have at least 3 nodes per axis");
    li.unit = make_units(unitLength, unitTime, unitDensity);
    for (int k = 0; k < 3; ++k) li.lbF[k] = lbF[k];
    return dem_init(d, li);
}

static int dem_init(hyb_dem* d, const LatticeInfo& lbi) {
    const LatticeInfo* lb = &lbi;
    const hyb_config& c = d->cfg;
    d->unit = lb->unit;
    const Units& U = d->unit;
    // DEM.cpp:185-233
    for (int k = 0; k < 3; ++k) d->demSize[k] = double(k == 0 ? lb->X : (k == 1 ? lb->Y : lb->Z)) * U.Length;
    const V3 demF = V3(lb->lbF[0], lb->lbF[1], lb->lbF[2]) * U.Accel;
    if (d->P) {
        if (d->multiStep == 0) return fail(HYB_E_UNSUPPORTED, "multiStep = 0 (automatic) is not supported");
        d->deltat = U.Time / (double(d->multiStep));
    } else {
        d->deltat = U.Time;
        d->multiStep = 1;
    }
    DEMParams& p = d->p;
    std::memset(&p, 0, sizeof(p));
    p.P = d->P;
    p.E = d->E;
    p.clumps = d->P != d->E;
    // walls (DEM.cpp:435-638, the six box planes)
    d->walls.clear();
    const V3 ps[6] = {V3(0.5 * U.Length, 0, 0), V3((double(lb->X) - 1.5) * U.Length, 0, 0), V3(0, 0.5 * U.Length, 0),
                      V3(0, (double(lb->Y) - 1.5) * U.Length, 0), V3(0, 0, 0.5 * U.Length),
                      V3(0, 0, (double(lb->Z) - 1.5) * U.Length)};
    const V3 ns[6] = {V3(1, 0, 0), V3(-1, 0, 0), V3(0, 1, 0), V3(0, -1, 0), V3(0, 0, 1), V3(0, 0, -1)};
    for (int k = 0; k < 6; ++k) {
        const int bc = lb->boundary[k];
        if (bc == 5 || bc == 6 || bc == 7 || bc == 8) {
            WallDev w;
            w.n = ns[k];
            w.p = ps[k];
            w.rotCenter = V3();
            w.omega = V3();
            w.vel = V3();
            w.moving = (bc == 6 || bc == 8);
            w.slip = (bc == 5 || bc == 6);
            d->walls.push_back(w);
        }
    }
    // DRUM's two extra y walls (DEM.cpp:783-868), after the box walls; only their
    // static forms (boundary 5, 7): the rotating ones (6, 8) are not supported
    if (c.problem == HYB_PROBLEM_DRUM) {
        for (int k = 2; k < 4; ++k) {
            const int bc = lb->boundary[k];
            if (bc == 6 || bc == 8)
                return fail(HYB_E_UNSUPPORTED, "DRUM: rotating y walls (boundary 6, 8) are not supported");
            if (bc != 5 && bc != 7) continue;
            WallDev w;
            w.p = k == 2 ? V3(0.0 + 0.6 + 0.35, 0.50001 * U.Length, 0.0)
                         : V3(0.0 + 0.6 + 0.35, (double(lb->Y) - 1.50001) * U.Length, 0.0);
            w.n = k == 2 ? V3(0.0, 1.0, 0.0) : V3(0.0, -1.0, 0.0);
            w.rotCenter = V3();
            w.omega = V3();
            w.vel = V3();
            w.moving = 0;
            w.slip = bc == 5;
            d->walls.push_back(w);
        }
    }
    // cylinders (DEM.cpp:870-935): the NET problem's static pipe along x and DRUM's
    // rotating drum along y; AVALANCHE (a free-surface obstacle) is not supported
    d->cylinders.clear();
    if (c.problem == HYB_PROBLEM_DRUM) {
        CylDev cy;
        cy.p1 = V3(0.0 + 0.6 + 0.35, 0.0, 1.26);
        cy.p2 = V3(0.0 + 0.6 + 0.35, 1.0, 1.26);
        cy.R = 1.243;
        cy.omega = V3(0.0, c.drumSpeed, 0.0);
        cy.moving = 1;
        cy.slip = 0;
        const V3 axes = cy.p2 - cy.p1;  // initAxes (utils.cpp:154-161)
        cy.naxes = axes / axes.norm();
        d->cylinders.push_back(cy);
    }
    if (c.problem == HYB_PROBLEM_NET) {
        CylDev cy;
        cy.p1 = V3(0.0, 4.0, 4.0);
        cy.p2 = V3(1.0, 4.0, 4.0);
        cy.R = 4.0;
        cy.omega = V3();
        cy.moving = 0;
        cy.slip = 0;
        const V3 axes = cy.p2 - cy.p1;  // initAxes (utils.cpp:154-161)
        cy.naxes = axes / axes.norm();
        d->cylinders.push_back(cy);
    }
    p.ncyl = (int)d->cylinders.size();
    for (int k = 0; k < p.ncyl; ++k) p.cyl[k] = d->cylinders[k];
    // periodic pairs (DEM.cpp:937-988)
    d->pbcs.clear();
    for (int k = 0; k < 3; ++k) {
        if (lb->boundary[2 * k] == 4) {
            if (lb->boundary[2 * k + 1] != 4) return fail(HYB_E_ARG, "Periodic boundaries error!");
            PbcDev pb;
            V3 base, v;
            const double sz = double(k == 0 ? lb->X : (k == 1 ? lb->Y : lb->Z));
            if (k == 0) { base = V3(0.5 * U.Length, 0, 0); v = V3((sz - 2.0) * U.Length, 0, 0); }
            if (k == 1) { base = V3(0, 0.5 * U.Length, 0); v = V3(0, (sz - 2.0) * U.Length, 0); }
            if (k == 2) { base = V3(0, 0, 0.5 * U.Length); v = V3(0, 0, (sz - 2.0) * U.Length); }
            pb.v = v;
            pb.p1 = base;
            pb.p2 = base + v;
            pb.n1 = v / v.norm();
            pb.n2 = -1.0 * pb.n1;
            d->pbcs.push_back(pb);
        }
    }
    p.npbc = (int)d->pbcs.size();
    p.nwalls = (int)d->walls.size();
    for (int k = 0; k < p.npbc; ++k) p.pbc[k] = d->pbcs[k];
    for (int k = 0; k < p.nwalls; ++k) p.walls[k] = d->walls[k];
    // neighbour-cell sizing (DEM.cpp:1270-1312)
    double maxRad = 0.0;
    for (int e = 0; e < d->E; ++e) maxRad = std::max(maxRad, d->rec[(size_t)e * HYB_ELEMENT_RECORD + 2] * c.scale);
    for (int k = 0; k < 3; ++k) {
        p.cellWidth[k] = d->P ? std::min(maxRad * 5.0, d->demSize[k]) : d->demSize[k];
        p.nCells[k] = (unsigned)int(std::ceil(d->demSize[k] / p.cellWidth[k]));
        p.cellWidth[k] = d->demSize[k] / double(p.nCells[k]);
        p.nCells[k] += 2;
    }
    p.nebrRange = d->P ? std::max(maxRad * 3.0, 0.5 * std::min(p.cellWidth[0], std::min(p.cellWidth[1], p.cellWidth[2])))
                       : 0.0;
    p.trigger = 0.25 * p.nebrRange;
    p.deltat = d->deltat;
    const double dt = d->deltat;
    const double cc[5] = {dt, dt * dt / 2.0, dt * dt * dt / 6.0, dt * dt * dt * dt / 24.0, dt * dt * dt * dt * dt / 120.0};
    for (int k = 0; k < 5; ++k) { p.c1[k] = cc[k]; p.c2[k] = cc[k]; }
    const double g1[6] = {95.0 / 288.0, 1.0, 25.0 / 24.0, 35.0 / 72.0, 5.0 / 48.0, 1.0 / 120.0};
    const double g2[6] = {3.0 / 16.0, 251.0 / 360.0, 1.0, 11.0 / 18.0, 1.0 / 6.0, 1.0 / 60.0};
    const double co1[6] = {g1[0] * cc[0], g1[1] * cc[0] / cc[0], g1[2] * cc[0] / cc[1], g1[3] * cc[0] / cc[2],
                           g1[4] * cc[0] / cc[3], g1[5] * cc[0] / cc[4]};
    const double co2[6] = {g2[0] * cc[1], g2[1] * cc[1] / cc[0], g2[2] * cc[1] / cc[1], g2[3] * cc[1] / cc[2],
                           g2[4] * cc[1] / cc[3], g2[5] * cc[1] / cc[4]};
    for (int k = 0; k < 6; ++k) { p.co1[k] = co1[k]; p.co2[k] = co2[k]; }
    p.knConst = d->knConst; p.ksConst = d->ksConst; p.dampCoeff = d->dampCoeff;
    p.viscTang = c.viscTang; p.fricPart = c.frictionCoefPart; p.fricWall = c.frictionCoefWall;
    p.linearStiff = c.linearStiff;
    p.hertzian = c.contactModel == HYB_CONTACT_HERTZIAN;
    if (c.tangentialModel != HYB_TANGENTIAL_VISCOUS && c.tangentialModel != HYB_TANGENTIAL_SPRING)
        return fail(HYB_E_ARG, "unknown tangential contact model");
    p.spring = c.tangentialModel == HYB_TANGENTIAL_SPRING;
    if (c.wallContacts != HYB_WALL_CONTACTS_FIRST && c.wallContacts != HYB_WALL_CONTACTS_ALL)
        return fail(HYB_E_ARG, "unknown wall-contact mode");
    p.allWalls = c.wallContacts == HYB_WALL_CONTACTS_ALL;
    p.numVisc = c.numVisc;
    p.demF[0] = demF.x; p.demF[1] = demF.y; p.demF[2] = demF.z;
    // objects: periodic copies of those within 2 nebrRange of a periodic plane,
    // made only when elements exist (DEM.cpp:252-256, 990-1062)
    d->objects.resize((size_t)d->stdObjects);
    if (d->P) {
        struct G { int origin; V3 v; };
        std::vector<G> side, corner;
        for (const PbcDev& pb : d->pbcs)
            for (int o = 0; o < d->stdObjects; ++o) {
                const double leftDist = pb.n1.dot(d->objects[o].x0 - pb.p1);
                const double rightDist = pb.n2.dot(d->objects[o].x0 - pb.p2);
                if (leftDist < 2.0 * p.nebrRange && leftDist > 0.0) side.push_back(G{d->objects[o].index, pb.v});
                else if (rightDist < 2.0 * p.nebrRange && rightDist > 0.0)
                    side.push_back(G{d->objects[o].index, -1.0 * pb.v});
            }
        for (size_t a = 0; a < side.size(); ++a)
            for (size_t k = a + 1; k < side.size(); ++k)
                if (side[a].origin == side[k].origin) corner.push_back(G{side[a].origin, side[a].v + side[k].v});
        side.insert(side.end(), corner.begin(), corner.end());
        for (size_t g = 0; g < side.size(); ++g) {
            ObjDev ob = d->objects[side[g].origin];
            ob.x0 = ob.x0 + side[g].v;
            ob.index = d->stdObjects + (int)g;
            d->objects.push_back(ob);
        }
    }
    p.nobj = d->P ? (int)d->objects.size() : 0;
    // capacities
    const int np = p.npbc;
    const int ghostsPer = np + np * (np - 1) / 2;
    p.Gcap = std::max(1, d->P * std::max(ghostsPer, 1));
    const int scap = d->P + p.Gcap;
    p.pairCap = c.pairCapacity > 0 ? c.pairCapacity : std::max(1024, scap * 16 + 1024);
    DEMBuffers& b = d->b;
    std::memset(&b, 0, sizeof(b));
    b.scap = scap;
    b.ncells = (int)(p.nCells[0] * p.nCells[1] * p.nCells[2]);
    b.W = b.ncells * 14;
    // element state, host side (elmt.cpp:17-83 with wSolver = true), and the
    // particles of each element in prototype order (elmt.cpp:122-137)
    const int P = d->P;
    const int NE = d->E;
    // compositeProperties (DEM.cpp:404-433); -1/2 is integer division there, so
    // the second coordinate of the three-sphere prototype's last two vectors is 0
    const double s3 = std::sqrt(3), s2 = std::sqrt(2), s6 = std::sqrt(6);
    const V3 proto[5][4] = {{},
                            {V3()},
                            {V3(0.5, 0.0, 0.0), V3(-0.5, 0.0, 0.0)},
                            {V3(0.0, 1.0, 0.0), V3(-s3 / 2, -1 / 2, 0.0), V3(s3 / 2, -1 / 2, 0.0)},
                            {V3(0.0, 0.0, 1.0), V3(0.0, 2.0 * s2 / 3.0, -1.0 / 3.0),
                             V3(2.0 * s6 / 6.0, -2.0 * s2 / 6.0, -1.0 / 3.0),
                             V3(-2.0 * s6 / 6.0, -2.0 * s2 / 6.0, -1.0 / 3.0)}};
    std::vector<double> el((size_t)EF_COUNT * std::max(NE, 1), 0.0);
    std::vector<int> pelem, efirst;
    std::vector<double> pproto;
    auto at = [&](int f, int e) -> double& { return el[(size_t)f * NE + e]; };
    auto set3 = [&](int f, int e, const V3& v) { at(f, e) = v.x; at(f + 1, e) = v.y; at(f + 2, e) = v.z; };
    auto set4 = [&](int f, int e, const Q4& q) { at(f, e) = q.q0; at(f + 1, e) = q.q1; at(f + 2, e) = q.q2; at(f + 3, e) = q.q3; };
    const V3 translate(c.translate[0], c.translate[1], c.translate[2]);
    for (int e = 0; e < NE; ++e) {
        const double* r = &d->rec[(size_t)e * HYB_ELEMENT_RECORD];
        const int size = (int)r[1];
        const double radius = r[2] * c.scale;
        const V3 x0 = V3(r[3] * c.scale, r[4] * c.scale, r[5] * c.scale) + translate;
        const V3 x1(r[6], r[7], r[8]);
        const V3 w0(r[9], r[10], r[11]);
        const Q4 q0(r[12], r[13], r[14], r[15]);
        const Q4 q1(r[16], r[17], r[18], r[19]);
        set3(EF_X, e, x0); set3(EF_X + 3, e, x1);
        set3(EF_XP, e, x0); set3(EF_XP + 3, e, x1);
        set4(EF_Q, e, q0); set4(EF_Q + 4, e, q1);
        set4(EF_QP, e, q0); set4(EF_QP + 4, e, q1);
        set3(EF_W, e, w0);
        set3(EF_WG, e, w0); set3(EF_WL, e, w0); set3(EF_WPG, e, w0); set3(EF_WPL, e, w0);
        const double singleMass = 4.0 / 3.0 * c.density * M_PI * radius * radius * radius;
        const double m = size * singleMass;
        V3 I = size * 2.0 / 5.0 * singleMass * radius * radius * V3(1.0, 1.0, 1.0);
        for (int n = 0; n < size; ++n) {  // Steiner terms (vector.cpp transport: y^2+z^2, z^2+x^2, x^2+y^2)
            const V3& pr = proto[size][n];
            I += singleMass * radius * radius * V3(pr.y * pr.y + pr.z * pr.z, pr.z * pr.z + pr.x * pr.x,
                                                   pr.x * pr.x + pr.y * pr.y);
        }
        at(EF_M, e) = m;
        at(EF_RAD, e) = radius;
        set3(EF_I, e, I);
        set3(EF_FG, e, demF * m);
        efirst.push_back((int)pelem.size());
        for (int n = 0; n < size; ++n) {
            pelem.push_back(e);
            pproto.push_back(proto[size][n].x); pproto.push_back(proto[size][n].y); pproto.push_back(proto[size][n].z);
        }
    }
    efirst.push_back((int)pelem.size());
    // device buffers
    int rc;
    if ((rc = errword_alloc(d->err))) return rc;
    errword_clear(d->err);
    p.herr = d->err.dev;
    std::vector<Blk>& R = d->saved;
    R.clear();
    if ((rc = ralloc(R, "el", &b.el, el.size())) || (rc = ralloc(R, "st", &b.st, 1)) ||
        (rc = ralloc(R, "ghostIndex", &b.ghostIndex, p.Gcap)) || (rc = ralloc(R, "ghostVec", &b.ghostVec, 3 * (size_t)p.Gcap)) ||
        (rc = ralloc(R, "pairs", &b.pairs, 2 * (size_t)p.pairCap)) || (rc = ralloc(R, "nearWall", &b.nearWall, std::max(P, 1))) ||
        (rc = ralloc(R, "cellTable", &b.cellTable, (size_t)scap + b.ncells)) || (rc = ralloc(R, "cellOf", &b.cellOf, scap)) ||
        (rc = ralloc(R, "cellCount", &b.cellCount, b.ncells + 1)) ||
        (rc = dalloc(&b.work, (size_t)std::max(std::max(b.W, scap), std::max(p.pairCap, b.ncells + 1)) + 1)) ||
        (rc = ralloc(R, "csrOff", &b.csrOff, (size_t)P + 1)) || (rc = ralloc(R, "csr", &b.csr, 2 * (size_t)p.pairCap)) ||
        (rc = dalloc(&b.cursor, scap)) || (rc = dalloc(&b.spos, 4 * (size_t)scap)) || (rc = ralloc(R, "ghostOfCount", &b.ghostOfCount, std::max(P, 1))) ||
        (rc = ralloc(R, "ghostOf", &b.ghostOf, 8 * (size_t)std::max(P, 1))) ||
        (rc = ralloc(R, "side", &b.side, 3 * (size_t)std::max(P, 1))) || (rc = ralloc(R, "ghostB", &b.ghostB, p.Gcap)) ||
        (rc = ralloc(R, "posI", &b.posI, p.pairCap)) || (rc = ralloc(R, "posJ", &b.posJ, p.pairCap)) ||
        (rc = dalloc(&b.prec, 2 * (size_t)p.pairCap)) || (rc = ralloc(R, "vmax", &b.vmax, 1)) || (rc = dalloc(&b.energy, 6)) ||
        (rc = dalloc(&b.wcf, (size_t)std::max(P, 1) * wall_slots(p))))
        return rc;
    HIPCHK(hipMemset(b.wcf, 0, sizeof(ContactForce) * (size_t)std::max(P, 1) * wall_slots(p)));
    if (p.allWalls) {
        if ((rc = ralloc(R, "nearWallMask", &b.nearWallMask, (size_t)std::max(P, 1)))) return rc;
        HIPCHK(hipMemset(b.nearWallMask, 0, sizeof(uint32_t) * (size_t)std::max(P, 1)));
    }
    if (p.ncyl && (rc = ralloc(R, "nearCyl", &b.nearCyl, (size_t)std::max(P, 1)))) return rc;
    if (p.clumps) {
        if ((rc = dalloc(&b.pelem, (size_t)P)) || (rc = dalloc(&b.efirst, (size_t)NE + 1)) ||
            (rc = dalloc(&b.pproto, 3 * (size_t)P)) || (rc = ralloc(R, "pcor", &b.pcor, 9 * (size_t)P)) ||
            (rc = ralloc(R, "ppred", &b.ppred, 9 * (size_t)P)) || (rc = dalloc(&b.pmn, 6 * (size_t)p.pairCap)))
            return rc;
        HIPCHK(hipMemcpy(b.pelem, pelem.data(), sizeof(int) * pelem.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.efirst, efirst.data(), sizeof(int) * efirst.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.pproto, pproto.data(), sizeof(double) * pproto.size(), hipMemcpyHostToDevice));
    }
    if (p.spring) {
        const size_t Pm = (size_t)std::max(P, 1), Wm = Pm * (size_t)wall_slots(p);  // a spring per wall slot
        if ((rc = ralloc(R, "hist", &b.hist, 3 * (size_t)p.pairCap)) || (rc = ralloc(R, "whist", &b.whist, 4 * Wm)) ||
            (rc = dalloc(&b.histOld, 3 * (size_t)p.pairCap)) || (rc = dalloc(&b.pairsOld, 2 * (size_t)p.pairCap)) ||
            (rc = dalloc(&b.csrOffOld, Pm + 1)) || (rc = dalloc(&b.csrOld, 2 * (size_t)p.pairCap)) ||
            (rc = dalloc(&b.ghostIndexOld, (size_t)p.Gcap)))
            return rc;
        HIPCHK(hipMemset(b.hist, 0, 3 * sizeof(double) * (size_t)p.pairCap));
        std::vector<double> wh(4 * Wm, 0.0);
        for (size_t e = 0; e < Wm; ++e) wh[4 * e + 3] = -1.0;  // no wall in contact
        HIPCHK(hipMemcpy(b.whist, wh.data(), wh.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (p.nobj) {
        ObjDev* objs = nullptr;
        if ((rc = ralloc(R, "objs", &objs, (size_t)p.nobj)) ||
            (rc = ralloc(R, "nearObj", &b.nearObj, (size_t)P * NEAR_OBJ)) ||
            (rc = ralloc(R, "nearObjCount", &b.nearObjCount, (size_t)P)) ||
            (rc = dalloc(&b.ocf, (size_t)std::max(P, 1) * NEAR_OBJ)))
            return rc;
        HIPCHK(hipMemset(b.ocf, 0, sizeof(ContactForce) * (size_t)std::max(P, 1) * NEAR_OBJ));
        HIPCHK(hipMemcpy(objs, d->objects.data(), sizeof(ObjDev) * (size_t)p.nobj, hipMemcpyHostToDevice));
        b.objs = objs;
    }
    HIPCHK(hipMemcpyAsync(b.el, el.data(), el.size() * sizeof(double), hipMemcpyHostToDevice, d->stream));
    DEMState st;
    std::memset(&st, 0, sizeof(st));
    st.maxDisp = 0.5 * p.nebrRange;  // DEM.cpp:1307
    st.nSlots = P;
    HIPCHK(hipMemcpyAsync(b.st, &st, sizeof(st), hipMemcpyHostToDevice, d->stream));
    HIPCHK(hipMemsetAsync(b.nearWall, 0xFF, sizeof(uint32_t) * std::max(P, 1), d->stream));
    HIPCHK(hipMemsetAsync(b.vmax, 0, sizeof(double), d->stream));
    if (b.nearCyl) HIPCHK(hipMemsetAsync(b.nearCyl, 0xFF, sizeof(uint32_t) * std::max(P, 1), d->stream));
    HIPCHK(hipMemsetAsync(b.csrOff, 0, sizeof(uint32_t) * ((size_t)P + 1), d->stream));
    HIPCHK(hipMemsetAsync(b.ghostOfCount, 0, sizeof(int) * std::max(P, 1), d->stream));
    if (P) {
        launch_dem_particles(p, b, d->stream);   // generateParticles (elmt.cpp:122-137)
        launch_dem_rebuild(p, b, 1, d->stream);  // evalNeighborTable at init (DEM.cpp:258-262)
        launch_dem_predict(p, b, d->stream);     // the first sub-step's predictor
    }
    HIPCHK(hipGetLastError());
    d->demTime = 0.0;
    d->demTimeStep = 0;
    d->inited = true;
    return HYB_OK;
}

static void save_free(hyb_dem* d) {
    dfree(d->saveElems); dfree(d->saveInfo); dfree(d->savePairs); dfree(d->saveWalls); dfree(d->saveObjs);
    d->saveSlots = 0;
}

int hyb_dem_set_save_count(hyb_dem* d, int32_t saveCount) {
    if (!d || !d->inited) return fail(HYB_E_STATE, "hyb_dem_set_save_count before hyb_dem_init");
    if (saveCount < 0) return fail(HYB_E_ARG, "saveCount must be >= 0");
    if (saveCount == d->saveCount) return HYB_OK;
    HIPCHK(hipStreamSynchronize(d->stream));
    save_free(d);
    d->saveEvents.clear();
    d->saveCount = saveCount;
    if (!saveCount) return HYB_OK;
    // at most ceil(multiStep / saveCount) saving sub-steps in one DEM step
    const int slots = (d->multiStep + saveCount - 1) / saveCount;
    const size_t P = (size_t)std::max(d->P, 1);
    int rc;
    if ((rc = dalloc(&d->saveElems, (size_t)slots * P * 9)) || (rc = dalloc(&d->saveInfo, (size_t)slots)) ||
        (rc = dalloc(&d->savePairs, (size_t)slots * d->p.pairCap)) || (rc = dalloc(&d->saveWalls, (size_t)slots * P * wall_slots(d->p))) ||
        (d->p.nobj && (rc = dalloc(&d->saveObjs, (size_t)slots * P * NEAR_OBJ))))
        return rc;
    d->saveSlots = slots;
    return HYB_OK;
}

int hyb_dem_step(hyb_dem* d) {  // DEM.cpp:331-376
    if (!d || !d->inited) return fail(HYB_E_STATE, "hyb_dem_step before hyb_dem_init");
    if (int rc = dem_pending(d)) return rc;
    d->saveEvents.clear();
    for (int it = 0; it < d->multiStep; ++it) {
        d->demTimeStep++;
        d->demTime += d->deltat;
        const bool save = d->saveCount > 0 && d->demTimeStep % (uint32_t)d->saveCount == 0;
        const int slot = (int)d->saveEvents.size();
        if (save) {
            if (slot >= d->saveSlots) return fail(HYB_E_STATE, "hyb_dem_step: more saving sub-steps than slots");
            d->saveEvents.emplace_back(d->demTime, d->demTimeStep);
        }
        if (d->P) {
            launch_dem_rebuild(d->p, d->b, 0, d->stream);
            if (save) {
                const size_t P = (size_t)d->P;
                DEMBuffers bs = d->b;
                bs.pstat = d->savePairs + (size_t)slot * d->p.pairCap;
                bs.wstat = d->saveWalls + (size_t)slot * P * wall_slots(d->p);
                bs.ostat = d->saveObjs ? d->saveObjs + (size_t)slot * P * NEAR_OBJ : nullptr;
                launch_dem_save_elements(d->p, d->b, d->saveElems + (size_t)slot * P * 9, d->saveInfo + slot, d->stream);
                launch_dem_pairs(d->p, bs, d->stream);
                launch_dem_gather(d->p, bs, d->stream);
            } else {
                launch_dem_pairs(d->p, d->b, d->stream);
                launch_dem_gather(d->p, d->b, d->stream);
            }
        }
        d->substeps++;
    }
    HIPCHK(hipGetLastError());
    return HYB_OK;
}

int hyb_dem_destroy(hyb_dem* d) {
    if (!d) return HYB_OK;
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    DEMBuffers& b = d->b;
    dfree(b.el); dfree(b.st); dfree(b.ghostIndex); dfree(b.ghostVec); dfree(b.pairs); dfree(b.nearWall); dfree(b.nearWallMask);
    dfree(b.cellTable); dfree(b.cellOf); dfree(b.cellCount); dfree(b.work); dfree(b.csrOff); dfree(b.csr);
    dfree(b.cursor); dfree(b.spos); dfree(b.ghostOfCount); dfree(b.ghostOf); dfree(b.side); dfree(b.ghostB);
    dfree(b.posI); dfree(b.posJ); dfree(b.prec); dfree(b.vmax); dfree(b.energy); dfree(b.nearObj); dfree(b.nearObjCount);
    dfree(b.wcf); dfree(b.ocf); dfree(b.nearCyl);
    dfree(b.hist); dfree(b.whist); dfree(b.histOld); dfree(b.pairsOld); dfree(b.csrOffOld); dfree(b.csrOld);
    dfree(b.ghostIndexOld);
    dfree(b.pelem); dfree(b.efirst); dfree(b.pproto); dfree(b.pcor); dfree(b.ppred); dfree(b.pmn);
    if (b.objs) (void)hipFree((void*)b.objs);
    save_free(d);
    errword_free(d->err);
    delete d;
    return HYB_OK;
}

// ------------------------------------------------------------------ LB
int hyb_lb_create(const hyb_config* cfg, hyb_lb** out) {  // LB.cpp:85-188 + LB.cpp:42-83
    if (!cfg || !out) return fail(HYB_E_ARG, "hyb_lb_create: null argument");
    hyb_lb* l = new hyb_lb();
    l->cfg = *cfg;
    const hyb_config& c = *cfg;
    if (!(c.unitLength > 0) || !(c.unitDensity > 0)) { delete l; return fail(HYB_E_ARG, "unit length/density must be > 0"); }
    l->nonNewtonian = c.nonNewtonian != 0;
    if (c.rheology != HYB_RHEOLOGY_BINGHAM && c.rheology != HYB_RHEOLOGY_FRICTIONAL) {
        delete l;
        return fail(HYB_E_ARG, "unknown rheology");
    }
    if (!(c.frictionCoefFluid >= 0.0)) { delete l; return fail(HYB_E_ARG, "frictionCoefFluid must be >= 0"); }
    l->frictional = c.rheology == HYB_RHEOLOGY_FRICTIONAL;
    l->frictionFluid = c.frictionCoefFluid;
    l->turbulenceOn = c.turbulence != 0;
    l->freeSurface = c.freeSurface != 0;
    l->forceField = c.forceField != 0;
    l->X = int(std::floor(c.lbSize[0] / c.unitLength + 0.5));
    l->Y = int(std::floor(c.lbSize[1] / c.unitLength + 0.5));
    l->Z = int(std::floor(c.lbSize[2] / c.unitLength + 0.5));
    if (l->X < 3 || l->Y < 3 || l->Z < 3) { delete l; return fail(HYB_E_ARG, "lattice must have at least 3 nodes per axis"); }
    l->N = (size_t)l->X * l->Y * l->Z;
    double unitTime = c.unitTime;
    if (unitTime == 0.0) {
        const double tauOpt = 1.0;
        unitTime = (l->nonNewtonian || l->frictional) ? c.unitDensity * c.unitLength * c.unitLength / c.plasticVisc * (1.0 / 3.0) * (tauOpt - 0.5)
                                   : c.unitDensity * c.unitLength * c.unitLength / c.initVisc * (1.0 / 3.0) * (tauOpt - 0.5);
    }
    l->unit = make_units(c.unitLength, unitTime, c.unitDensity);
    l->initDynVisc = c.initVisc / l->unit.DynVisc;
    l->plasticVisc = c.plasticVisc / l->unit.DynVisc;
    l->yieldStress = c.yieldStress / l->unit.Stress;
    l->initDensity = 1.0;
    for (int k = 0; k < 3; ++k) {
        l->initVelocity[k] = c.initVelocity[k] / l->unit.Speed;
        l->lbF[k] = c.lbF[k] / l->unit.Accel;
    }
    for (int k = 0; k < 6; ++k) {
        l->boundary[k] = c.boundary[k];
        if (c.boundary[k] < 4 || c.boundary[k] > 8) {
            delete l;
            return fail(HYB_E_UNSUPPORTED, "box boundaries must be periodic (4), slip (5, 6) or no-slip (7, 8) walls");
        }
    }
    for (int k = 0; k < 3; ++k)
        if ((c.boundary[2 * k] == 4) != (c.boundary[2 * k + 1] == 4)) { delete l; return fail(HYB_E_ARG, "Periodic boundaries error!"); }
    l->slipCoefficient = c.slipCoefficient;
    l->turbConst = c.turbConst;
    if (c.problem != HYB_PROBLEM_NONE && c.problem != HYB_PROBLEM_SPLASH && c.problem != HYB_PROBLEM_NET &&
        c.problem != HYB_PROBLEM_DRUM && c.problem != HYB_PROBLEM_LIQUID_BOX) {
        delete l;
        return fail(HYB_E_UNSUPPORTED, "problemName not supported");
    }
    if (c.problem != HYB_PROBLEM_NONE && !l->freeSurface) {
        delete l;
        return fail(HYB_E_UNSUPPORTED, "gas regions without free-surface tracking (freeSurfaceSolve = 0) are not supported");
    }
    if (l->freeSurface && l->N >= 0xFFFFFFFFull) {
        delete l;
        return fail(HYB_E_CAPACITY, "free-surface lattices are limited to 2^32 - 1 nodes");
    }
    if (c.collision != HYB_COLLISION_BGK && c.collision != HYB_COLLISION_TRT) { delete l; return fail(HYB_E_ARG, "unknown collision model"); }
    l->prec = c.precision == HYB_FP32_SHIFTED ? Prec::F32S : Prec::F64;
#ifdef HYB_PLANE_PAD
    l->planePad = HYB_PLANE_PAD;
#endif
    l->stream = shared_stream();
    *out = l;
    return HYB_OK;
}

static int lb_alloc(hyb_lb* l, int slotCap) {
    const size_t N = l->N;
    int rc;
    if ((rc = errword_alloc(l->err))) return rc;
    errword_clear(l->err);
    const size_t pb = l->popBytes();
    std::vector<Blk>& R = l->saved;
    R.clear();
    HIPCHK(hipMalloc(&l->f[0], pb));
    HIPCHK(hipMalloc(&l->f[1], pb));
    R.push_back(Blk{"f0", l->f[0], pb});
    R.push_back(Blk{"f1", l->f[1], pb});
    if ((rc = ralloc(R, "type", &l->type, N)) || (rc = ralloc(R, "link", &l->link, N)) ||
        (rc = ralloc(R, "solid", &l->solid, N)) ||
        (rc = dalloc(&l->maxXYZ, 3 * (size_t)l->Z)) || (rc = dalloc(&l->dcount, (size_t)l->Z + 2)) ||
        (rc = dalloc(&l->nslots, 2)))
        return rc;
    if (l->varVisc() && (rc = ralloc(R, "visc", &l->visc, N))) return rc;
    if (l->hasSlip() && (rc = ralloc(R, "slipw", &l->slipw, N))) return rc;
    if (l->freeSurface) {
        FSBuffers& b = l->fsb;
        const size_t eb = l->prec == Prec::F64 ? 8 : 4;
        HIPCHK(hipMalloc(&b.rho[0], eb * N));
        HIPCHK(hipMalloc(&b.rho[1], eb * N));
        R.push_back(Blk{"rho0", b.rho[0], eb * N});
        R.push_back(Blk{"rho1", b.rho[1], eb * N});
        if ((rc = ralloc(R, "mass", &b.mass, N)) || (rc = ralloc(R, "uS", &b.uS, 3 * N)) ||
            (rc = ralloc(R, "hFi", &b.hFi, 3 * N)) || (rc = ralloc(R, "claim", &b.claim, N)) ||
            (rc = ralloc(R, "list0", &b.list[0], N)) || (rc = ralloc(R, "list1", &b.list[1], N)) ||
            (rc = ralloc(R, "listA", &b.listA, N)) || (rc = ralloc(R, "listB", &b.listB, N)) ||
            (rc = ralloc(R, "mnew", &b.mnew, N)) || (rc = ralloc(R, "state", &b.state, N)) ||
            (rc = ralloc(R, "ctr", &b.ctr, 1)))
            return rc;
        if (l->hasSlip() && (rc = ralloc(R, "actPrev", &b.actPrev, N))) return rc;
    }
    l->slotCap = slotCap;
    if (slotCap > 0) {
        if ((rc = dalloc(&l->cp, slotCap)) || (rc = dalloc(&l->cpf, slotCap)) || (rc = dalloc(&l->boxes, slotCap)) ||
            (rc = dalloc(&l->segs, (size_t)slotCap * l->segCap)) || (rc = dalloc(&l->segCount, slotCap)))
            return rc;
        HIPCHK(hipMemset(l->segCount, 0, sizeof(uint32_t) * (size_t)slotCap));
    }
    return HYB_OK;
}

// Particle-node classification of one coupling step: every slot marks the
// active nodes strictly inside it with this step's tag (LB.cpp:475-495).
static void lb_couple_launch(hyb_lb* l, hyb_dem* d) {
    if (++l->ctag > SOLID_TAG_MAX) {  // tag wrap-around: clear the ownership words once
        launch_fill_u32(l->solid, 0u, l->N, l->stream);
        l->ctag = 1;
    }
    const LBKParams p = l->params();
    launch_dem_couple_table(d->p, d->b, l->cp, l->cpf, l->nslots, l->unit, l->stream);
    // without free surface no inactive node becomes active and the ownership of an inactive node is never
    // read (views mask it), so every node inside a sphere is marked without reading its link word
    launch_classify_set(l->nslots, l->slotCap, l->cp, l->boxes, l->nslots + 1, l->solid,
                        l->freeSurface ? l->link : nullptr, p, l->stream);
}

int hyb_lb_init(hyb_lb* l, hyb_dem* d) {  // LB.cpp:190-219
    if (!l) return fail(HYB_E_ARG, "hyb_lb_init: null handle");
    if (d && !d->inited) return fail(HYB_E_STATE, "hyb_lb_init: DEM handle not initialised");
    const int slotCap = d ? d->P + d->p.Gcap : 0;
    if (slotCap >= (int)SOLID_SLOT_MASK) return fail(HYB_E_CAPACITY, "too many particle slots for the ownership word");
    const bool cylinders = d && !d->cylinders.empty();
    if (l->cfg.problem == HYB_PROBLEM_DRUM && !cylinders)
        return fail(HYB_E_ARG, "hyb_lb_init: DRUM needs the DEM handle that holds the drum");
    // run-sum list capacity per slot: two runs per lattice row of the largest
    // sphere's bounding box (a run splits at most once at a wavefront boundary)
    // plus margin for rows cut by overlapping slots
    if (d && d->P) {
        double rmax = 0.0;
        for (int e = 0; e < d->E; ++e) rmax = std::max(rmax, d->rec[(size_t)e * HYB_ELEMENT_RECORD + 2]);
        rmax *= d->cfg.scale / l->unit.Length;
        const double side = 2.0 * std::ceil(rmax) + 3.0;
        if (!(side < 4096.0)) return fail(HYB_E_ARG, "particle radius too large for the lattice");
        l->segCap = (uint32_t)(2.0 * side * side + 64.0);
    }
    int rc = lb_alloc(l, d && d->P ? slotCap : 0);
    if (rc) return rc;
    // the curved-wall rule tests the type of the node behind as it was when streamed
    if (cylinders && l->freeSurface && !l->fsb.actPrev && (rc = ralloc(l->saved, "actPrev", &l->fsb.actPrev, l->N)))
        return rc;
    CylSet cs;
    cs.n = 0;
    if (cylinders) {
        cs.n = (int)std::min<size_t>(d->cylinders.size(), 2);
        for (int k = 0; k < cs.n; ++k) cs.c[k] = d->cylinders[k];
        l->cyl = d->cylinders[0];  // the curved-wall distances use the first cylinder (LB.cpp:597)
        if ((rc = ralloc(l->saved, "curvw", &l->curvw, l->N)) || (rc = ralloc(l->saved, "mac0", &l->macBuf[0], 4 * l->N)) ||
            (rc = ralloc(l->saved, "mac1", &l->macBuf[1], 4 * l->N)))
            return rc;
        HIPCHK(hipMemset(l->macBuf[0], 0, 32 * l->N));
        HIPCHK(hipMemset(l->macBuf[1], 0, 32 * l->N));
        l->macCur = 0;
    }
    const LBKParams p = l->params();
    hipStream_t s = l->stream;
    // node types: box boundaries, DEM plane walls, objects (LB.cpp:324-339), gas/interface
    std::vector<WallDev> walls = d ? d->walls : std::vector<WallDev>();
    WallDev* dw = nullptr;
    ObjDev* dob = nullptr;
    struct Release {  // the two upload buffers, on every return path
        WallDev*& w;
        ObjDev*& o;
        hipStream_t s;
        ~Release() {
            if (w || o) (void)hipStreamSynchronize(s);
            if (w) (void)hipFree(w);
            if (o) (void)hipFree(o);
        }
    } release{dw, dob, s};
    if (!walls.empty()) {
        HIPCHK(hipMalloc((void**)&dw, walls.size() * sizeof(WallDev)));
        HIPCHK(hipMemcpyAsync(dw, walls.data(), walls.size() * sizeof(WallDev), hipMemcpyHostToDevice, s));
    }
    std::vector<ObjDev> objs = d ? d->objects : std::vector<ObjDev>();
    if (!objs.empty()) {
        HIPCHK(hipMalloc((void**)&dob, objs.size() * sizeof(ObjDev)));
        HIPCHK(hipMemcpyAsync(dob, objs.data(), objs.size() * sizeof(ObjDev), hipMemcpyHostToDevice, s));
    }
    double region[7] = {l->cfg.liquidBox[0], l->cfg.liquidBox[1], l->cfg.liquidBox[2], l->cfg.liquidBox[3],
                        l->cfg.liquidBox[4], l->cfg.liquidBox[5], l->cfg.avalanchePosit};
    if (l->cfg.problem == HYB_PROBLEM_DRUM) {
        // the liquid cylinder of LB.cpp:734-741: axis point (x, z) and radius, lattice units
        const double fluidMass = l->cfg.fluidMass;
        const double fluidRadius = 1.01 + 0.00035298 * sqrt(7891604 - 5713 * (1389.9 - fluidMass));
        region[0] = 0.45 / l->unit.Length;
        region[1] = 0.0 / l->unit.Length;
        region[2] = (-1.0 + 0.0 + 0.04) / l->unit.Length;
        region[3] = fluidRadius / l->unit.Length;
    }
    launch_lb_init_types(l->type, l->boundary, (int)walls.size(), dw, (int)objs.size(), dob, cs, l->cfg.problem,
                         region, p, s);
    // the ownership words are cleared below, so their memory is the tagging scratch
    launch_lb_init_interface(l->type, (uint8_t*)l->solid, p, s);
    launch_lb_init_links(l->type, l->link, l->slipw, l->curvw, p, s);
    HIPCHK(hipMemsetAsync(l->maxXYZ, 0, 3 * sizeof(int) * l->Z, s));
    launch_lb_max_active(l->link, l->maxXYZ, p, s);
    {
        std::vector<int> mz(3 * (size_t)l->Z);
        HIPCHK(hipMemcpyAsync(mz.data(), l->maxXYZ, mz.size() * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        int m3[3] = {0, 0, 0};
        for (int z = 0; z < l->Z; ++z)
            for (int k = 0; k < 3; ++k) m3[k] = std::max(m3[k], mz[3 * z + k]);
        HIPCHK(hipMemcpy(l->maxXYZ, m3, sizeof(m3), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemsetAsync(l->f[0], 0, l->popBytes(), s));
    HIPCHK(hipMemsetAsync(l->f[1], 0, l->popBytes(), s));
    if (l->freeSurface) {
        FSBuffers& b = l->fsb;
        HIPCHK(hipMemsetAsync(b.mass, 0, 8 * l->N, s));
        HIPCHK(hipMemsetAsync(b.uS, 0, 24 * l->N, s));
        HIPCHK(hipMemsetAsync(b.hFi, 0, 24 * l->N, s));
        HIPCHK(hipMemsetAsync(b.claim, 0, 8 * l->N, s));
        if (b.actPrev) HIPCHK(hipMemsetAsync(b.actPrev, 0, 4 * l->N, s));
        HIPCHK(hipMemsetAsync(b.ctr, 0, sizeof(FSCounters), s));
        launch_fs_build_list(l->type, b, 0, p, s);
    }
    launch_lb_init_pops(l->prec, l->f[0], l->link, l->maxXYZ, l->visc, l->initVelocity, l->lbF,
                        l->freeSurface ? l->fsb.rho[0] : nullptr, l->fsb.mass, l->fsb.uS, p, s);
    launch_fill_u32(l->solid, 0u, l->N, s);
    HIPCHK(hipMemsetAsync(l->nslots, 0, 2 * sizeof(int), s));
    HIPCHK(hipMemsetAsync(l->dcount, 0, sizeof(unsigned long long) * (l->Z + 2), s));
    launch_count_active(l->link, l->dcount, p, s);
    // particle nodes at init (LB.cpp:475-495)
    if (d && d->P) lb_couple_launch(l, d);
    launch_count_mass(l->link, d && d->P && l->slotCap > 0 ? l->solid : nullptr, l->dcount + l->Z, l->params(), s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    {
        std::vector<unsigned long long> cz(l->Z + 2);
        HIPCHK(hipMemcpy(cz.data(), l->dcount, cz.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        l->activeCount = 0;
        for (int z = 0; z < l->Z; ++z) l->activeCount += cz[z];
        // total mass of active nodes outside particles (LB.cpp:209-216): fluid nodes carry mass rho0,
        // interface nodes half of it; with rho0 = 1 every partial sum of the reference's node loop is
        // a multiple of 1/2, so the counts give its sum exactly; DRUM takes the target liquid mass
        // instead (LB.cpp:206-208)
        const double tm = (double)cz[l->Z] * l->initDensity + (double)cz[l->Z + 1] * (0.5 * l->initDensity);
        l->totalMass = l->cfg.problem == HYB_PROBLEM_DRUM ? l->cfg.fluidMass / l->unit.Mass : tm;
    }
    l->steps = 0;
    l->cur = 0;
    l->time = 0;
    l->fsCur = 0;
    l->rhoCur = 0;
    l->fluidMassBuf = 0;
    l->fsTag = 0;
    l->fsSteps = 0;
    l->dem = d;
    l->inited = true;
    return HYB_OK;
}

int hyb_lb_fs_step(hyb_lb* l) {  // LB.cpp:235-245
    if (!l || !l->inited) return fail(HYB_E_STATE, "hyb_lb_fs_step before init");
    if (int rc = lb_pending(l)) return rc;
    if (!l->freeSurface) return HYB_OK;
    if (++l->fsTag >= 0x40000000u) {  // tag wrap-around (30-bit tags): clear the tagged words once
        launch_fill_u32((uint32_t*)l->fsb.claim, 0u, 2 * l->N, l->stream);
        if (l->fsb.actPrev) launch_fill_u32(l->fsb.actPrev, 0u, l->N, l->stream);
        l->fsTag = 1;
    }
    l->actTag = l->fsTag;
    const LBKParams p = l->params();
    // new interface nodes get hydroForce + lbF; lbF is zeroed by the first collision
    // when forceFieldSolve = 0 (LB.cpp:1074-1076)
    double Fnew[3];
    for (int k = 0; k < 3; ++k) Fnew[k] = (l->forceField || l->steps == 0) ? l->lbF[k] : 0.0;
    launch_fs_step(l->prec, l->f[l->cur], l->f[l->cur ^ 1], l->type, l->link, l->slipw, l->visc, l->fsb, l->fsCur, l->rhoCur,
                   l->fsTag, Fnew, l->steps > 0, p, l->stream);
    if (l->cfg.problem == HYB_PROBLEM_DRUM)  // LB.cpp:239-243
        launch_fs_mass_fix(l->prec, l->type, l->dem && l->dem->P && l->slotCap > 0 ? l->solid : nullptr, l->ctag, l->fsb, l->fsCur,
                           l->rhoCur, l->totalMass, p, l->stream);
    HIPCHK(hipGetLastError());
    l->fsCur ^= 1;
    l->fluidMassBuf = l->rhoCur;
    l->fsSteps++;
    return HYB_OK;
}

int hyb_lb_couple(hyb_lb* l, hyb_dem* d, int32_t* newNeighborList) {  // LB.cpp:247-280
    if (!l || !l->inited) return fail(HYB_E_STATE, "hyb_lb_couple before init");
    if (int rc = lb_pending(l)) return rc;
    if (int rc = dem_pending(d)) return rc;
    // The GPU re-tests every node against every particle slot at each coupling
    // step, so the rebuild handshake flag is consumed unconditionally.
    if (newNeighborList) *newNeighborList = 0;
    if (!d || !d->P) return HYB_OK;
    lb_couple_launch(l, d);
    HIPCHK(hipGetLastError());
    return HYB_OK;
}

int hyb_lb_step(hyb_lb* l, hyb_dem* d) {  // LB.cpp:221-233 fused into one kernel
    if (!l || !l->inited) return fail(HYB_E_STATE, "hyb_lb_step before init");
    if (int rc = lb_pending(l)) return rc;
    const LBKParams p = l->params();
    const bool withParticles = d && d->P && l->slotCap > 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (l->profiling) {
        if (l->evUsed + 2 > l->evs.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t ev;
                HIPCHK(hipEventCreate(&ev));
                l->evs.push_back(ev);
            }
        }
        e0 = l->evs[l->evUsed++];
        e1 = l->evs[l->evUsed++];
        HIPCHK(hipEventRecord(e0, l->stream));
    }
    FSView fv{};
    if (l->freeSurface) {
        fv.pre = l->f[l->cur ^ 1];
        fv.rhoPrev = l->fsb.rho[l->rhoCur];
        fv.rhoCur = l->fsb.rho[l->rhoCur ^ 1];
        fv.mass = l->fsb.mass;
        fv.uS = l->fsb.uS;
        fv.hFi = l->fsb.hFi;
        fv.linkw = l->link;
    }
    const HydroSink sink{l->segs, l->segCount, l->segCap};
    launch_lb_fused(l->prec, l->steps > 0, l->varVisc(), l->cfg.collision, l->f[l->cur], l->f[l->cur ^ 1], l->link,
                    withParticles ? l->solid : nullptr, l->cp, l->cpf, withParticles ? &sink : nullptr, l->visc,
                    l->capture ? l->mac : nullptr, l->freeSurface ? &fv : nullptr, p, l->stream);
    if (l->profiling) HIPCHK(hipEventRecord(e1, l->stream));
    if (withParticles) launch_dem_hydro_combine(d->p, d->b, sink, l->unit, l->stream);
    HIPCHK(hipGetLastError());
    l->cur ^= 1;
    if (l->curvw) l->macCur ^= 1;
    if (l->freeSurface) l->rhoCur ^= 1;
    l->actTag = 0;  // the next pull streams populations of this collision
    l->steps++;
    // the streaming that the next pull completes adds the curved-wall rule's extra
    // mass to the interface at the end of this step (LB.cpp:1315-1317, 1477)
    if (l->curvw && l->freeSurface) {
        const LBKParams q = l->params();
        launch_fs_curved_extra(l->prec, l->f[l->cur], l->link, l->fsb, l->fsCur, l->fsb.rho[l->fluidMassBuf],
                               l->fsSteps > 0, q, l->stream);
        HIPCHK(hipGetLastError());
    }
    l->macValid = l->capture;
    return HYB_OK;
}

int hyb_lb_set_force_field(hyb_lb* l, const float* F) {
    if (!l || !l->inited) return fail(HYB_E_STATE, "hyb_lb_set_force_field before init");
    HIPCHK(hipStreamSynchronize(l->stream));
    if (!F) {
        if (l->ffield) {
            unregister(l->saved, l->ffield);
            (void)hipFree(l->ffield);
        }
        l->ffield = nullptr;
        return HYB_OK;
    }
    const size_t N = l->N;
    std::vector<float> planes(3 * N);
    for (size_t i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) planes[(size_t)k * N + i] = F[3 * i + k];
    int rc;
    if (!l->ffield && (rc = ralloc(l->saved, "ffield", &l->ffield, 3 * N))) return rc;
    HIPCHK(hipMemcpy(l->ffield, planes.data(), 12 * N, hipMemcpyHostToDevice));
    return HYB_OK;
}

int hyb_lb_set_populations(hyb_lb* l, const double* f) {
    if (!l || !f) return fail(HYB_E_ARG, "hyb_lb_set_populations: null argument");
    if (!l->inited || l->steps != 0) return fail(HYB_E_STATE, "populations can only be set before the first step");
    HIPCHK(hipStreamSynchronize(l->stream));
    const size_t N = l->N, PS = l->planeStride();
    std::vector<uint32_t> lk(N);
    HIPCHK(hipMemcpy(lk.data(), l->link, N * 4, hipMemcpyDeviceToHost));
    if (l->prec == Prec::F64) {
        std::vector<double> pl(19 * PS, 0.0);
        for (size_t i = 0; i < N; ++i)
            if (lk[i] & LINK_ACTIVE)
                for (int j = 0; j < 19; ++j) pl[(size_t)j * PS + i] = f[i * 19 + j];
        HIPCHK(hipMemcpy(l->f[l->cur], pl.data(), pl.size() * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> pl(19 * PS, 0.0f);
        for (size_t i = 0; i < N; ++i)
            if (lk[i] & LINK_ACTIVE)
                for (int j = 0; j < 19; ++j) pl[(size_t)j * PS + i] = (float)(f[i * 19 + j] - wgt(j));
        HIPCHK(hipMemcpy(l->f[l->cur], pl.data(), pl.size() * 4, hipMemcpyHostToDevice));
    }
    l->macValid = false;
    return HYB_OK;
}

// ------------------------------------------------------------------ adjoint
static int adj_check(hyb_lb* l) {
    if (l->prec != Prec::F64) return fail(HYB_E_UNSUPPORTED, "the adjoint runs on the FP64 path");
    if (l->freeSurface) return fail(HYB_E_UNSUPPORTED, "the adjoint does not differentiate the free surface");
    if (l->varVisc()) return fail(HYB_E_UNSUPPORTED, "the adjoint needs a constant viscosity");
    if (l->hasSlip()) return fail(HYB_E_UNSUPPORTED, "the adjoint supports periodic and no-slip boundaries only");
    if (l->curvw) return fail(HYB_E_UNSUPPORTED, "the adjoint does not differentiate curved walls");
    return HYB_OK;
}

// Releases the checkpoints and the reverse pass's buffers.
static void adj_free(hyb_lb* l) {
    for (void* q : l->adjCk) (void)hipFree(q);
    for (void* q : l->adjSeg) (void)hipFree(q);
    l->adjCk.clear();
    l->adjSeg.clear();
    dfree(l->adjLam[0]); dfree(l->adjLam[1]); dfree(l->adjGF); dfree(l->adjTP);
    l->adjSteps = -1;
}

// Elapsed device time between the handle's two adjoint events.
static int adj_elapsed(hyb_lb* l, double* ms) {
    float f = 0.0f;
    HIPCHK(hipEventSynchronize(l->adjEv[1]));
    HIPCHK(hipEventElapsedTime(&f, l->adjEv[0], l->adjEv[1]));
    *ms = (double)f;
    return HYB_OK;
}

int hyb_adj_forward(hyb_lb* l, hyb_dem* d, int32_t steps, int32_t every) {
    if (!l || !l->inited) return fail(HYB_E_STATE, "hyb_adj_forward before init");
    if (steps < 1 || every < 1) return fail(HYB_E_ARG, "steps and checkpoint interval must be >= 1");
    if (l->steps != 0) return fail(HYB_E_STATE, "hyb_adj_forward starts from the initial state (before any step)");
    int rc = adj_check(l);
    if (rc) return rc;
    adj_free(l);
    const size_t pb = l->popBytes();
    const int M = (steps + every - 1) / every;  // checkpoints of s_0, s_k, ..., s_{(M-1)k}
    const int K = std::min(every, steps);        // states of one interval in the reverse pass
    auto grab = [&](std::vector<void*>& v, int n) -> int {
        for (int m = 0; m < n; ++m) {
            void* q = nullptr;
            if (hipMalloc(&q, pb) != hipSuccess) return fail(HYB_E_HIP, "adjoint buffer allocation failed");
            v.push_back(q);
        }
        return HYB_OK;
    };
    if ((rc = grab(l->adjCk, M)) || (rc = grab(l->adjSeg, K)) || (rc = dalloc(&l->adjLam[0], pb / 8)) ||
        (rc = dalloc(&l->adjLam[1], pb / 8)) || (rc = dalloc(&l->adjGF, 3 * l->N)) ||
        (rc = dalloc(&l->adjTP, adj_tau_slots(l->params())))) {
        adj_free(l);
        return rc;
    }
    for (hipEvent_t& ev : l->adjEv)
        if (!ev) HIPCHK(hipEventCreate(&ev));
    HIPCHK(hipMemcpyAsync(l->adjCk[0], l->f[l->cur], pb, hipMemcpyDeviceToDevice, l->stream));
    HIPCHK(hipEventRecord(l->adjEv[0], l->stream));
    for (int t = 1; t <= steps; ++t) {
        if ((rc = hyb_lb_step(l, d))) return rc;
        if (t % every == 0 && t / every < M)
            HIPCHK(hipMemcpyAsync(l->adjCk[t / every], l->f[l->cur], pb, hipMemcpyDeviceToDevice, l->stream));
    }
    HIPCHK(hipEventRecord(l->adjEv[1], l->stream));
    if ((rc = adj_elapsed(l, &l->adjFwdMs))) return rc;
    l->adjSteps = steps;
    l->adjEvery = every;
    return HYB_OK;
}

int hyb_adj_device_time(hyb_lb* l, double* forward_ms, double* backward_ms) {
    if (!l) return fail(HYB_E_ARG, "hyb_adj_device_time: null handle");
    if (forward_ms) *forward_ms = l->adjFwdMs;
    if (backward_ms) *backward_ms = l->adjBwdMs;
    return HYB_OK;
}

int hyb_adj_backward(hyb_lb* l, hyb_dem* d, const double* seed, double* grad_force, double* grad_tau, double* grad_f0) {
    if (!l || !l->inited || !seed) return fail(HYB_E_ARG, "hyb_adj_backward: bad arguments");
    if (l->adjSteps < 1 || l->steps != l->adjSteps) return fail(HYB_E_STATE, "hyb_adj_backward needs hyb_adj_forward just before");
    int rc = adj_check(l);
    if (rc) return rc;
    const size_t N = l->N, PS = l->planeStride(), pb = l->popBytes();
    const int T = l->adjSteps, K = l->adjEvery;
    const LBKParams p = l->params();
    const bool withParticles = d && d->P && l->slotCap > 0;
    const uint32_t* solid = withParticles ? l->solid : nullptr;
    hipStream_t s = l->stream;
    // the buffers hyb_adj_forward allocated: segment of K states, two adjoint buffers, gradients
    std::vector<void*>& seg = l->adjSeg;
    double* lam[2] = {l->adjLam[0], l->adjLam[1]};
    double* gF = l->adjGF;
    double* tp = l->adjTP;
    const size_t slots = adj_tau_slots(p);
    if ((int)seg.size() < std::min(K, T) || !lam[0] || !lam[1] || !gF || !tp)
        return fail(HYB_E_STATE, "hyb_adj_backward: reverse-pass buffers missing");
    // the seed goes through lam[1] (node-major) into lam[0] (planes)
    HIPCHK(hipMemsetAsync(lam[0], 0, pb, s));
    HIPCHK(hipMemcpyAsync(lam[1], seed, 19 * 8 * N, hipMemcpyHostToDevice, s));
    launch_nodes_to_planes(lam[1], lam[0], N, PS, 19, s);
    HIPCHK(hipMemsetAsync(lam[1], 0, pb, s));
    HIPCHK(hipMemsetAsync(gF, 0, 24 * N, s));
    HIPCHK(hipMemsetAsync(tp, 0, 8 * slots, s));
    int cur = 0;
    const int M = (int)l->adjCk.size();
    HIPCHK(hipEventRecord(l->adjEv[0], s));
    for (int m = M - 1; m >= 0; --m) {
        const int t0 = m * K;
        const int len = std::min(K, T - t0);
        HIPCHK(hipMemcpyAsync(seg[0], l->adjCk[m], pb, hipMemcpyDeviceToDevice, s));
        for (int i = 1; i < len; ++i)  // recompute s_{t0+i}
            launch_lb_fused(Prec::F64, t0 + i > 1, false, l->cfg.collision, seg[i - 1], seg[i], l->link, solid, l->cp,
                            l->cpf, nullptr, l->visc, nullptr, nullptr, p, s);
        for (int t = t0 + len; t > t0; --t) {
            launch_adj_step(l->cfg.collision, (const double*)seg[t - 1 - t0], lam[cur], lam[cur ^ 1], l->link, solid,
                            l->cp, gF, tp, t > 1, p, s);
            cur ^= 1;
        }
    }
    HIPCHK(hipEventRecord(l->adjEv[1], s));
    HIPCHK(hipGetLastError());
    if (int rc2 = adj_elapsed(l, &l->adjBwdMs)) return rc2;
    if (grad_f0) {  // node-major copy in the free adjoint buffer
        launch_planes_to_nodes(lam[cur], lam[cur ^ 1], l->link, N, PS, 19, s);
        HIPCHK(hipMemcpyAsync(grad_f0, lam[cur ^ 1], 19 * 8 * N, hipMemcpyDeviceToHost, s));
    }
    if (grad_force) {
        launch_planes_to_nodes(gF, lam[cur], nullptr, N, N, 3, s);
        HIPCHK(hipMemcpyAsync(grad_force, lam[cur], 24 * N, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    if (grad_tau) {
        std::vector<double> t(slots);
        HIPCHK(hipMemcpy(t.data(), tp, 8 * slots, hipMemcpyDeviceToHost));
        double sum = 0.0, comp = 0.0;  // Neumaier, fixed slot order
        for (double v : t) {
            const double u = sum + v;
            comp += std::fabs(sum) >= std::fabs(v) ? (sum - u) + v : (v - u) + sum;
            sum = u;
        }
        *grad_tau = sum + comp;
    }
    return HYB_OK;
}

int hyb_lb_destroy(hyb_lb* l) {
    if (!l) return HYB_OK;
    if (l->stream) (void)hipStreamSynchronize(l->stream);
    for (auto& ev : l->evs) (void)hipEventDestroy(ev);
    if (l->f[0]) (void)hipFree(l->f[0]);
    if (l->f[1]) (void)hipFree(l->f[1]);
    if (l->scratch) (void)hipFree(l->scratch);
    dfree(l->type); dfree(l->link); dfree(l->slipw); dfree(l->solid); dfree(l->visc); dfree(l->mac); dfree(l->maxXYZ);
    dfree(l->curvw); dfree(l->macBuf[0]); dfree(l->macBuf[1]);
    dfree(l->dcount); dfree(l->cp); dfree(l->cpf); dfree(l->boxes); dfree(l->nslots); dfree(l->segs);
    dfree(l->segCount);
    adj_free(l);
    for (hipEvent_t ev : l->adjEv)
        if (ev) (void)hipEventDestroy(ev);
    if (l->ffield) (void)hipFree(l->ffield);
    errword_free(l->err);
    FSBuffers& b = l->fsb;
    if (b.rho[0]) (void)hipFree(b.rho[0]);
    if (b.rho[1]) (void)hipFree(b.rho[1]);
    dfree(b.mass); dfree(b.uS); dfree(b.hFi); dfree(b.claim); dfree(b.list[0]); dfree(b.list[1]); dfree(b.listA);
    dfree(b.listB); dfree(b.mnew); dfree(b.state); dfree(b.ctr); dfree(b.actPrev);
    delete l;
    return HYB_OK;
}

int hyb_cycle(hyb_lb* l, hyb_dem* d, int32_t n) {  // hybird.cpp:35-66 without IO
    if (!l || !d) return fail(HYB_E_ARG, "hyb_cycle: null handle");
    if (int rc = dem_pending(d)) return rc;
    if (int rc = lb_pending(l)) return rc;
    const bool demSolve = l->cfg.demSolve != 0, lbSolve = l->cfg.lbSolve != 0;
    for (int k = 0; k < n; ++k) {
        ++l->time;
        int rc;
        if (demSolve && (rc = hyb_dem_step(d))) return rc;
        if (lbSolve && d->demTime > d->demInitialRepeat) {
            if (l->freeSurface && (rc = hyb_lb_fs_step(l))) return rc;
            int32_t flag = 1;
            if (demSolve && (rc = hyb_lb_couple(l, d, &flag))) return rc;
            if ((rc = hyb_lb_step(l, d))) return rc;
        } else if (lbSolve) {
            int32_t flag = 1;
            if (demSolve && (rc = hyb_lb_couple(l, d, &flag))) return rc;
        }
    }
    return HYB_OK;
}

int hyb_sync(hyb_lb* l) {
    hipStream_t s = l ? l->stream : shared_stream();
    HIPCHK(hipStreamSynchronize(s));
    if (l) {
        if (int rc = dem_pending(l->dem)) return rc;
        if (int rc = lb_pending(l)) return rc;
    }
    return HYB_OK;
}

// ------------------------------------------------------------------ host views
static int put(const void* src, size_t bytes, void* host, size_t cap, size_t* need) {
    if (need) *need = bytes;
    if (host && cap >= bytes && bytes) std::memcpy(host, src, bytes);
    return HYB_OK;
}

static int ensure_mac(hyb_lb* l, hyb_dem* d) {
    // n, u (shifted) and hydroForce of the last collision, recomputed from that
    // collision's input with the same kernel code (no population write).
    if (l->macValid) return HYB_OK;
    int rc;
    if (!l->mac && (rc = ralloc(l->saved, "mac", &l->mac, 7 * l->N))) return rc;
    LBKParams p = l->params();
    // the recomputed collision pulls with the n, u of the collision before it
    p.macPrev = l->macBuf[l->macCur ^ 1];
    p.macNext = nullptr;
    if (l->steps == 0) {
        // state after init: recompute the hydrostatic values the init kernel used
        std::vector<double> h(7 * l->N, NAN);
        std::vector<uint32_t> lk(l->N);
        std::vector<int> mx(3);
        HIPCHK(hipMemcpy(lk.data(), l->link, l->N * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(mx.data(), l->maxXYZ, 12, hipMemcpyDeviceToHost));
        const V3 F(l->lbF[0], l->lbF[1], l->lbF[2]);
        for (size_t i = 0; i < l->N; ++i) {
            if (!(lk[i] & LINK_ACTIVE)) continue;
            const int x = (int)(i % l->X), y = (int)((i / l->X) % l->Y), z = (int)(i / ((size_t)l->X * l->Y));
            double n;
            if (lk[i] & LINK_INTERFACE) n = l->initDensity;
            else {
                const V3 dh = V3(x, y, z) - V3(mx[0], mx[1], mx[2]);
                n = l->initDensity + 3.0 * 1.0 * 1.0 * l->initDensity * dh.dot(F);
            }
            const V3 u = (1.0 * F) / 2.0 / n + V3(l->initVelocity[0], l->initVelocity[1], l->initVelocity[2]);
            h[i] = n;
            h[l->N + i] = u.x; h[2 * l->N + i] = u.y; h[3 * l->N + i] = u.z;
            h[4 * l->N + i] = 0.0; h[5 * l->N + i] = 0.0; h[6 * l->N + i] = 0.0;
        }
        HIPCHK(hipMemcpy(l->mac, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    } else if (l->freeSurface) {
        // the free-surface pass prepares some collision inputs in place, so they
        // cannot be recomputed afterwards
        return fail(HYB_E_STATE, "free-surface mode: enable hyb_lb_set_capture before stepping to read n/u/hydroForce");
    } else {
        const bool withParticles = d && d->P && l->slotCap > 0;
        launch_fill_u32((uint32_t*)l->mac, 0xFFFFFFFFu, 14 * l->N, l->stream);  // NaN pattern for inactive nodes
        launch_lb_fused(l->prec, l->steps > 1, l->varVisc(), l->cfg.collision, l->f[l->cur ^ 1], nullptr, l->link,
                        withParticles ? l->solid : nullptr, l->cp, l->cpf, nullptr, l->visc, l->mac, nullptr, p,
                        l->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(l->stream));
    }
    l->macValid = true;
    return HYB_OK;
}

// Density buffer k of the free-surface state, as doubles.
static int fs_density(hyb_lb* l, int k, std::vector<double>& out) {
    out.resize(l->N);
    if (l->prec == Prec::F64) {
        HIPCHK(hipMemcpy(out.data(), l->fsb.rho[k], 8 * l->N, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> r(l->N);
        HIPCHK(hipMemcpy(r.data(), l->fsb.rho[k], 4 * l->N, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < l->N; ++i) out[i] = r[i];
    }
    return HYB_OK;
}

// Node mass of active nodes: without free surface the initial values (LB.cpp:
// 909-944); with free surface the tracked interface mass and, for fluid nodes,
// the density that the last free-surface step assigned (LB.cpp:1582-1584).
static int mass_view(hyb_lb* l, const std::vector<uint32_t>& lk, std::vector<double>& v) {
    const size_t N = l->N;
    if (!l->freeSurface || l->fsSteps == 0) {
        std::vector<double> m;
        if (l->freeSurface) {
            m.resize(N);
            HIPCHK(hipMemcpy(m.data(), l->fsb.mass, 8 * N, hipMemcpyDeviceToHost));
        }
        for (size_t i = 0; i < N; ++i)
            if (lk[i] & LINK_ACTIVE)
                v[i] = (lk[i] & LINK_INTERFACE) ? (l->freeSurface ? m[i] : 0.5 * l->initDensity) : l->initDensity;
        return HYB_OK;
    }
    std::vector<double> m(N), r;
    HIPCHK(hipMemcpy(m.data(), l->fsb.mass, 8 * N, hipMemcpyDeviceToHost));
    int rc = fs_density(l, l->fluidMassBuf, r);
    if (rc) return rc;
    for (size_t i = 0; i < N; ++i)
        if (lk[i] & LINK_ACTIVE) v[i] = (lk[i] & LINK_INTERFACE) ? m[i] : r[i];
    return HYB_OK;
}

int hyb_lb_get(hyb_lb* l, const char* cname, void* host, size_t cap, size_t* need) {
    if (!l || !cname) return fail(HYB_E_ARG, "hyb_lb_get: null argument");
    if (!l->inited) return fail(HYB_E_STATE, "hyb_lb_get before init");
    HIPCHK(hipStreamSynchronize(l->stream));
    if (int rc = lb_pending(l)) return rc;
    const std::string name(cname);
    const size_t N = l->N;
    const LBKParams p = l->params();
    std::vector<uint32_t> lk(N);
    HIPCHK(hipMemcpy(lk.data(), l->link, N * 4, hipMemcpyDeviceToHost));
    if (name == "types") {
        std::vector<uint8_t> t(N);
        HIPCHK(hipMemcpy(t.data(), l->type, N, hipMemcpyDeviceToHost));
        return put(t.data(), N, host, cap, need);
    }
    if (name == "pflag" || name == "solidIndex") {  // particle nodes are active nodes (LB.cpp:475-495)
        std::vector<uint32_t> so(N), lw(N);
        HIPCHK(hipMemcpy(so.data(), l->solid, N * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(lw.data(), l->link, N * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; ++i)
            so[i] = ((lw[i] & LINK_ACTIVE) && (so[i] >> SOLID_SLOT_BITS) == l->ctag) ? (so[i] & SOLID_SLOT_MASK) : 0u;
        if (name == "pflag") {
            std::vector<uint8_t> pf(N);
            for (size_t i = 0; i < N; ++i) pf[i] = so[i] ? 1 : 0;
            return put(pf.data(), N, host, cap, need);
        }
        for (size_t i = 0; i < N; ++i) so[i] = so[i] ? so[i] - 1 : 0xFFFFFFFFu;
        return put(so.data(), N * 4, host, cap, need);
    }
    if (name == "f") {
        if (need) *need = N * 19 * 8;
        if (!host || cap < N * 19 * 8) return HYB_OK;
        const void* src = l->f[l->cur];
        if (l->steps > 0) {
            if (!l->scratch) HIPCHK(hipMalloc(&l->scratch, l->popBytes()));
            if (l->freeSurface)
                launch_fs_stream_view(l->prec, l->f[l->cur], l->scratch, l->type, l->link, l->fsb.uS, p, l->stream);
            else
                launch_lb_stream(l->prec, l->f[l->cur], l->scratch, l->link, p, l->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(l->stream));
            src = l->scratch;
        }
        double* out = (double*)host;
        const size_t PS = l->planeStride();
        if (l->prec == Prec::F64) {
            std::vector<double> f(19 * PS);
            HIPCHK(hipMemcpy(f.data(), src, 19 * PS * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; ++i)
                for (int j = 0; j < 19; ++j) out[i * 19 + j] = (lk[i] & LINK_ACTIVE) ? f[(size_t)j * PS + i] : NAN;
        } else {
            std::vector<float> f(19 * PS);
            HIPCHK(hipMemcpy(f.data(), src, 19 * PS * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; ++i)
                for (int j = 0; j < 19; ++j)
                    out[i * 19 + j] = (lk[i] & LINK_ACTIVE) ? (double)f[(size_t)j * PS + i] + wgt(j) : NAN;
        }
        return HYB_OK;
    }
    if (name == "n" && l->freeSurface) {
        if (need) *need = N * 8;
        if (!host || cap < N * 8) return HYB_OK;
        std::vector<double> r;
        int rc = fs_density(l, l->rhoCur, r);
        if (rc) return rc;
        double* out = (double*)host;
        for (size_t i = 0; i < N; ++i) out[i] = (lk[i] & LINK_ACTIVE) ? r[i] : NAN;
        return HYB_OK;
    }
    if (name == "n" || name == "u" || name == "hydroForce") {
        const size_t w = name == "n" ? 1 : 3;
        if (need) *need = N * w * 8;
        if (!host || cap < N * w * 8) return HYB_OK;
        int rc = ensure_mac(l, l->dem);
        if (rc) return rc;
        std::vector<double> m(7 * N);
        HIPCHK(hipMemcpy(m.data(), l->mac, 7 * N * 8, hipMemcpyDeviceToHost));
        double* out = (double*)host;
        const size_t off = name == "n" ? 0 : (name == "u" ? 1 : 4);
        for (size_t i = 0; i < N; ++i)
            for (size_t k = 0; k < w; ++k) out[i * w + k] = (lk[i] & LINK_ACTIVE) ? m[(off + k) * N + i] : NAN;
        return HYB_OK;
    }
    if (name == "visc" || name == "mass") {
        std::vector<double> v(N, NAN);
        if (name == "visc") {
            if (l->visc) HIPCHK(hipMemcpy(v.data(), l->visc, N * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; ++i) v[i] = (lk[i] & LINK_ACTIVE) ? (l->visc ? v[i] : l->initDynVisc) : NAN;
        } else {
            int rc = mass_view(l, lk, v);
            if (rc) return rc;
        }
        return put(v.data(), N * 8, host, cap, need);
    }
    if (name == "totalMass") return put(&l->totalMass, 8, host, cap, need);
    if (name == "fluidMass") {  // IO::totFluidMass (IO.cpp:987-999), summed in node order
        std::vector<double> v(N, NAN);
        int rc = mass_view(l, lk, v);
        if (rc) return rc;
        std::vector<uint32_t> so(N);
        HIPCHK(hipMemcpy(so.data(), l->solid, N * 4, hipMemcpyDeviceToHost));
        double m = 0.0;
        for (size_t i = 0; i < N; ++i) {
            const bool owned = (so[i] >> SOLID_SLOT_BITS) == l->ctag && (so[i] & SOLID_SLOT_MASK);
            if ((lk[i] & LINK_ACTIVE) && !owned) m += v[i];
        }
        return put(&m, 8, host, cap, need);
    }
    if (name == "status") {
        // IO's status-line quantities (IO.cpp:835-850, 969-985): max |u|^2 over active
        // nodes (u of the last collision), active nodes, active nodes with viscosity
        // above 0.95 (maxTau - 0.5) / 3 (only with variable viscosity)
        int rc = ensure_mac(l, l->dem);
        if (rc) return rc;
        const double maxVisc = (1.8 - 0.5) / 3 / 1.0;
        unsigned long long* acc = nullptr;
        HIPCHK(hipMalloc((void**)&acc, 3 * 8));
        HIPCHK(hipMemsetAsync(acc, 0, 3 * 8, l->stream));
        launch_status(l->mac, l->varVisc() ? l->visc : nullptr, l->link, N, 0.95 * maxVisc, acc, l->stream);
        unsigned long long a[3] = {0, 0, 0};
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(a, acc, 3 * 8, hipMemcpyDeviceToHost, l->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(l->stream);
        (void)hipFree(acc);
        if (e != hipSuccess) return fail(HYB_E_HIP, std::string("status: ") + hipGetErrorString(e));
        double v3[3];
        std::memcpy(&v3[0], &a[0], 8);
        v3[1] = (double)a[1];
        v3[2] = (double)a[2];
        return put(v3, sizeof(v3), host, cap, need);
    }
    if (name == "wallFHydro") {  // walls[].FHydro of the last LB step (LB.cpp:1186-1189, 1321-1341, 1485-1487)
        const hyb_dem* d = l->dem;
        const size_t W = d ? d->walls.size() : 0;
        std::vector<double> F(3 * W, 0.0);
        if (W && l->steps > 0) {
            WallSet ws;
            ws.n = (int)std::min<size_t>(W, 6);
            ws.unitLength = l->unit.Length;
            for (int k = 0; k < ws.n; ++k) ws.w[k] = d->walls[k];
            unsigned long long* acc = nullptr;
            const size_t nacc = 3 * 4 * W;
            HIPCHK(hipMalloc((void**)&acc, nacc * 8));
            HIPCHK(hipMemsetAsync(acc, 0, nacc * 8, l->stream));
            launch_wall_fhydro(l->prec, l->f[l->cur], l->type, ws, acc, p, l->stream);
            std::vector<unsigned long long> a(nacc);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpyAsync(a.data(), acc, nacc * 8, hipMemcpyDeviceToHost, l->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(l->stream);
            (void)hipFree(acc);
            if (e != hipSuccess) return fail(HYB_E_HIP, std::string("wallFHydro: ") + hipGetErrorString(e));
            for (size_t k = 0; k < 3 * W; ++k) F[k] = exact_value_host(&a[4 * k]) * l->unit.Force;
        }
        return put(F.data(), F.size() * 8, host, cap, need);
    }
    return fail(HYB_E_ARG, "unknown LB field " + name);
}

// ------------------------------------------------------------------ no-allocation entry point
// walls[].FHydro into caller-provided device memory, for callers that cannot
// allocate (code under graph capture, serving loops). The exact-sum limbs live
// in `scratch`, a fixed device array (one copy per GPU, kept by the process):
// its device address is looked up once per device with hipGetSymbolAddress and
// only that address is handed to kernels and fills. Calls on one GPU share the
// array, so calls that overlap on one GPU race.
static __device__ __align__(256) unsigned long long scratch[(size_t)HYB_WALL_FHYDRO_MAX_WALLS * 3 * 4];

static int scratch_address(unsigned long long** out) {
    constexpr int MAX_DEVICES = 64;
    static std::mutex lock;
    static void* addr[MAX_DEVICES] = {};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(HYB_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> guard(lock);
    if (!addr[dev]) {
        void* q = nullptr;
        HIPCHK(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        addr[dev] = q;
    }
    *out = (unsigned long long*)addr[dev];
    return HYB_OK;
}

int hyb_lb_wall_fhydro_into(hyb_lb* l, double* d_out, int32_t n_out) {
    if (!l || !d_out) return fail(HYB_E_ARG, "hyb_lb_wall_fhydro_into: null argument");
    if (!l->inited) return fail(HYB_E_STATE, "hyb_lb_wall_fhydro_into before init");
    if (int rc = lb_pending(l)) return rc;
    const hyb_dem* d = l->dem;
    const int W = d ? (int)d->walls.size() : 0;
    if (W > HYB_WALL_FHYDRO_MAX_WALLS)
        return fail(HYB_E_CAPACITY, "hyb_lb_wall_fhydro_into: " + std::to_string(W) + " walls, at most " +
                                        std::to_string(HYB_WALL_FHYDRO_MAX_WALLS));
    if (n_out < 3 * W) return fail(HYB_E_ARG, "hyb_lb_wall_fhydro_into: output holds fewer than 3 values per wall");
    if (!W) return HYB_OK;
    if (l->steps == 0) {  // no LB step yet: zero force, as the view
        launch_exact_finalize(nullptr, 3 * W, 0.0, d_out, l->stream);
        HIPCHK(hipGetLastError());
        return HYB_OK;
    }
    unsigned long long* acc = nullptr;
    if (int rc = scratch_address(&acc)) return rc;
    WallSet ws;
    ws.n = W;
    ws.unitLength = l->unit.Length;
    for (int k = 0; k < W; ++k) ws.w[k] = d->walls[k];
    // a fill kernel, not a memset node, so the sequence can be captured into a graph
    launch_fill_u32((uint32_t*)acc, 0u, (size_t)W * 3 * 4 * 2, l->stream);
    launch_wall_fhydro(l->prec, l->f[l->cur], l->type, ws, acc, l->params(), l->stream);
    launch_exact_finalize(acc, 3 * W, l->unit.Force, d_out, l->stream);
    HIPCHK(hipGetLastError());
    return HYB_OK;
}

int hyb_dem_saved(hyb_dem* d, int32_t* count) {
    if (!d || !count) return fail(HYB_E_ARG, "hyb_dem_saved: null argument");
    *count = (int32_t)d->saveEvents.size();
    return HYB_OK;
}

int hyb_dem_saved_get(hyb_dem* d, int32_t event, const char* cpart, void* host, size_t cap, size_t* need) {
    if (!d || !cpart) return fail(HYB_E_ARG, "hyb_dem_saved_get: null argument");
    if (event < 0 || event >= (int32_t)d->saveEvents.size()) return fail(HYB_E_ARG, "hyb_dem_saved_get: no such capture");
    HIPCHK(hipStreamSynchronize(d->stream));
    const std::string part(cpart);
    const size_t P = (size_t)d->P;
    double nPairs = 0.0;
    if (P) HIPCHK(hipMemcpy(&nPairs, d->saveInfo + event, sizeof(double), hipMemcpyDeviceToHost));
    if (part == "info") {
        const double v[10] = {d->saveEvents[event].first, (double)d->saveEvents[event].second, nPairs, (double)d->P,
                              (double)d->stdObjects, d->dampCoeff, d->deltat, d->demSize[0], d->demSize[1],
                              d->demSize[2]};
        return put(v, sizeof(v), host, cap, need);
    }
    auto rows = [&](const void* dev, size_t bytes) -> int {
        if (need) *need = bytes;
        if (host && cap >= bytes && bytes) HIPCHK(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
        return HYB_OK;
    };
    if (part == "elements") return rows(d->saveElems + (size_t)event * P * 9, P * 9 * sizeof(double));
    if (part == "pairs")
        return rows(d->savePairs + (size_t)event * d->p.pairCap, (size_t)nPairs * sizeof(StatRec));
    if (part == "walls") {
        const size_t ws = (size_t)wall_slots(d->p);
        return rows(d->saveWalls + (size_t)event * P * ws, P * ws * sizeof(StatRec));
    }
    if (part == "objects")
        return rows(d->saveObjs ? (const void*)(d->saveObjs + (size_t)event * P * NEAR_OBJ) : nullptr,
                    d->saveObjs ? P * NEAR_OBJ * sizeof(StatRec) : 0);
    return fail(HYB_E_ARG, "unknown capture part " + part);
}

int hyb_dem_get(hyb_dem* d, const char* cname, void* host, size_t cap, size_t* need) {
    if (!d || !cname) return fail(HYB_E_ARG, "hyb_dem_get: null argument");
    if (!d->inited) return fail(HYB_E_STATE, "hyb_dem_get before init");
    HIPCHK(hipStreamSynchronize(d->stream));
    const std::string name(cname);
    const int P = d->P, NE = d->E;
    DEMState st;
    std::memset(&st, 0, sizeof(st));
    if (P) HIPCHK(hipMemcpy(&st, d->b.st, sizeof(st), hipMemcpyDeviceToHost));
    if (int rc = dem_pending(d)) return rc;
    std::vector<double> el;
    if (P) {
        el.resize((size_t)EF_COUNT * NE);
        HIPCHK(hipMemcpy(el.data(), d->b.el, el.size() * 8, hipMemcpyDeviceToHost));
    }
    auto at = [&](int f, int e) { return el[(size_t)f * NE + e]; };
    std::vector<double> v;
    std::vector<uint32_t> u;
    const int G = P ? st.nGhost : 0;
    std::vector<uint32_t> gi(G);
    std::vector<double> gv(3 * (size_t)G);
    if (G) {
        HIPCHK(hipMemcpy(gi.data(), d->b.ghostIndex, 4 * (size_t)G, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(gv.data(), d->b.ghostVec, 24 * (size_t)G, hipMemcpyDeviceToHost));
    }
    if (name == "nearObjectTable") {  // DEM.cpp:1515-1532: (particle, object) pairs, particle-major
        std::vector<int> cnt(P ? (size_t)P : 0);
        std::vector<uint32_t> lst(P ? (size_t)P * NEAR_OBJ : 0);
        if (P && d->p.nobj) {
            HIPCHK(hipMemcpy(cnt.data(), d->b.nearObjCount, 4 * (size_t)P, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(lst.data(), d->b.nearObj, 4 * lst.size(), hipMemcpyDeviceToHost));
            for (int e = 0; e < P; ++e)
                for (int k = 0; k < cnt[e]; ++k) { u.push_back((uint32_t)e); u.push_back(lst[(size_t)e * NEAR_OBJ + k]); }
        }
        return put(u.data(), u.size() * 4, host, cap, need);
    }
    if (name == "objects") {  // index, r, x0, x1 of every object (originals, then periodic copies)
        for (const ObjDev& o : d->objects) {
            v.push_back((double)o.index); v.push_back(o.r);
            v.push_back(o.x0.x); v.push_back(o.x0.y); v.push_back(o.x0.z);
            v.push_back(o.x1.x); v.push_back(o.x1.y); v.push_back(o.x1.z);
        }
        return put(v.data(), v.size() * 8, host, cap, need);
    }
    if (name == "walls" || name == "objectForces") {
        // walls: p, n, FParticle (DEM.cpp:1123-1125, 1936-1979) and FHydro (0 here; the
        // lattice's "wallFHydro"); objectForces: FParticle of every object (DEM.cpp:2100-2134).
        // Contacts of the last sub-step, applied in the reference's list order.
        const bool wallView = name == "walls";
        const size_t rows = wallView ? d->walls.size() : d->objects.size();
        std::vector<V3> F(rows, V3(0.0, 0.0, 0.0));
        if (P && d->substeps > 0) {
            const size_t per = wallView ? (size_t)wall_slots(d->p) : NEAR_OBJ;
            std::vector<ContactForce> cf((size_t)P * per);
            std::vector<uint32_t> idx((size_t)P * per);
            std::vector<int> cnt((size_t)P, (int)per);
            if (wallView) {
                HIPCHK(hipMemcpy(cf.data(), d->b.wcf, sizeof(ContactForce) * cf.size(), hipMemcpyDeviceToHost));
                if (d->p.allWalls) {  // record w of a particle is its contact with wall w, if that wall is near
                    std::vector<uint32_t> mask((size_t)P);
                    HIPCHK(hipMemcpy(mask.data(), d->b.nearWallMask, 4 * mask.size(), hipMemcpyDeviceToHost));
                    for (size_t e = 0; e < (size_t)P; ++e)
                        for (size_t w = 0; w < per; ++w) idx[e * per + w] = (mask[e] >> w & 1u) ? (uint32_t)w : 0xFFFFFFFFu;
                } else {
                    HIPCHK(hipMemcpy(idx.data(), d->b.nearWall, 4 * idx.size(), hipMemcpyDeviceToHost));
                }
            } else if (d->p.nobj) {
                HIPCHK(hipMemcpy(cf.data(), d->b.ocf, sizeof(ContactForce) * cf.size(), hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(idx.data(), d->b.nearObj, 4 * idx.size(), hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(cnt.data(), d->b.nearObjCount, 4 * cnt.size(), hipMemcpyDeviceToHost));
            } else {
                cnt.assign((size_t)P, 0);
            }
            for (int e = 0; e < P; ++e)
                for (int k = 0; k < cnt[e]; ++k) {
                    const uint32_t w = idx[(size_t)e * per + k];
                    const ContactForce& c = cf[(size_t)e * per + k];
                    if (w >= rows || c.flag == 0.0) continue;
                    F[w] = F[w] - V3(c.fn[0], c.fn[1], c.fn[2]);
                    if (c.flag == 3.0) F[w] = F[w] + V3(c.ft[0], c.ft[1], c.ft[2]);
                }
        }
        for (size_t w = 0; w < rows; ++w) {
            if (wallView) {
                const WallDev& wd = d->walls[w];
                v.push_back(wd.p.x); v.push_back(wd.p.y); v.push_back(wd.p.z);
                v.push_back(wd.n.x); v.push_back(wd.n.y); v.push_back(wd.n.z);
            }
            v.push_back(F[w].x); v.push_back(F[w].y); v.push_back(F[w].z);
            if (wallView) { v.push_back(0.0); v.push_back(0.0); v.push_back(0.0); }
        }
        return put(v.data(), v.size() * 8, host, cap, need);
    }
    if (name == "energy") {  // DEM::particleEnergy, updated after every sub-step (DEM.cpp:371-372)
        double e6[6] = {0, 0, 0, 0, 0, 0};
        if (d->substeps > 0 && d->b.energy) {
            launch_dem_energy(d->p, d->b, d->demSize, d->b.energy, d->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(e6, d->b.energy, sizeof(e6), hipMemcpyDeviceToHost, d->stream));
            HIPCHK(hipStreamSynchronize(d->stream));
        }
        return put(e6, sizeof(e6), host, cap, need);
    }
    if (name == "e_x" || name == "e_xp" || name == "e_w" || name == "e_wp") {
        // at step end the predicted copies equal the corrected ones (elmt.cpp:190-224)
        const int base = (name == "e_x" || name == "e_xp") ? EF_X : EF_W;
        for (int e = 0; e < NE; ++e)
            for (int k = 0; k < 18; ++k) v.push_back(at(base + k, e));
        return put(v.data(), v.size() * 8, host, cap, need);
    }
    if (name == "e_q" || name == "e_qp") {
        for (int e = 0; e < NE; ++e)
            for (int k = 0; k < 24; ++k) v.push_back(at(EF_Q + k, e));
        return put(v.data(), v.size() * 8, host, cap, need);
    }
    if (name == "e_forces") {
        for (int e = 0; e < NE; ++e)
            for (int k = 0; k < 21; ++k) v.push_back(at(EF_FH + k, e));
        return put(v.data(), v.size() * 8, host, cap, need);
    }
    if (name == "e_misc") {
        for (int e = 0; e < NE; ++e) {
            v.push_back(at(EF_M, e)); v.push_back(at(EF_RAD, e));
            for (int k = 0; k < 3; ++k) v.push_back(at(EF_I + k, e));
            for (int k = 0; k < 12; ++k) v.push_back(at(EF_WG + k, e));
            v.push_back(at(EF_FVOL, e)); v.push_back(d->rec[(size_t)e * HYB_ELEMENT_RECORD + 1]);
        }
        return put(v.data(), v.size() * 8, host, cap, need);
    }
    if (name == "p_data" || name == "p_idx") {
        // per standard particle: position, velocity, radiusVec after the corrector, element, prototype index
        std::vector<double> pc;
        std::vector<int> pe, pf;
        if (P && d->p.clumps) {
            pc.resize(9 * (size_t)P);
            pe.resize((size_t)P);
            pf.resize((size_t)NE + 1);
            HIPCHK(hipMemcpy(pc.data(), d->b.pcor, 8 * pc.size(), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(pe.data(), d->b.pelem, 4 * pe.size(), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(pf.data(), d->b.efirst, 4 * pf.size(), hipMemcpyDeviceToHost));
        }
        for (int s = 0; s < P + G; ++s) {
            const int o = s < P ? s : (int)gi[s - P];  // standard particle
            const int e = d->p.clumps ? pe[o] : o;
            if (name == "p_data") {
                // ghosts: origin position + period vector, the origin's velocity and radiusVec (elmt.cpp:292-298)
                for (int k = 0; k < 3; ++k) {
                    const double x = d->p.clumps ? pc[9 * (size_t)o + k] : at(EF_X + k, o);
                    v.push_back(s < P ? x : x + gv[3 * (s - P) + k]);
                }
                for (int k = 0; k < 3; ++k) v.push_back(d->p.clumps ? pc[9 * (size_t)o + 3 + k] : at(EF_X + 3 + k, o));
                for (int k = 0; k < 3; ++k) v.push_back(d->p.clumps ? pc[9 * (size_t)o + 6 + k] : 0.0);
                v.push_back(at(EF_RAD, e));
            } else {
                u.push_back((uint32_t)s); u.push_back((uint32_t)e);
                u.push_back(d->p.clumps ? (uint32_t)(o - pf[e]) : 0u); u.push_back(0u);
            }
        }
        if (name == "p_data") return put(v.data(), v.size() * 8, host, cap, need);
        return put(u.data(), u.size() * 4, host, cap, need);
    }
    if (name == "neighborTable") {
        u.resize(2 * (size_t)(P ? st.nPairs : 0));
        if (!u.empty()) HIPCHK(hipMemcpy(u.data(), d->b.pairs, u.size() * 4, hipMemcpyDeviceToHost));
        return put(u.data(), u.size() * 4, host, cap, need);
    }
    if (name == "nearWallTable") {
        std::vector<uint32_t> nw(P);
        if (P) HIPCHK(hipMemcpy(nw.data(), d->p.allWalls ? d->b.nearWallMask : d->b.nearWall, 4 * (size_t)P,
                                hipMemcpyDeviceToHost));
        for (int e = 0; e < P; ++e) {
            if (d->p.allWalls) {  // every near wall, in wall order
                for (int w = 0; w < d->p.nwalls; ++w)
                    if (nw[e] >> w & 1u) { u.push_back((uint32_t)e); u.push_back((uint32_t)w); }
            } else if (nw[e] != 0xFFFFFFFFu) {
                u.push_back((uint32_t)e); u.push_back(nw[e]);
            }
        }
        return put(u.data(), u.size() * 4, host, cap, need);
    }
    if (name == "ghostIndex") return put(gi.data(), gi.size() * 4, host, cap, need);
    if (name == "ghostVec") return put(gv.data(), gv.size() * 8, host, cap, need);
    if (name == "dem_scalars") {
        const DEMParams& p = d->p;
        v = {P ? st.maxDisp : 0.0, p.nebrRange, p.cellWidth[0], p.cellWidth[1], p.cellWidth[2], (double)p.nCells[0],
             (double)p.nCells[1], (double)p.nCells[2], d->demTime, (double)d->demTimeStep, d->deltat, (double)d->multiStep,
             (double)P, d->dampCoeff, d->knConst, d->ksConst, (double)st.newNeighborList};
        return put(v.data(), v.size() * 8, host, cap, need);
    }
    if (name == "rebuilds") {
        const double r = st.rebuilds;
        return put(&r, 8, host, cap, need);
    }
    return fail(HYB_E_ARG, "unknown DEM field " + name);
}

int hyb_copy_bandwidth(size_t bytes, int32_t width, int32_t reps, double* gbps) {
    if (!gbps || reps < 1 || bytes < 4096 || bytes % 16) return fail(HYB_E_ARG, "hyb_copy_bandwidth: bad arguments");
    if (width != 4 && width != 8 && width != 16) return fail(HYB_E_ARG, "hyb_copy_bandwidth: width must be 4, 8 or 16");
    hipStream_t s = shared_stream();
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = [&]() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
    };
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        release();
        return fail(HYB_E_HIP, std::string("hyb_copy_bandwidth: ") + hipGetErrorString(e));
    }
    launch_copy_probe(width, a, b, bytes, s);  // warm-up
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r) launch_copy_probe(width, a, b, bytes, s);
    (void)hipEventRecord(e1, s);
    e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) e = hipGetLastError();
    release();
    if (e != hipSuccess) return fail(HYB_E_HIP, std::string("hyb_copy_bandwidth: ") + hipGetErrorString(e));
    *gbps = 2.0 * (double)bytes * reps / ((double)ms * 1e-3) / 1e9;
    return HYB_OK;
}

int hyb_lb_set_profiling(hyb_lb* l, int32_t enabled) {
    if (!l) return fail(HYB_E_ARG, "null handle");
    l->profiling = enabled != 0;
    l->evUsed = 0;
    l->profMs = 0.0;
    l->profLaunches = 0;
    return HYB_OK;
}

int hyb_lb_set_capture(hyb_lb* l, int32_t enabled) {
    if (!l) return fail(HYB_E_ARG, "null handle");
    if (enabled && !l->mac) {
        int rc = ralloc(l->saved, "mac", &l->mac, 7 * l->N);
        if (rc) return rc;
    }
    l->capture = enabled != 0;
    if (!l->capture && l->freeSurface) l->macValid = false;
    return HYB_OK;
}

int hyb_lb_kernel_time(hyb_lb* l, double* total_ms, int64_t* launches) {
    if (!l) return fail(HYB_E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(l->stream));
    for (size_t k = 0; k + 1 < l->evUsed; k += 2) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, l->evs[k], l->evs[k + 1]));
        l->profMs += ms;
        l->profLaunches += 1;
    }
    l->evUsed = 0;
    if (total_ms) *total_ms = l->profMs;
    if (launches) *launches = l->profLaunches;
    return HYB_OK;
}

int hyb_lb_active_nodes(hyb_lb* l, int64_t* count) {
    if (!l || !count) return fail(HYB_E_ARG, "null argument");
    *count = (int64_t)l->activeCount;
    return HYB_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ state file
// Layout: "HYBSTATE", u32 version, u32 sections (bit 0: lattice, bit 1: DEM);
// per section its scalar record, u32 block count, then per block u32 name
// length, name, u64 bytes, raw bytes (the device buffer as it is).
namespace {

constexpr char STATE_MAGIC[8] = {'H', 'Y', 'B', 'S', 'T', 'A', 'T', 'E'};
constexpr uint32_t STATE_VERSION = 1;
constexpr size_t STATE_CHUNK = size_t(64) << 20;

struct LbScalars {
    int32_t X, Y, Z, prec;
    uint64_t N, popBytes;
    int32_t cur, fsCur, rhoCur, fluidMassBuf;
    uint32_t ctag, fsTag, actTag, time;
    int64_t steps, fsSteps;
    uint64_t activeCount;
    double totalMass;
    uint8_t capture, macValid, havePart, macCur, pad[4];
};

struct DemScalars {
    int32_t P, pairCap, Gcap, multiStep;
    double demTime;
    uint32_t demTimeStep, E;
    int64_t substeps;
};

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

bool wr(FILE* f, const void* p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
bool rd(FILE* f, void* p, size_t n) { return std::fread(p, 1, n, f) == n; }

int write_blocks(FILE* f, const std::vector<Blk>& reg) {
    const uint32_t nb = (uint32_t)reg.size();
    if (!wr(f, &nb, 4)) return fail(HYB_E_ARG, "state file: write error");
    std::vector<char> buf;
    for (const Blk& b : reg) {
        const uint32_t nl = (uint32_t)b.name.size();
        const uint64_t nbytes = b.bytes;
        if (!wr(f, &nl, 4) || !wr(f, b.name.data(), nl) || !wr(f, &nbytes, 8))
            return fail(HYB_E_ARG, "state file: write error");
        for (size_t off = 0; off < b.bytes; off += STATE_CHUNK) {
            const size_t n = std::min(STATE_CHUNK, b.bytes - off);
            buf.resize(n);
            HIPCHK(hipMemcpy(buf.data(), (const char*)b.ptr + off, n, hipMemcpyDeviceToHost));
            if (!wr(f, buf.data(), n)) return fail(HYB_E_ARG, "state file: write error");
        }
    }
    return HYB_OK;
}

// Reads the blocks of one section into the registered buffers; `lazy` may
// allocate a buffer the handle creates only on demand (returns nullptr if the
// name is not one of those).
template <class Lazy>
int read_blocks(FILE* f, std::vector<Blk>& reg, Lazy&& lazy) {
    uint32_t nb = 0;
    if (!rd(f, &nb, 4)) return fail(HYB_E_ARG, "state file: truncated");
    std::vector<char> buf;
    std::vector<bool> seen(reg.size(), false);
    for (uint32_t k = 0; k < nb; ++k) {
        uint32_t nl = 0;
        uint64_t nbytes = 0;
        if (!rd(f, &nl, 4) || nl > 64) return fail(HYB_E_ARG, "state file: corrupt block header");
        std::string name(nl, '\0');
        if (!rd(f, &name[0], nl) || !rd(f, &nbytes, 8)) return fail(HYB_E_ARG, "state file: truncated");
        int at = -1;
        for (size_t i = 0; i < reg.size(); ++i)
            if (reg[i].name == name) at = (int)i;
        if (at < 0) {
            const int rc = lazy(name);
            if (rc) return rc;
            for (size_t i = 0; i < reg.size(); ++i)
                if (reg[i].name == name) at = (int)i;
            seen.resize(reg.size(), false);
        }
        if (at < 0) return fail(HYB_E_ARG, "state file: buffer '" + name + "' does not exist in this configuration");
        if (reg[at].bytes != nbytes)
            return fail(HYB_E_ARG, "state file: buffer '" + name + "' has " + std::to_string(nbytes) + " bytes, this handle " +
                                       std::to_string(reg[at].bytes));
        for (size_t off = 0; off < nbytes; off += STATE_CHUNK) {
            const size_t n = std::min(STATE_CHUNK, (size_t)nbytes - off);
            buf.resize(n);
            if (!rd(f, buf.data(), n)) return fail(HYB_E_ARG, "state file: truncated");
            HIPCHK(hipMemcpy((char*)reg[at].ptr + off, buf.data(), n, hipMemcpyHostToDevice));
        }
        seen[at] = true;
    }
    for (size_t i = 0; i < reg.size(); ++i)
        if (!seen[i] && reg[i].name != "mac" && reg[i].name != "ffield")
            return fail(HYB_E_ARG, "state file: buffer '" + reg[i].name + "' missing");
    return HYB_OK;
}

}  // namespace

extern "C" {

int hyb_state_save(hyb_lb* l, hyb_dem* d, const char* path) {
    if (!path || (!l && !d)) return fail(HYB_E_ARG, "hyb_state_save: null argument");
    if ((l && !l->inited) || (d && !d->inited)) return fail(HYB_E_STATE, "hyb_state_save before init");
    if (l) HIPCHK(hipStreamSynchronize(l->stream));
    if (d) HIPCHK(hipStreamSynchronize(d->stream));
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return fail(HYB_E_ARG, std::string("hyb_state_save: cannot open ") + path);
    const uint32_t sections = (l ? 1u : 0u) | (d ? 2u : 0u);
    if (!wr(fh.f, STATE_MAGIC, 8) || !wr(fh.f, &STATE_VERSION, 4) || !wr(fh.f, &sections, 4))
        return fail(HYB_E_ARG, "state file: write error");
    int rc;
    if (l) {
        LbScalars s;
        std::memset(&s, 0, sizeof(s));
        s.X = l->X; s.Y = l->Y; s.Z = l->Z; s.prec = (int32_t)l->prec;
        s.N = l->N; s.popBytes = l->popBytes();
        s.cur = l->cur; s.fsCur = l->fsCur; s.rhoCur = l->rhoCur; s.fluidMassBuf = l->fluidMassBuf;
        s.ctag = l->ctag; s.fsTag = l->fsTag; s.actTag = l->actTag; s.time = l->time;
        s.steps = l->steps; s.fsSteps = l->fsSteps;
        s.activeCount = l->activeCount; s.totalMass = l->totalMass;
        s.capture = l->capture; s.macValid = l->macValid; s.havePart = l->havePart; s.macCur = (uint8_t)l->macCur;
        if (!wr(fh.f, &s, sizeof(s))) return fail(HYB_E_ARG, "state file: write error");
        if ((rc = write_blocks(fh.f, l->saved))) return rc;
    }
    if (d) {
        DemScalars s;
        std::memset(&s, 0, sizeof(s));
        s.P = d->P; s.E = (uint32_t)d->E; s.pairCap = d->p.pairCap; s.Gcap = d->p.Gcap; s.multiStep = d->multiStep;
        s.demTime = d->demTime; s.demTimeStep = d->demTimeStep; s.substeps = d->substeps;
        if (!wr(fh.f, &s, sizeof(s))) return fail(HYB_E_ARG, "state file: write error");
        if ((rc = write_blocks(fh.f, d->saved))) return rc;
    }
    if (std::fflush(fh.f) != 0) return fail(HYB_E_ARG, "state file: write error");
    return HYB_OK;
}

int hyb_state_load(hyb_lb* l, hyb_dem* d, const char* path) {
    if (!path || (!l && !d)) return fail(HYB_E_ARG, "hyb_state_load: null argument");
    if ((l && !l->inited) || (d && !d->inited)) return fail(HYB_E_STATE, "hyb_state_load before init");
    if (l) HIPCHK(hipStreamSynchronize(l->stream));
    if (d) HIPCHK(hipStreamSynchronize(d->stream));
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return fail(HYB_E_ARG, std::string("hyb_state_load: cannot open ") + path);
    char magic[8];
    uint32_t version = 0, sections = 0;
    if (!rd(fh.f, magic, 8) || std::memcmp(magic, STATE_MAGIC, 8) != 0) return fail(HYB_E_ARG, "not a state file");
    if (!rd(fh.f, &version, 4) || version != STATE_VERSION || !rd(fh.f, &sections, 4))
        return fail(HYB_E_ARG, "state file: unsupported version");
    if ((sections & 1u) != (l ? 1u : 0u) || (sections & 2u) != (d ? 2u : 0u))
        return fail(HYB_E_ARG, "state file: holds a different set of handles");
    int rc;
    if (l) {
        LbScalars s;
        if (!rd(fh.f, &s, sizeof(s))) return fail(HYB_E_ARG, "state file: truncated");
        if (s.X != l->X || s.Y != l->Y || s.Z != l->Z || s.prec != (int32_t)l->prec || s.popBytes != l->popBytes())
            return fail(HYB_E_ARG, "state file: lattice size or precision differs");
        // buffer selectors index two-element arrays
        if ((s.cur | s.fsCur | s.rhoCur | s.fluidMassBuf) & ~1 || s.macCur > 1)
            return fail(HYB_E_ARG, "state file: corrupt buffer selector");
        auto lazy = [&](const std::string& name) -> int {
            if (name == "mac" && !l->mac) return ralloc(l->saved, "mac", &l->mac, 7 * l->N);
            if (name == "ffield" && !l->ffield) return ralloc(l->saved, "ffield", &l->ffield, 3 * l->N);
            return HYB_OK;
        };
        if ((rc = read_blocks(fh.f, l->saved, lazy))) {
            l->inited = false;  // partly overwritten: unusable until destroyed
            return rc;
        }
        l->macCur = s.macCur;
        errword_clear(l->err);
        l->cur = s.cur; l->fsCur = s.fsCur; l->rhoCur = s.rhoCur; l->fluidMassBuf = s.fluidMassBuf;
        l->ctag = s.ctag; l->fsTag = s.fsTag; l->actTag = s.actTag; l->time = s.time;
        l->steps = s.steps; l->fsSteps = s.fsSteps;
        l->activeCount = s.activeCount; l->totalMass = s.totalMass;
        l->capture = s.capture != 0; l->macValid = s.macValid != 0; l->havePart = s.havePart != 0;
        // per-step scratch: the run-sum lists start empty
        if (l->segCount) HIPCHK(hipMemset(l->segCount, 0, sizeof(uint32_t) * (size_t)l->slotCap));
        adj_free(l);
    }
    if (d) {
        // the lattice (if any) already holds the file's state: a DEM section that fails leaves the pair
        // inconsistent, so both handles are then unusable until destroyed
        DemScalars s;
        if (!rd(fh.f, &s, sizeof(s))) {
            if (l) l->inited = false;
            return fail(HYB_E_ARG, "state file: truncated");
        }
        if (s.P != d->P || s.E != (uint32_t)d->E || s.pairCap != d->p.pairCap || s.Gcap != d->p.Gcap ||
            s.multiStep != d->multiStep) {
            if (l) l->inited = false;
            return fail(HYB_E_ARG, "state file: particle set or DEM configuration differs");
        }
        if ((rc = read_blocks(fh.f, d->saved, [](const std::string&) { return (int)HYB_OK; }))) {
            if (l) l->inited = false;
            d->inited = false;
            return rc;
        }
        errword_clear(d->err);
        d->demTime = s.demTime; d->demTimeStep = s.demTimeStep; d->substeps = s.substeps;
    }
    return HYB_OK;
}

}  // extern "C"
