// Host-side launch interface of the discrete-element kernels (dem.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"
#include "lb_kernels.h"

namespace hyb {

// Element state is stored field-major: field f of element e at el[f * Ecap + e].
// Offsets (in doubles) of each field; vectors take 3 slots, quaternions 4.
enum ElemField : int {
    EF_X = 0,      // x0..x5  (6 x 3)
    EF_XP = 18,    // xp0..xp5
    EF_Q = 36,     // q0..q5  (6 x 4)
    EF_QP = 60,    // qp0..qp5
    EF_W = 84,     // w0..w5  (6 x 3)
    EF_WP = 102,   // wp0..wp5
    EF_WG = 120,   // wGlobal
    EF_WL = 123,   // wLocal
    EF_WPG = 126,  // wpGlobal
    EF_WPL = 129,  // wpLocal
    EF_FH = 132, EF_MH = 135, EF_FP = 138, EF_MP = 141, EF_FW = 144, EF_MW = 147, EF_FG = 150,
    EF_M = 153, EF_RAD = 154, EF_I = 155, EF_FVOL = 158,
    EF_COUNT = 159
};

struct DEMParams {
    int P;                 // standard particles: the spheres of all elements (elements with size > 1 hold several)
    int E;                 // elements
    int clumps;            // some element has size > 1: particles live in their own arrays (DEMBuffers::pelem ...)
    int Gcap;              // ghost capacity
    int pairCap;           // neighbour-pair capacity
    int npbc, nwalls;
    int nobj;              // objects (originals and periodic copies), DEMBuffers::objs
    PbcDev pbc[3];
    WallDev walls[6];
    double cellWidth[3];
    unsigned nCells[3];
    double nebrRange, trigger, deltat;
    double c1[5], c2[5], co1[6], co2[6];
    double knConst, ksConst, dampCoeff, viscTang, fricPart, fricWall, linearStiff;
    int hertzian;
    double numVisc;
    double demF[3];
    int ncyl;              // cylinders (DEM.cpp:870-935)
    CylDev cyl[2];
    int spring;            // tangential spring with history (HYB_TANGENTIAL_SPRING) instead of the viscous law
    int allWalls;          // every wall within nebrRange is a contact partner (HYB_WALL_CONTACTS_ALL), not only the first
    int* herr;             // the handle's error word (mapped pinned host memory), nullptr: none
};

// Wall-contact records per particle: one per wall with DEMParams::allWalls, else one.
__host__ __device__ inline int wall_slots(const DEMParams& p) { return p.allWalls && p.nwalls > 0 ? p.nwalls : 1; }

// DEMState::error codes: 1 ghost capacity, 2 particle outside the cell grid,
// 3 stencil outside the cell table, 4 pair capacity, 5 more than NEAR_OBJ objects
// near one particle, 6 non-finite element velocity.
constexpr int DEM_ERR_NUMERIC = 6;

// Objects a particle can have in range at once (a rebuild that finds more
// reports DEMState::error 5).
constexpr int NEAR_OBJ = 16;

// Device-side mutable scalars of the DEM.
struct DEMState {
    double maxDisp;
    int nGhost;            // current ghost count
    int nSlots;            // P + nGhost
    int nPairs;
    int nNearWall;
    int newNeighborList;
    int error;             // nonzero: see the DEMState::error codes; every DEM kernel then does nothing
    int rebuilds;
};

// Device buffers of the neighbour machinery. Per-particle arrays are indexed by
// the standard particle (particles of one element are consecutive, in prototype
// order); ghost slots follow them.
struct DEMBuffers {
    double* el;            // [EF_COUNT][E]
    DEMState* st;
    uint32_t* ghostIndex;  // [Gcap]
    double* ghostVec;      // [Gcap][3]
    uint32_t* pairs;       // [pairCap][2]
    uint32_t* nearWall;    // [P] wall index or 0xFFFFFFFF (the first wall within range)
    uint32_t* nearWallMask;  // [P] bit w: wall w within range (DEMParams::allWalls only, else nullptr)
    int* cellTable;        // [Scap + ncells]
    int* cellOf;           // [Scap]
    int* cellCount;        // [ncells + 1]
    int* work;             // [max(W, Scap, pairCap) + 1] scan scratch
    uint32_t* csrOff;      // [P + 1]
    uint32_t* csr;         // [2 * pairCap] pair indices per particle, ascending
    int* cursor;           // [Scap] slots sorted by cell (descending within a cell)
    double* spos;          // [Scap][4] slot position and cluster, packed for the pair walk
    int* ghostOfCount;     // [P]
    int* ghostOf;          // [P][8] ghost slots of each particle, ascending
    int* side;             // [P][3] side-ghost index of each particle per periodic pair, -1 if none
    int* ghostB;           // [Gcap] periodic pair that created a side ghost
    uint32_t* posI;        // [pairCap] csr slot of a pair in its first particle's list (0xFFFFFFFF: ghost)
    uint32_t* posJ;        // [pairCap] csr slot in its second particle's list
    struct PairRec* prec;  // [2 * pairCap] per csr slot: signed force/torque increments of one pair
    double* vmax;          // [1] max |x1|^2 over the elements after the corrector, kept as the bit pattern of a
                           // non-negative double (integer atomic max); the displacement check reads and clears it
    double* energy;        // [6] particleEnergy: kin, rotKin, trKin, grav, elastic, total
    const ObjDev* objs;    // [nobj]
    uint32_t* nearObj;     // [P][NEAR_OBJ] objects within nebrRange of each particle, ascending
    int* nearObjCount;     // [P]
    uint32_t* nearCyl;     // [P] first cylinder within nebrRange or 0xFFFFFFFF
    // forces of the last sub-step's wall and object contacts (for walls[].FParticle
    // and objects[].FParticle, summed on request in the reference's order)
    struct ContactForce* wcf;  // [P][WS] contact of each particle with its near wall(s); WS = allWalls ? nwalls : 1
    struct ContactForce* ocf;  // [P][NEAR_OBJ] contacts with its near objects
    // contact statistics of a saving sub-step (nullptr on the others)
    struct StatRec* pstat;  // [pairCap] per pair
    struct StatRec* wstat;  // [P][WS] wall contacts of each particle
    struct StatRec* ostat;  // [P][NEAR_OBJ] object contacts of each element
    // tangential springs (spring model only, else nullptr): per pair xi of the pair's second particle
    // relative to its first [pairCap][3]; per particle and wall slot (wall_slots) xi against the wall and that
    // wall's index, -1 without contact [P][WS][4];
    // copies of the pair list, its springs, the per-particle lists and the ghosts taken at the start of a
    // rebuild, from which the new pairs take over their springs
    double* hist;
    double* whist;
    double* histOld;
    uint32_t* pairsOld;
    uint32_t* csrOffOld;
    uint32_t* csrOld;
    uint32_t* ghostIndexOld;
    // multi-sphere elements (DEMParams::clumps only, else nullptr; with single spheres the particle is its
    // element): element of each particle, first particle of each element, the particle's prototype vector
    // (DEM.cpp:404-433), particle state after the corrector (position, velocity, radiusVec, elmt.cpp:277-290)
    // and after the predictor (elmt.cpp:261-275), and per csr slot the moment of the normal force that
    // elements of size > 1 take (DEM.cpp:1862-1867)
    int* pelem;            // [P]
    int* efirst;           // [E + 1]
    double* pproto;        // [P][3]
    double* pcor;          // [P][9]
    double* ppred;         // [P][9]
    double* pmn;           // [2 * pairCap][3]
    int scap;
    int ncells;
    int W;                 // ncells * 14
};

// Increments one contact adds to one element: the normal force (applied when
// flag & 1), the normal-force moment of an element of size > 1 (flag & 4, held
// in DEMBuffers::pmn), then torque and tangential force (when flag & 2).
struct PairRec {
    double fn[3];
    double ft[3];
    double m[3];
    int flag;
    int pad;
};

// What one wall or object contact did to the wall/object: FParticle -= fn, then,
// with tangential motion, FParticle += ft (DEM.cpp:1936-1979, 2100-2134). flag
// bit 0: in contact; bit 1: tangential force applied.
struct ContactForce {
    double fn[3], ft[3];
    double flag, pad;
};

// One contact at a saving sub-step, with the fields of the MercuryDPM .fstat
// line (DEM.cpp:2479-2565): the partners (a, b: the pair's two particle slots;
// wall index and element; object slot (periodic copies included) and element), overlap, normal and
// tangential force magnitudes, normal and tangential unit vectors. flag = 1
// when the contact exists and has tangential motion (the reference writes only
// those); 0 otherwise.
struct StatRec {
    double a, b, ov, fn, ft;
    double en[3], et[3];
    double flag;
};

// The particle state a saving sub-step exports (after the predictor): per
// particle the predicted position and velocity and its element's wGlobal;
// info[0] = pairs.
void launch_dem_save_elements(const DEMParams& p, const DEMBuffers& b, double* out, double* info, hipStream_t s);
// force = 1: rebuild unconditionally (initialisation); 0: displacement check
// (DEM.cpp:1314-1324, 342-347), rebuilding when it fires.
void launch_dem_rebuild(const DEMParams& p, const DEMBuffers& b, int force, hipStream_t s);
// A sub-step's contacts and corrector: per-pair contact records, then the
// per-element gather (corrector, next predictor).
void launch_dem_pairs(const DEMParams& p, const DEMBuffers& b, hipStream_t s);
void launch_dem_gather(const DEMParams& p, const DEMBuffers& b, hipStream_t s);
void launch_dem_predict(const DEMParams& p, const DEMBuffers& b, hipStream_t s);
// Particle positions, velocities and radius vectors of every particle from its
// element's corrected and predicted state (multi-sphere elements only).
void launch_dem_particles(const DEMParams& p, const DEMBuffers& b, hipStream_t s);
void launch_dem_energy(const DEMParams& p, const DEMBuffers& b, const double demSize[3], double* out, hipStream_t s);
void launch_dem_couple_table(const DEMParams& p, const DEMBuffers& b, CoupleRec* cp, CoupleRecF* cpf, int* nslots,
                             const Units& u, hipStream_t s);
// FHydro, MHydro and fluidVolume of every element from the run sums of its own
// and its ghosts' slots (hs: filled by the fused collision, emptied here).
void launch_dem_hydro_combine(const DEMParams& p, const DEMBuffers& b, const HydroSink& hs, const Units& u,
                              hipStream_t s);

}  // namespace hyb
