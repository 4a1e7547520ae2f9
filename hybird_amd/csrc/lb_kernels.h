// Host-side launch interface of the lattice-Boltzmann kernels (lb.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"

namespace hyb {

// Parameters every LB kernel receives by value.
struct LBKParams {
    int X, Y, Z;
    size_t PS;           // population plane stride (elements), N padded against channel aliasing
    int per[3];          // periodic along x, y, z
    double F[3];         // body force (lattice units), zero when forceFieldSolve = 0
    double visc0;        // initial kinematic viscosity (lattice units)
    double plasticVisc, yieldStress, turbConst;
    int nonNewtonian, turbulence;
    int frictional;       // frictional apparent viscosity instead of Bingham (yield stress = frictionFluid x gauge pressure)
    double frictionFluid;
    int collision;       // HYB_COLLISION_BGK / HYB_COLLISION_TRT
    double trtMagic;
    double unitLength, unitSpeed, unitAngVel;
    double initDensity;
    uint32_t ctag;       // tag of the current coupling step in the node ownership words
    const float* ffield; // optional per-node body force [3][N] added to F (nullptr: none)
    // slip walls: S1 = slipCoefficient, S2 = 1 - S1 (LB.cpp:1154-1155), per-node slip
    // directions (bit j) and the link words (neighbour activity)
    double slipS1, slipS2;
    const uint32_t* slipw;
    const uint32_t* linkw;
    const uint32_t* actPrev;  // free surface with slip walls: activity before the last pass, see FSBuffers
    uint32_t actTag;          // tag of that pass; 0 = no pass since the last collision
    // curved walls of a cylinder (LB.cpp:1278-1319): per-node curved-link
    // directions (bit j), density and Guo-shifted velocity of the previous collision
    // (macPrev, [4][N]: n, ux, uy, uz; read by the pull) and of this one (macNext,
    // written for every active node), node types, per-node viscosity (nullptr: visc0)
    const uint32_t* curvw;
    const double* macPrev;
    double* macNext;
    const uint8_t* typew;
    const double* viscw;
    CylDev cyl;
    // box sides (bit 2 * axis + high) whose boundary layer is a plane wall: a curved
    // wall node there takes that wall's (zero) velocity, every other one the cylinder's
    int wallSides;
    // the handle's error word (mapped pinned host memory; nullptr: none): a
    // collision that finds a non-finite density stores LB_ERR_NUMERIC there
    int* herr;
};

constexpr int LB_ERR_NUMERIC = 6;

constexpr int SOLID_SLOT_BITS = 20;
constexpr uint32_t SOLID_SLOT_MASK = (1u << SOLID_SLOT_BITS) - 1u;
constexpr uint32_t SOLID_TAG_MAX = (1u << (32 - SOLID_SLOT_BITS)) - 1u;

// Per-particle record the coupling reads (particle slots include ghosts).
// Raw (DEM, physical-unit) values: centre, vector centre-of-element -> centre,
// element velocity and angular velocity; radius.
struct CoupleRec {
    double x0[3];
    double rv[3];
    double v[3];
    double w[3];
    double r;
    int32_t elem;
    int32_t pad;
};

// The same record for the FP32 path, prescaled to lattice units:
// c = x0/L - radiusVec/L (so radius = node - c), v/Speed, w/AngVel.
struct CoupleRecF {
    float c[3];
    float v[3];
    float w[3];
    float pad[3];
};

// Box of lattice nodes a particle slot may cover (inclusive bounds).
struct NodeBox {
    int lo[3];
    int hi[3];
};

// Hydrodynamic sums of one particle slot over one run of its nodes: a
// wavefront's consecutive x nodes that the slot owns (LB.cpp:1897-1902):
// diffVel (3), radius x diffVel (3), node mass (fluidVolume), added in x order.
struct HydroSeg {
    double s[7];
    uint32_t slot;
    uint32_t pad;
};

// Where the collision appends the run sums: slot k's runs go to
// segs[k * cap + i] for i = atomicAdd(&count[k], 1) < cap. The consumer (the
// per-element combine) zeroes count[k] after reading. segs == nullptr: no sums.
struct HydroSink {
    HydroSeg* segs;
    uint32_t* count;
    uint32_t cap;
};

// LB error-word code of a run that found its slot's segment list full.
constexpr int LB_ERR_SEGMENTS = 4;

enum class Prec { F64, F32S };

struct FSView;
// f planes: FP64 -> double[19][PS]; FP32S -> float[19][PS] holding f - w (PS >= N).
// fs: free-surface state (nullptr without free surface), see fs_kernels.h
// hs: per-slot hydrodynamic run sums of the particle nodes (nullptr: none)
void launch_lb_fused(Prec prec, bool pull, bool varVisc, int collision, const void* fin, void* fout,
                     const uint32_t* link, const uint32_t* solid, const CoupleRec* cp, const CoupleRecF* cpf,
                     const HydroSink* hs, double* visc, double* mac, const FSView* fs, const LBKParams& p,
                     hipStream_t s);
void launch_lb_stream(Prec prec, const void* fin, void* fout, const uint32_t* link, const LBKParams& p, hipStream_t s);
// out[3] (zeroed by the caller): max |u|^2 bits from mac (may be nullptr), active
// nodes, active nodes with visc > viscMin (visc may be nullptr).
void launch_status(const double* mac, const double* visc, const uint32_t* link, size_t N, double viscMin,
                   unsigned long long* out, hipStream_t s);
// DEM plane walls (positions in physical units) for the lattice-side wall force.
struct WallSet {
    int n;
    double unitLength;
    WallDev w[6];
};
// Exact sums of the hydrodynamic force on moving no-slip walls from the
// post-collision state f: acc[nwalls][3][4] (zeroed by the caller).
void launch_wall_fhydro(Prec prec, const void* f, const uint8_t* type, const WallSet& ws, unsigned long long* acc,
                        const LBKParams& p, hipStream_t s);
// out[k] = (value of the exact sum acc[4k..4k+3]) * scale for k < n, on the device
// (acc == nullptr writes zeros).
void launch_exact_finalize(const unsigned long long* acc, int n, double scale, double* out, hipStream_t s);
// Cylinders of the lattice (nodes outside become curved walls).
struct CylSet {
    int n;
    CylDev c[2];
};
// region: liquid box (6 values), then avalanchePosit (NET)
void launch_lb_init_types(uint8_t* type, const int* boundary, int nwalls, const WallDev* walls, int nobj,
                          const ObjDev* objs, const CylSet& cs, int problem, const double* region, const LBKParams& p,
                          hipStream_t s);
// tmp: N bytes of device scratch (its contents are overwritten)
void launch_lb_init_interface(uint8_t* type, uint8_t* tmp, const LBKParams& p, hipStream_t s);
// slipw: per-node slip-wall directions (nullptr when no slip walls exist)
// curvw: per-node curved-link directions (nullptr without cylinders)
void launch_lb_init_links(const uint8_t* type, uint32_t* link, uint32_t* slipw, uint32_t* curvw, const LBKParams& p,
                          hipStream_t s);
void launch_lb_max_active(const uint32_t* link, int* maxXYZ, const LBKParams& p, hipStream_t s);
// With free surface, also the initial density (rho), interface mass and interface velocity.
void launch_lb_init_pops(Prec prec, void* f, const uint32_t* link, const int* maxXYZ, double* visc, const double* u0,
                         const double* Finit, void* rho, double* mass, double* uS, const LBKParams& p, hipStream_t s);
// link == nullptr: mark every node inside a sphere, active or not (for lattices whose
// inactive nodes stay inactive and are never read as particle nodes)
void launch_classify_set(const int* nslots, int max_slots, const CoupleRec* cp, NodeBox* boxes, int* nboxes,
                         uint32_t* solid, const uint32_t* link, const LBKParams& p, hipStream_t s);
// Non-temporal device copy of `bytes` with 4-, 8- or 16-byte accesses per lane
// (false: unsupported width).
bool launch_copy_probe(int width, const void* src, void* dst, size_t bytes, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void launch_count_active(const uint32_t* link, unsigned long long* count, const LBKParams& p, hipStream_t s);
// count[0], count[1] += active fluid and interface nodes outside particles (solid: ownership words or nullptr).
void launch_count_mass(const uint32_t* link, const uint32_t* solid, unsigned long long* count, const LBKParams& p,
                       hipStream_t s);

}  // namespace hyb
