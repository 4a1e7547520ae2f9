// Host-side launch interface of the free-surface pass (fs.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"
#include "lb_kernels.h"

namespace hyb {

// Device counters of the free-surface pass.
struct FSCounters {
    uint32_t n[2];              // interface list lengths (current / next, ping-pong)
    uint32_t nA, nB;            // gas -> interface and fluid -> interface lists of this step
    unsigned long long acc[4];  // exact fixed-point mass-surplus accumulator (32-bit limbs, LSB 2^-96)
    unsigned long long extra[4];  // extra mass of the curved-wall rule in the last streaming (same format)
    unsigned long long tot[4];    // mass of the active nodes outside particles (DRUM mass fix, same format)
};

// Device state of the free-surface mode. Densities use the population type
// (double for FP64, float for the FP32 form); masses and velocities are FP64.
struct FSBuffers {
    double* mass = nullptr;                // [N] interface-node mass (fluid-node mass is its last density)
    void* rho[2] = {nullptr, nullptr};     // [N] density of the last two collisions
    double* uS = nullptr;                  // [3][N] shifted velocity of interface nodes (last collision)
    double* hFi = nullptr;                 // [3][N] node force of interface nodes (last collision)
    unsigned long long* claim = nullptr;   // [N] (step tag << 32) | value, see fs.hip
    uint32_t* list[2] = {nullptr, nullptr};
    uint32_t* listA = nullptr;
    uint32_t* listB = nullptr;
    double* mnew = nullptr;                // [N] new mass per interface-list position
    uint8_t* state = nullptr;              // [N] filled / emptied / unchanged per interface-list position
    FSCounters* ctr = nullptr;
    // [N] (step tag << 2) | class (0 inactive, 1 fluid, 2 interface), written for
    // every node whose type the pass may change (only allocated with slip walls or
    // cylinders): the slip rule of the next collision's streaming tests the
    // neighbours' activity before the pass, the curved-wall rule whether the node
    // behind was fluid
    uint32_t* actPrev = nullptr;
};

// What the fused collision kernel reads and writes in free-surface mode.
struct FSView {
    const void* pre;       // output population buffer: populations prepared by the free-surface pass
    const void* rhoPrev;   // density of the previous collision = mass of fluid nodes
    void* rhoCur;          // density of this collision
    const double* mass;    // interface-node mass
    double* uS;            // shifted velocity of interface nodes
    double* hFi;           // node force of interface nodes
    uint32_t* linkw;       // link words: the prepared flag is cleared when consumed
};

// Builds the interface list of the initial lattice.
// Before a free-surface pass with cylinders: the curved-wall rule's extra mass of
// the last streaming, spread over the interface list `cur` (fpost: post-collision
// state; rhoFluid: fluid-node masses, or initDensity when !massKnown).
void launch_fs_curved_extra(Prec prec, const void* fpost, const uint32_t* link, FSBuffers& b, int cur,
                            const void* rhoFluid, bool massKnown, const LBKParams& p, hipStream_t s);
// DRUM's enforceMassConservation (LB.cpp:1806-1822) after a pass started with
// interface list `cur`: the mass of the active nodes outside particles (solid:
// ownership words of coupling step ctag, nullptr without particles) against
// totalMass; the interface nodes of the new list gain -0.01 (mass - totalMass) / count.
void launch_fs_mass_fix(Prec prec, const uint8_t* type, const uint32_t* solid, uint32_t ctag, FSBuffers& b, int cur,
                        int rcur, double totalMass, const LBKParams& p, hipStream_t s);
void launch_fs_build_list(const uint8_t* type, FSBuffers& b, int cur, const LBKParams& p, hipStream_t s);
// One free-surface step (updateMass + updateInterface, LB.cpp:1492-1804).
// cur: which interface list is current; rcur: which density buffer holds the
// last collision's density; Fnew: body force used to initialise new interface
// nodes; pull: the populations are post-collision (false before the first collision).
void launch_fs_step(Prec prec, const void* fpost, void* fout, uint8_t* type, uint32_t* link, uint32_t* slipw,
                    double* visc, FSBuffers& b, int cur, int rcur, uint32_t tag, const double* Fnew, bool pull,
                    const LBKParams& p, hipStream_t s);
// Post-stream populations for the host view in free-surface mode (gas links
// reconstructed from the stored interface velocity, LB.cpp:1237-1246).
void launch_fs_stream_view(Prec prec, const void* fin, void* fout, const uint8_t* type, const uint32_t* link,
                           const double* uS, const LBKParams& p, hipStream_t s);

}  // namespace hyb
