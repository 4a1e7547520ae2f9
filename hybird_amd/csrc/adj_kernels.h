// Host-side launch interface of the adjoint step (adj.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "lb_kernels.h"

namespace hyb {

// Number of per-workgroup slots of the tau-gradient partial sums.
size_t adj_tau_slots(const LBKParams& p);
// One reverse step: lamOut = lambda_t from lamNext = lambda_{t+1} and the
// recomputed post-collision state sPrev = s_{t-1} (pullFwd: t > 1, else sPrev = f0).
// Accumulates dL/dF into gradF [3][N] and dL/dtau into tauPart [adj_tau_slots].
void launch_adj_step(int collision, const double* sPrev, const double* lamNext, double* lamOut, const uint32_t* link,
                     const uint32_t* solid, const CoupleRec* cp, double* gradF, double* tauPart, bool pullFwd,
                     const LBKParams& p, hipStream_t s);

// Layout changes between [N][Q] (node-major, host views) and [Q][PS] (planes);
// planes_to_nodes writes zero for inactive nodes when link is given.
void launch_nodes_to_planes(const double* src, double* dst, size_t N, size_t PS, int Q, hipStream_t s);
void launch_planes_to_nodes(const double* src, double* dst, const uint32_t* link, size_t N, size_t PS, int Q,
                            hipStream_t s);

}  // namespace hyb
