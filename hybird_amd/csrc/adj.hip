// Discrete adjoint of the lattice-Boltzmann step (FP64) for gfx950.
//
// Forward model (particles frozen, no free surface): g_1 = f0, and for t >= 1
//   s_t = C(g_t; F, tau)      collision of every active node (the fused kernel)
//   g_{t+1} = P(s_t)          pull streaming with half-way bounce-back
// where C contains the penalty force of the frozen particles (LB.cpp:1851-1919),
// the Guo velocity shift, the equilibrium and BGK or TRT relaxation with Guo
// forcing (node.cpp:85-184), and F is the uniform body force plus an optional
// per-node force field. For a loss L(g_{T+1}) the adjoint runs backwards:
//   mu_t = P^T lambda_{t+1}   (P^T pulls plane j from the neighbour x + c_j;
//                              at a static wall it reads the node's own plane opp(j))
//   lambda_t = (dC/dg)^T mu_t,  dL/dF(x) += (dC/dF)^T mu_t,  dL/dtau += (dC/dtau)^T mu_t
// and dL/df0 = lambda_1. One kernel per reverse step recomputes C's
// intermediates from g_t (pulled from the recomputed s_{t-1}) and applies the
// hand-derived vector-Jacobian product; the gradient of tau is reduced per
// workgroup into its own slot (deterministic).
#include <hip/hip_runtime.h>

#include "adj_kernels.h"
#include "common.h"
#include "hybird_amd.h"
#include "lattice.h"
#include "lb_kernels.h"

namespace hyb {

namespace {

constexpr int ABX = 256;

template <int COLL>
__global__ void __launch_bounds__(ABX)
    k_adj_step(const double* __restrict__ sPrev, const double* __restrict__ lamNext, double* __restrict__ lamOut,
               const uint32_t* __restrict__ link, const uint32_t* __restrict__ solid, const CoupleRec* __restrict__ cp,
               double* __restrict__ gradF, double* __restrict__ tauPart, LBKParams p, int pullFwd) {
#pragma clang fp contract(off)
    __shared__ double red[ABX / 64];
    const int x = blockIdx.x * ABX + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    const size_t N = (size_t)p.X * p.Y * p.Z;
    double tauAcc = 0.0;
    const bool inRange = x < p.X;
    const size_t self = inRange ? (size_t)x + (size_t)p.X * ((size_t)y + (size_t)p.Y * (size_t)z) : 0;
    const uint32_t lw = inRange ? link[self] : 0u;
    if (lw & LINK_ACTIVE) {
        Coord c;
        c.x = x; c.y = y; c.z = z; c.idx = self;
        size_t nb[19];
#pragma unroll
        for (int j = 1; j < 19; ++j) nb[j] = nb_index(p, c, j);
        // g_t: populations entering this node's collision
        double g[19], mu[19];
        g[0] = sPrev[self];
#pragma unroll
        for (int i = 1; i < 19; ++i) {
            const int j = opp(i);  // f_i arrives from the neighbour in direction j = opp(i)
            if (!pullFwd) g[i] = sPrev[(size_t)i * p.PS + self];
            else if (lw & (1u << j)) g[i] = sPrev[(size_t)j * p.PS + self];
            else g[i] = sPrev[(size_t)i * p.PS + nb[j]];
        }
        // mu_t = P^T lambda_{t+1}
        mu[0] = lamNext[self];
#pragma unroll
        for (int j = 1; j < 19; ++j) {
            if (lw & (1u << j)) mu[j] = lamNext[(size_t)opp(j) * p.PS + self];
            else mu[j] = lamNext[(size_t)j * p.PS + nb[j]];
        }
        // ---- forward intermediates of C (same formulas as the fused kernel)
        double n = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            n += g[i];
            jx += (double)cx(i) * g[i];
            jy += (double)cy(i) * g[i];
            jz += (double)cz(i) * g[i];
        }
        const V3 u(jx / n, jy / n, jz / n);
        V3 hF, dv;
        double pm = 0.0;
        const uint32_t sid = solid ? owner_slot(solid[self], p.ctag) : 0u;
        if (sid) {
            const CoupleRec& r = cp[sid - 1];
            const V3 pos((double)x, (double)y, (double)z);
            const V3 radius = pos - V3(r.x0[0], r.x0[1], r.x0[2]) / p.unitLength + V3(r.rv[0], r.rv[1], r.rv[2]) / p.unitLength;
            const V3 lv = V3(r.v[0], r.v[1], r.v[2]) / p.unitSpeed +
                          (V3(r.w[0], r.w[1], r.w[2]).cross(radius)) / p.unitAngVel;
            pm = (lw & LINK_INTERFACE) ? 0.5 * p.initDensity : p.initDensity;
            dv = u - lv;
            hF = (-pm / n) * dv;
        }
        V3 Fb(p.F[0], p.F[1], p.F[2]);
        if (p.ffield) Fb = Fb + V3((double)p.ffield[self], (double)p.ffield[N + self], (double)p.ffield[2 * N + self]);
        const V3 T = Fb + hF;
        const V3 us = u + T / (2.0 * n);
        const double omP = 1.0 / (0.5 + 3.0 * p.visc0);
        const double tau = 1.0 / omP;
        const double omM = COLL == HYB_COLLISION_TRT ? 1.0 / (0.5 + p.trtMagic / (tau - 0.5)) : omP;
        const double q = us.norm2(), rr = us.dot(T);
        double a[19], b[19], feq[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            a[i] = (double)cx(i) * us.x + (double)cy(i) * us.y + (double)cz(i) * us.z;
            b[i] = (double)cx(i) * T.x + (double)cy(i) * T.y + (double)cz(i) * T.z;
            feq[i] = wgt(i) * n * (1.0 + 3.0 * a[i] + 4.5 * a[i] * a[i] - 1.5 * q);
        }
        // ---- reverse
        double gb[19];
        double nbar = 0.0, qbar = 0.0, rbar = 0.0, omPbar = 0.0, omMbar = 0.0;
        V3 usbar, Tbar;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            const int o = opp(i);
            const double mi = mu[i], mo = mu[o];
            gb[i] = mi * (1.0 - 0.5 * omP - 0.5 * omM) + mo * (0.5 * omM - 0.5 * omP);
            const double feqb = mi * (0.5 * omP + 0.5 * omM) + mo * (0.5 * omP - 0.5 * omM);
            const double spb = mi * (1.0 - 0.5 * omP), smb = mi * (1.0 - 0.5 * omM);
            const double sp = wgt(i) * (9.0 * a[i] * b[i] - 3.0 * rr), sm = wgt(i) * 3.0 * b[i];
            omPbar += mi * (-(0.5 * (g[i] + g[o]) - 0.5 * (feq[i] + feq[o])) - 0.5 * sp);
            omMbar += mi * (-(0.5 * (g[i] - g[o]) - 0.5 * (feq[i] - feq[o])) - 0.5 * sm);
            nbar += feqb * wgt(i) * (1.0 + 3.0 * a[i] + 4.5 * a[i] * a[i] - 1.5 * q);
            qbar += feqb * wgt(i) * n * -1.5;
            const double ab = feqb * wgt(i) * n * (3.0 + 9.0 * a[i]) + spb * wgt(i) * 9.0 * b[i];
            const double bb = spb * wgt(i) * 9.0 * a[i] + smb * wgt(i) * 3.0;
            rbar += spb * wgt(i) * -3.0;
            usbar += V3((double)cx(i) * ab, (double)cy(i) * ab, (double)cz(i) * ab);
            Tbar += V3((double)cx(i) * bb, (double)cy(i) * bb, (double)cz(i) * bb);
        }
        usbar += (2.0 * qbar) * us + rbar * T;
        Tbar += rbar * us;
        // tau: omP = 1/tau, omM = 1/(1/2 + Lambda/(tau - 1/2)) (or omP for BGK)
        if (COLL == HYB_COLLISION_TRT) {
            const double d = tau - 0.5;
            tauAcc = omPbar * (-omP * omP) + omMbar * (omM * omM * p.trtMagic / (d * d));
        } else {
            tauAcc = (omPbar + omMbar) * (-omP * omP);
        }
        // us = u + T / (2 n)
        V3 ubar = usbar;
        Tbar += usbar / (2.0 * n);
        nbar += -usbar.dot(T) / (2.0 * n * n);
        // T = F + hF
        gradF[self] += Tbar.x;
        gradF[N + self] += Tbar.y;
        gradF[2 * N + self] += Tbar.z;
        if (sid) {  // hF = -(m / n)(u - u_p)
            ubar += (-pm / n) * Tbar;
            nbar += (pm / (n * n)) * Tbar.dot(dv);
        }
        // u = j / n
        const V3 jbar = ubar / n;
        nbar += -ubar.dot(u) / n;
#pragma unroll
        for (int i = 0; i < 19; ++i)
            lamOut[(size_t)i * p.PS + self] =
                gb[i] + nbar + (double)cx(i) * jbar.x + (double)cy(i) * jbar.y + (double)cz(i) * jbar.z;
    }
    // workgroup sum of the tau gradient in a fixed order
    for (int off = 32; off > 0; off >>= 1) tauAcc += __shfl_xor(tauAcc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = tauAcc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < ABX / 64; ++w) s += red[w];
        const size_t slot = (size_t)blockIdx.x + (size_t)gridDim.x * ((size_t)blockIdx.y + (size_t)gridDim.y * blockIdx.z);
        tauPart[slot] += s;
    }
}

}  // namespace

size_t adj_tau_slots(const LBKParams& p) { return (size_t)((p.X + ABX - 1) / ABX) * p.Y * p.Z; }

void launch_adj_step(int collision, const double* sPrev, const double* lamNext, double* lamOut, const uint32_t* link,
                     const uint32_t* solid, const CoupleRec* cp, double* gradF, double* tauPart, bool pullFwd,
                     const LBKParams& p, hipStream_t s) {
    const dim3 g((p.X + ABX - 1) / ABX, p.Y, p.Z), b(ABX, 1, 1);
    if (collision == HYB_COLLISION_TRT)
        k_adj_step<HYB_COLLISION_TRT><<<g, b, 0, s>>>(sPrev, lamNext, lamOut, link, solid, cp, gradF, tauPart, p,
                                                      pullFwd ? 1 : 0);
    else
        k_adj_step<HYB_COLLISION_BGK><<<g, b, 0, s>>>(sPrev, lamNext, lamOut, link, solid, cp, gradF, tauPart, p,
                                                      pullFwd ? 1 : 0);
}

}  // namespace hyb
