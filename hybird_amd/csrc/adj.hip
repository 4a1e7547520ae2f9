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
// at least this many workgroups per CU: caps the registers so that several
// wavefronts per SIMD hide the latency of the 38 neighbour loads
#ifndef ADJ_MIN_BLOCKS
#define ADJ_MIN_BLOCKS 2
#endif
// issue the adjoint populations' loads together with the state's (more bytes in
// flight per wavefront) instead of after the forward intermediates: measured on
// the 128^3 frozen bed, backward/forward 2.80 against 2.93
#ifndef ADJ_HOIST
#define ADJ_HOIST 1
#endif
// non-temporal loads of the state and adjoint populations and stores of the
// new adjoint populations (each is touched once per step)
#ifndef ADJ_NT
#define ADJ_NT 1
#endif

__device__ __forceinline__ double ld_nt(const double* p) {
#if ADJ_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
__device__ __forceinline__ void st_nt(double* p, double v) {
#if ADJ_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// Neighbour index along direction j (periodic rewiring) as a 32-bit offset.
__device__ __forceinline__ uint32_t nb32(const Nb3& nx, const Nb3& ny, const Nb3& nz, uint32_t sX, uint32_t sXY, int j) {
    return (uint32_t)pick(nx, cx(j)) + sX * (uint32_t)pick(ny, cy(j)) + sXY * (uint32_t)pick(nz, cz(j));
}

// The vector-Jacobian product runs over the 9 direction pairs (i, opp i = i + 1)
// and the rest direction, using c_opp = -c_i and w_opp = w_i, so that only the
// incoming populations and the adjoint populations stay in registers.
template <int COLL>
__global__ void __launch_bounds__(ABX, ADJ_MIN_BLOCKS)
    k_adj_step(const double* __restrict__ sPrev, const double* __restrict__ lamNext, double* __restrict__ lamOut,
               const uint32_t* __restrict__ link, const uint32_t* __restrict__ solid, const CoupleRec* __restrict__ cp,
               double* __restrict__ gradF, double* __restrict__ tauPart, LBKParams p, int pullFwd) {
#pragma clang fp contract(off)
    __shared__ double red[ABX / 64];
    const int x = blockIdx.x * ABX + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    const size_t N = (size_t)p.X * p.Y * p.Z;
    double tauAcc = 0.0;
    const bool inRange = x < p.X;
    const uint32_t self = inRange ? (uint32_t)x + (uint32_t)p.X * ((uint32_t)y + (uint32_t)p.Y * (uint32_t)z) : 0u;
    const uint32_t lw = inRange ? link[self] : 0u;
    if (lw & LINK_ACTIVE) {
        const Nb3 nx = nb3(x, p.X, p.per[0]), ny = nb3(y, p.Y, p.per[1]), nz = nb3(z, p.Z, p.per[2]);
        const uint32_t sX = (uint32_t)p.X, sXY = (uint32_t)p.X * (uint32_t)p.Y;
        // g_t: populations entering this node's collision; f_i arrives from the
        // neighbour in direction opp(i) (or bounces back from a static wall there)
        double g[19], m[19];
        if (pullFwd) {
            // the 19 loads back to back; a static-wall neighbour's index is valid
            // memory and its value is replaced by the bounce-back below
#pragma unroll
            for (int i = 0; i < 19; ++i)
                g[i] = ld_nt(&sPrev[(size_t)i * p.PS + (i == 0 ? self : nb32(nx, ny, nz, sX, sXY, opp(i)))]);
            if (lw & LINK_WALLS) {
#pragma unroll
                for (int i = 1; i < 19; ++i)
                    if (lw & (1u << opp(i))) g[i] = sPrev[(size_t)opp(i) * p.PS + self];
            }
        } else {
#pragma unroll
            for (int i = 0; i < 19; ++i) g[i] = sPrev[(size_t)i * p.PS + self];
        }
        // mu_t = P^T lambda_{t+1}: plane j from the neighbour in direction j
        const auto load_mu = [&]() {
#pragma unroll
            for (int j = 0; j < 19; ++j)
                m[j] = ld_nt(&lamNext[(size_t)j * p.PS + (j == 0 ? self : nb32(nx, ny, nz, sX, sXY, j))]);
            if (lw & LINK_WALLS) {
#pragma unroll
                for (int j = 1; j < 19; ++j)
                    if (lw & (1u << j)) m[j] = lamNext[(size_t)opp(j) * p.PS + self];
            }
        };
        if (ADJ_HOIST) load_mu();
        // the force gradient accumulated so far, read here so that its load overlaps the populations' loads
        const double gF0 = gradF[self], gF1 = gradF[N + self], gF2 = gradF[2 * N + self];
        // ---- forward intermediates of C (the fused kernel's formulas)
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
        const double hp = 0.5 * omP, hm = 0.5 * omM;
        // d(s)/d(omP) and d(s)/d(omM) per direction pair (even / odd parts); g is
        // no longer needed after this, which keeps the adjoint loads below in registers
        double ev[10], od[10];
        ev[0] = -(g[0] - wgt(0) * n * (1.0 - 1.5 * q)) - 0.5 * (wgt(0) * -3.0 * rr);
        od[0] = 0.0;
#pragma unroll
        for (int i = 1; i < 19; i += 2) {
            const int o = i + 1;
            const double w = wgt(i);
            const double a = (double)cx(i) * us.x + (double)cy(i) * us.y + (double)cz(i) * us.z;
            const double b = (double)cx(i) * T.x + (double)cy(i) * T.y + (double)cz(i) * T.z;
            const double base = 1.0 + 4.5 * a * a - 1.5 * q;
            const double fei = w * n * (base + 3.0 * a), feo = w * n * (base - 3.0 * a);
            ev[(i + 1) / 2] = -(0.5 * (g[i] + g[o]) - 0.5 * (fei + feo)) - 0.5 * (w * (9.0 * a * b - 3.0 * rr));
            od[(i + 1) / 2] = -(0.5 * (g[i] - g[o]) - 0.5 * (fei - feo)) - 0.5 * (w * 3.0 * b);
        }
        // Opaque copies of the velocity and force: the reverse loop recomputes c.us
        // and c.T per pair instead of keeping the values of the loop above alive.
        double usx = us.x, usy = us.y, usz = us.z, Tx = T.x, Ty = T.y, Tz = T.z;
        asm volatile("" : "+v"(usx), "+v"(usy), "+v"(usz), "+v"(Tx), "+v"(Ty), "+v"(Tz));
        if (!ADJ_HOIST) load_mu();
        // ---- reverse
        double nbar = 0.0, qbar = 0.0, rbar = 0.0, omPbar = 0.0, omMbar = 0.0;
        V3 usbar, Tbar;
        {  // rest direction: c = 0, so a = b = 0, no odd part
            const double w = wgt(0), mi = m[0];
            const double feqb = mi * omP, spb = mi * (1.0 - hp);
            omPbar += mi * ev[0];
            nbar += feqb * w * (1.0 - 1.5 * q);
            qbar += feqb * w * n * -1.5;
            rbar += spb * w * -3.0;
            m[0] = mi * (1.0 - omP);  // adjoint of g_0 through the relaxation
        }
#pragma unroll
        for (int i = 1; i < 19; i += 2) {
            const int o = i + 1;  // opp(i)
            const double w = wgt(i);
            const double mi = m[i], mo = m[o];
            const double a = (double)cx(i) * usx + (double)cy(i) * usy + (double)cz(i) * usz;
            const double b = (double)cx(i) * Tx + (double)cy(i) * Ty + (double)cz(i) * Tz;
            const double base = 1.0 + 4.5 * a * a - 1.5 * q;
            const double feqbi = mi * (hp + hm) + mo * (hp - hm), feqbo = mo * (hp + hm) + mi * (hp - hm);
            const double spbi = mi * (1.0 - hp), spbo = mo * (1.0 - hp);
            const double smbi = mi * (1.0 - hm), smbo = mo * (1.0 - hm);
            omPbar += (mi + mo) * ev[(i + 1) / 2];
            omMbar += (mi - mo) * od[(i + 1) / 2];
            nbar += feqbi * w * (base + 3.0 * a) + feqbo * w * (base - 3.0 * a);
            qbar += (feqbi + feqbo) * w * n * -1.5;
            rbar += (spbi + spbo) * w * -3.0;
            const double abi = feqbi * w * n * (3.0 + 9.0 * a) + spbi * w * 9.0 * b;
            const double abo = feqbo * w * n * (3.0 - 9.0 * a) - spbo * w * 9.0 * b;  // a_o = -a, b_o = -b
            const double bbi = spbi * w * 9.0 * a + smbi * w * 3.0;
            const double bbo = -spbo * w * 9.0 * a + smbo * w * 3.0;
            const double da = abi - abo, db = bbi - bbo;
            usbar += V3((double)cx(i) * da, (double)cy(i) * da, (double)cz(i) * da);
            Tbar += V3((double)cx(i) * db, (double)cy(i) * db, (double)cz(i) * db);
            // adjoint of g through the relaxation (the moments' part is added below)
            m[i] = mi * (1.0 - hp - hm) + mo * (hm - hp);
            m[o] = mo * (1.0 - hp - hm) + mi * (hm - hp);
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
        gradF[self] = gF0 + Tbar.x;
        gradF[N + self] = gF1 + Tbar.y;
        gradF[2 * N + self] = gF2 + Tbar.z;
        if (sid) {  // hF = -(m / n)(u - u_p)
            ubar += (-pm / n) * Tbar;
            nbar += (pm / (n * n)) * Tbar.dot(dv);
        }
        // u = j / n
        const V3 jbar = ubar / n;
        nbar += -ubar.dot(u) / n;
#pragma unroll
        for (int i = 0; i < 19; ++i)
            st_nt(&lamOut[(size_t)i * p.PS + self],
                  m[i] + nbar + (double)cx(i) * jbar.x + (double)cy(i) * jbar.y + (double)cz(i) * jbar.z);
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

// [N][Q] node-major <-> [Q][PS] plane-major (inactive nodes read as zero on the way out)
__global__ void k_nodes_to_planes(const double* __restrict__ src, double* __restrict__ dst, size_t N, size_t PS, int Q) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N * Q; i += (size_t)gridDim.x * blockDim.x) {
        const size_t node = i / Q, q = i % Q;
        dst[q * PS + node] = src[i];
    }
}
__global__ void k_planes_to_nodes(const double* __restrict__ src, double* __restrict__ dst, const uint32_t* __restrict__ link,
                                  size_t N, size_t PS, int Q) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N * Q; i += (size_t)gridDim.x * blockDim.x) {
        const size_t node = i / Q, q = i % Q;
        dst[i] = (!link || (link[node] & LINK_ACTIVE)) ? src[q * PS + node] : 0.0;
    }
}

}  // namespace

void launch_nodes_to_planes(const double* src, double* dst, size_t N, size_t PS, int Q, hipStream_t s) {
    k_nodes_to_planes<<<4096, 256, 0, s>>>(src, dst, N, PS, Q);
}
void launch_planes_to_nodes(const double* src, double* dst, const uint32_t* link, size_t N, size_t PS, int Q,
                            hipStream_t s) {
    k_planes_to_nodes<<<4096, 256, 0, s>>>(src, dst, link, N, PS, Q);
}

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
