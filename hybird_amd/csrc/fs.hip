// Free-surface mass tracking and interface reclassification for gfx950.
//
// The reference walks its interface list serially (updateMass LB.cpp:1492-1590,
// updateInterface 1592-1618 -> findInterfaceMutants 1620-1650,
// smoothenInterface 1652-1701, updateMutants 1703-1742, removeIsolated
// 1744-1794, redistributeMass 1796-1804). Here every phase is one grid-stride
// kernel over a compacted node list; the phases are split where one reads
// what another writes, so that the outcome does not depend on thread order:
//
//   mass      interface nodes: stream their populations (gas links rebuilt
//             from the stored velocity), exchange mass with the neighbours,
//             mark filled (m > rho) / emptied (m < 0)
//   apply     new masses and types of the mutants
//   spread    gas neighbours of filled nodes are claimed by the filled node of
//             highest index (the reference visits filled nodes in descending
//             index order and the first visitor initialises the node); fluid
//             neighbours of emptied nodes become interface
//   mutate    claimed gas nodes get their equilibrium populations; unclaimed
//             mutants give their surplus
//   isolated  interface nodes without gas neighbours become fluid, then those
//             without fluid neighbours become gas; the survivors form the next list
//   spread    the surplus divided by the number of interface nodes
//
// The surplus terms are summed exactly in a fixed-point accumulator
// (integer atomics), so the sum does not depend on the order of the terms.
// Node ownership and list membership live in one tagged word per node,
// (tag << 32) | value, where the tag changes every step: nothing is cleared.
#include <hip/hip_runtime.h>

#include "common.h"
#include "fs_kernels.h"
#include "hybird_amd.h"
#include "lattice.h"
#include "lb_kernels.h"

namespace hyb {

namespace {

constexpr int FS_BLOCK = 256;
constexpr int FS_GRID = 1024;
enum : uint8_t { FS_STAY = 0, FS_FILLED = 1, FS_EMPTIED = 2 };

#define FS_FOR(i, n) \
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

template <typename T>
__device__ __forceinline__ double to_d(T v) {
    return (double)v;
}

// ---- 1. streaming of the interface nodes + updateMass (LB.cpp:1492-1590)
// SHIFT: populations are stored as f - w (rho0 = 1); f_opp - fs_j is the same in
// both forms because w_opp(j) = w_j.
template <typename T, bool SHIFT>
__global__ void __launch_bounds__(FS_BLOCK)
    k_fs_mass(const T* __restrict__ fpost, T* __restrict__ fout, const uint8_t* __restrict__ type,
              const double* __restrict__ mass, const T* __restrict__ rho, const double* __restrict__ uS,
              const uint32_t* __restrict__ list, FSCounters* __restrict__ ctr, int cur, double* __restrict__ mnew,
              uint8_t* __restrict__ state, LBKParams p, int pull) {
#pragma clang fp contract(off)
    const size_t N = (size_t)p.X * p.Y * p.Z;
    const uint32_t nI = ctr->n[cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // counters of this step's later phases
        ctr->n[cur ^ 1] = 0;
        ctr->nA = 0;
        ctr->nB = 0;
        for (int i = 0; i < 4; ++i) ctr->acc[i] = 0ull;
        for (int i = 0; i < 4; ++i) ctr->tot[i] = 0ull;
    }
    const double rho0 = p.initDensity;
    FS_FOR(pos, nI) {
        const size_t idx = list[pos];
        const Coord c = coord_of(p, idx);
        T fs[19], fi[19];
#pragma unroll
        for (int j = 0; j < 19; ++j) fs[j] = fpost[(size_t)j * p.PS + idx];
        const double mi = mass[idx];
        const V3 u(uS[idx], uS[N + idx], uS[2 * N + idx]);
        if (p.actPrev) {  // type class before this pass of the nodes it may change
            uint32_t* ap = const_cast<uint32_t*>(p.actPrev);
            ap[idx] = (p.actTag << 2) | 2u;
            for (int j = 1; j < 19; ++j) {
                const size_t L = nb_index(p, c, j);
                if (type[L] == T_GAS) ap[L] = p.actTag << 2;
            }
        }
        const double usq = u.norm2();
        // every neighbour's type and mass, and in the pull form its opposite population, loaded
        // together before any is used (a neighbour index is always a node of the lattice, so the loads a
        // neighbour's type makes unnecessary stay in bounds): one round trip instead of one per link
        uint8_t tn[19];
        T vn[19];
        double mL[19];
#pragma unroll
        for (int j = 1; j < 19; ++j) {
            const size_t L = nb_index(p, c, j);
            tn[j] = type[L];
            vn[j] = pull ? fpost[(size_t)opp(j) * p.PS + L] : (T)0;
            mL[j] = mass[L];
        }
        fi[0] = fs[0];
        double dm = 0.0;
#pragma unroll
        for (int j = 1; j < 19; ++j) {
            const int o = opp(j);
            const uint8_t tl = tn[j];
            double a = 0.0;
            T v;
            if (!pull) {
                v = fs[o];  // before the first collision f = fs
                if (tl == T_INTERFACE) a = 0.5 * (mL[j] + mi);
                else if (tl == T_FLUID) a = 1.0;
                else if (tl == T_WALL_MOVING || tl == T_CURVED) a = 1.0 * mi;
                else if (tl == T_SLIP_MOVING) {
                    const auto active = [&](size_t n) { const uint8_t t = type[n]; return t == T_FLUID || t == T_INTERFACE; };
                    bool half = false;
                    if (j > 6) {
                        const bool a1 = active(nb_index(p, c, slip1check(j))), a2 = active(nb_index(p, c, slip2check(j)));
                        half = (a1 && !a2) || (!a1 && a2);
                    }
                    a = half ? 0.0 + 1.0 * p.slipS2 * mi : 0.0 + 1.0 * mi;
                }
            } else if (tl == T_FLUID || tl == T_INTERFACE) {
                v = vn[j];  // fpost[opp][L], LB.cpp:1234
                a = (tl == T_INTERFACE) ? 0.5 * (mL[j] + mi) : 1.0;
            } else if (tl == T_GAS) {
                // LB.cpp:1237-1246 with the shifted velocity of the last collision
                const double vu = u.dot(V3((double)cx(j), (double)cy(j), (double)cz(j)));
                const double g = wgt(j) * rho0 * (2.0 + 9.0 * (vu * vu) - 3.0 * usq);
                if (SHIFT) v = (T)(-to_d(fs[j]) + (g - 2.0 * wgt(j)));
                else v = (T)(-to_d(fs[j]) + g);
            } else {
                // plane wall: bounce-back, slip reflection (LB.cpp:1321-1456); moving
                // walls have zero velocity. Mass exchange: a = m for a moving wall,
                // the slip rule for a moving slip wall (LB.cpp:1531-1575), else 0.
                v = fs[j];
                const auto active = [&](size_t n) { const uint8_t t = type[n]; return t == T_FLUID || t == T_INTERFACE; };
                if (is_slip_wall(tl)) v = wall_link<T>(fpost, p.PS, p, c, true, j, fs[j], active);
                if (tl == T_CURVED) {  // Mei-Luo-Shyy (LB.cpp:1278-1319), a = m (LB.cpp:1539-1545)
                    v = curved_link<T, SHIFT>(p, c, j, fs[j]);
                    a = 1.0 * mi;
                } else if (tl == T_WALL_MOVING) {
                    a = 1.0 * mi;
                } else if (tl == T_SLIP_MOVING) {
                    bool half = false;
                    if (j > 6) {
                        const bool a1 = active(nb_index(p, c, slip1check(j))), a2 = active(nb_index(p, c, slip2check(j)));
                        half = (a1 && !a2) || (!a1 && a2);
                    }
                    a = half ? 0.0 + 1.0 * p.slipS2 * mi : 0.0 + 1.0 * mi;
                }
            }
            fi[o] = v;
            dm += 1.0 * a * (to_d(v) - to_d(fs[j]));
        }
        const double mn = mi + dm;
        const double n = to_d(rho[idx]);
        state[pos] = mn > n ? FS_FILLED : (mn < 0.0 ? FS_EMPTIED : FS_STAY);
        mnew[pos] = mn;
#pragma unroll
        for (int j = 0; j < 19; ++j) fout[(size_t)j * p.PS + idx] = fi[j];
    }
}

// ---- 2. findInterfaceMutants (LB.cpp:1620-1650): new masses and types
__global__ void __launch_bounds__(FS_BLOCK)
    k_fs_apply(const uint32_t* __restrict__ list, const FSCounters* __restrict__ ctr, int cur,
               const double* __restrict__ mnew, const uint8_t* __restrict__ state, double* __restrict__ mass,
               uint8_t* __restrict__ type, uint32_t* __restrict__ link) {
    const uint32_t nI = ctr->n[cur];
    FS_FOR(pos, nI) {
        const size_t idx = list[pos];
        mass[idx] = mnew[pos];
        uint32_t lw = link[idx] | LINK_PRESTREAMED;
        const uint8_t s = state[pos];
        if (s == FS_FILLED) {
            type[idx] = T_FLUID;
            lw &= ~LINK_INTERFACE;
        } else if (s == FS_EMPTIED) {
            type[idx] = T_GAS;
            lw &= ~(LINK_ACTIVE | LINK_INTERFACE);
        }
        link[idx] = lw;
    }
}

// ---- 3. smoothenInterface (LB.cpp:1652-1701)
// Filled nodes claim their gas neighbours (highest filled index wins); emptied
// nodes turn their fluid neighbours (the node itself included, j = 0) into
// interface nodes with 99 % of their density as mass. Phase A reads only gas
// types and phase B changes only fluid types, so both run in one kernel.
template <typename T>
__global__ void __launch_bounds__(FS_BLOCK)
    k_fs_spread(const uint32_t* __restrict__ list, FSCounters* __restrict__ ctr, int cur,
                const uint8_t* __restrict__ state, uint8_t* __restrict__ type, uint32_t* __restrict__ link,
                unsigned long long* __restrict__ claim, uint32_t* __restrict__ listA, uint32_t* __restrict__ listB,
                double* __restrict__ mass, const T* __restrict__ rho, uint32_t tag, LBKParams p) {
#pragma clang fp contract(off)
    const uint32_t nI = ctr->n[cur];
    const unsigned long long tagv = (unsigned long long)tag << 32;
    FS_FOR(pos, nI) {
        const uint8_t s = state[pos];
        if (s == FS_STAY) continue;
        const size_t idx = list[pos];
        const Coord c = coord_of(p, idx);
        if (s == FS_FILLED) {
            for (int j = 1; j < 19; ++j) {
                const size_t L = nb_index(p, c, j);
                if (type[L] != T_GAS) continue;
                const unsigned long long old = atomicMax(&claim[L], tagv | (unsigned long long)(idx + 1));
                if (old < tagv) listA[atomicAdd(&ctr->nA, 1u)] = (uint32_t)L;
            }
        } else {
            for (int j = 0; j < 19; ++j) {
                const size_t L = j == 0 ? idx : nb_index(p, c, j);
                if (type[L] != T_FLUID) continue;
                const unsigned long long old = atomicMax(&claim[L], tagv | 1ull);
                if (old >= tagv) continue;
                listB[atomicAdd(&ctr->nB, 1u)] = (uint32_t)L;
                if (p.actPrev) const_cast<uint32_t*>(p.actPrev)[L] = (p.actTag << 2) | 1u;  // was fluid
                type[L] = T_INTERFACE;
                link[L] |= LINK_INTERFACE;
                const double n = to_d(rho[L]);
                mass[L] = 0.99 * n;
                exact_add(ctr->acc, 0.01 * n);
            }
        }
    }
}

// ---- 4. new interface nodes from gas (LB.cpp:1668-1682, node.cpp:26-61) and
// updateMutants (LB.cpp:1703-1742)
template <typename T, bool SHIFT>
__global__ void __launch_bounds__(FS_BLOCK)
    k_fs_mutate(const uint32_t* __restrict__ list, FSCounters* __restrict__ ctr, int cur,
                const uint8_t* __restrict__ state, uint8_t* __restrict__ type, uint32_t* __restrict__ link,
                const unsigned long long* __restrict__ claim, const uint32_t* __restrict__ listA,
                double* __restrict__ mass, T* __restrict__ rho, double* __restrict__ uS, double* __restrict__ hFi,
                double* __restrict__ visc, T* __restrict__ fout, uint32_t* __restrict__ slipw, uint32_t tag, double Fx,
                double Fy, double Fz, LBKParams p) {
#pragma clang fp contract(off)
    const size_t N = (size_t)p.X * p.Y * p.Z;
    const uint32_t nI = ctr->n[cur], nA = ctr->nA;
    const uint32_t n = nI > nA ? nI : nA;
    const double rho0 = p.initDensity;
    FS_FOR(pos, n) {
        if (pos < nA) {
            const size_t L = listA[pos];
            const size_t f = (size_t)(uint32_t)(claim[L] & 0xFFFFFFFFull) - 1;
            // nodeInitialize(rho0, u_f, 0.01 rho0, visc_f, hydroForce_f + lbF)
            const V3 F = V3(hFi[f], hFi[N + f], hFi[2 * N + f]) + V3(Fx, Fy, Fz);
            const V3 uf(uS[f], uS[N + f], uS[2 * N + f]);
            const V3 u = (1.0 * F) / 2.0 / rho0 + uf;
            const double usq = u.norm2();
#pragma unroll
            for (int j = 0; j < 19; ++j) {
                const double vu = u.dot(V3((double)cx(j), (double)cy(j), (double)cz(j)));
                const double feq = wgt(j) * rho0 * (1.0 + 3.0 * vu + 4.5 * vu * vu - 1.5 * usq);
                fout[(size_t)j * p.PS + L] = SHIFT ? (T)(feq - wgt(j)) : (T)feq;
            }
            const Coord c = coord_of(p, L);
            uint32_t lw = LINK_ACTIVE | LINK_INTERFACE | LINK_PRESTREAMED, sw = 0, cw = 0;
            for (int j = 1; j < 19; ++j) {
                const uint8_t tl = type[nb_index(p, c, j)];
                if (is_plane_wall(tl) || tl == T_CURVED) lw |= 1u << j;
                if (is_slip_wall(tl)) sw |= 1u << j;
                if (tl == T_CURVED) cw |= 1u << j;
            }
            if (sw) lw |= LINK_SLIP;
            if (cw) lw |= LINK_CURVED;
            if (slipw) slipw[L] = sw;
            if (p.curvw) const_cast<uint32_t*>(p.curvw)[L] = cw;
            link[L] = lw;
            type[L] = T_INTERFACE;
            mass[L] = 0.01 * rho0;
            rho[L] = (T)rho0;
            uS[L] = u.x; uS[N + L] = u.y; uS[2 * N + L] = u.z;
            hFi[L] = 0.0; hFi[N + L] = 0.0; hFi[2 * N + L] = 0.0;
            if (visc) visc[L] = visc[f];
            exact_add(ctr->acc, -(0.01 * rho0));
        }
        if (pos < nI) {
            const uint8_t s = state[pos];
            if (s == FS_STAY) continue;
            const size_t idx = list[pos];
            if ((claim[idx] >> 32) == tag) continue;  // became interface again this step
            if (s == FS_EMPTIED) {
                exact_add(ctr->acc, mass[idx]);
            } else {
                const double nn = to_d(rho[idx]);
                exact_add(ctr->acc, mass[idx] - nn);
                mass[idx] = nn;
            }
        }
    }
}

// Interface candidates after smoothenInterface: unchanged list entries, then the
// gas -> interface list, then the fluid -> interface list.
__device__ __forceinline__ bool fs_candidate(uint32_t q, uint32_t nI, uint32_t nA, const uint32_t* list,
                                             const uint8_t* state, const uint32_t* listA, const uint32_t* listB,
                                             size_t& idx) {
    if (q < nI) {
        if (state[q] != FS_STAY) return false;
        idx = list[q];
    } else if (q < nI + nA) {
        idx = listA[q - nI];
    } else {
        idx = listB[q - nI - nA];
    }
    return true;
}

// ---- 5. removeIsolated, first loop (LB.cpp:1744-1770): no gas neighbour -> fluid
template <typename T>
__global__ void __launch_bounds__(FS_BLOCK)
    k_fs_iso_fluid(const uint32_t* __restrict__ list, FSCounters* __restrict__ ctr, int cur,
                   const uint8_t* __restrict__ state, const uint32_t* __restrict__ listA,
                   const uint32_t* __restrict__ listB, uint8_t* __restrict__ type, uint32_t* __restrict__ link,
                   double* __restrict__ mass, const T* __restrict__ rho, LBKParams p) {
#pragma clang fp contract(off)
    const uint32_t nI = ctr->n[cur], nA = ctr->nA, nB = ctr->nB;
    FS_FOR(q, nI + nA + nB) {
        size_t idx;
        if (!fs_candidate(q, nI, nA, list, state, listA, listB, idx)) continue;
        const Coord c = coord_of(p, idx);
        uint8_t tn[19];  // all loaded before the test: one round trip
#pragma unroll
        for (int j = 1; j < 19; ++j) tn[j] = type[nb_index(p, c, j)];
        bool gasNb = false;
#pragma unroll
        for (int j = 1; j < 19; ++j) gasNb = gasNb || tn[j] == T_GAS;
        if (gasNb) continue;
        const double n = to_d(rho[idx]);
        exact_add(ctr->acc, mass[idx] - n);
        mass[idx] = n;
        type[idx] = T_FLUID;
        link[idx] &= ~LINK_INTERFACE;
    }
}

// ---- 6. removeIsolated, second loop (LB.cpp:1771-1792): no fluid neighbour ->
// gas; the remaining interface nodes form the next list
__global__ void __launch_bounds__(FS_BLOCK)
    k_fs_iso_gas(const uint32_t* __restrict__ list, FSCounters* __restrict__ ctr, int cur,
                 const uint8_t* __restrict__ state, const uint32_t* __restrict__ listA,
                 const uint32_t* __restrict__ listB, uint8_t* __restrict__ type, uint32_t* __restrict__ link,
                 const double* __restrict__ mass, uint32_t* __restrict__ next, LBKParams p) {
    const uint32_t nI = ctr->n[cur], nA = ctr->nA, nB = ctr->nB;
    FS_FOR(q, nI + nA + nB) {
        size_t idx;
        if (!fs_candidate(q, nI, nA, list, state, listA, listB, idx)) continue;
        if (type[idx] != T_INTERFACE) continue;
        const Coord c = coord_of(p, idx);
        uint8_t tn[19];  // all loaded before the test: one round trip
#pragma unroll
        for (int j = 1; j < 19; ++j) tn[j] = type[nb_index(p, c, j)];
        bool fluidNb = false;
#pragma unroll
        for (int j = 1; j < 19; ++j) fluidNb = fluidNb || tn[j] == T_FLUID;
        if (fluidNb) {
            next[atomicAdd(&ctr->n[cur ^ 1], 1u)] = (uint32_t)idx;
        } else {
            exact_add(ctr->acc, mass[idx]);
            type[idx] = T_GAS;
            link[idx] = 0u;
        }
    }
}

// ---- 7. redistributeMass (LB.cpp:1796-1804)
__global__ void __launch_bounds__(FS_BLOCK)
    k_fs_redistribute(const uint32_t* __restrict__ next, const FSCounters* __restrict__ ctr, int cur,
                      double* __restrict__ mass) {
#pragma clang fp contract(off)
    const uint32_t n = ctr->n[cur ^ 1];
    if (n == 0) return;
    const double add = exact_value(ctr->acc) / (double)n;
    FS_FOR(q, n) mass[next[q]] += add;
}

// ---- initial interface list
__global__ void k_fs_build_list(const uint8_t* __restrict__ type, uint32_t* __restrict__ list,
                                FSCounters* __restrict__ ctr, int cur, size_t N) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
        const bool it = type[i] == T_INTERFACE;
        const unsigned long long b = __ballot(it);
        const int lane = threadIdx.x & 63;
        uint32_t base = 0;
        if (lane == 0 && b) base = atomicAdd(&ctr->n[cur], (uint32_t)__popcll(b));
        base = __shfl(base, 0, 64);
        if (it) list[base + __popcll(b & ((1ull << lane) - 1ull))] = (uint32_t)i;
    }
}

// ---- post-stream populations for the host view (LB.cpp:1144-1488)
template <typename T, bool SHIFT>
__global__ void k_fs_stream_view(const T* __restrict__ fin, T* __restrict__ fout, const uint8_t* __restrict__ type,
                                 const uint32_t* __restrict__ link, const double* __restrict__ uS, LBKParams p) {
#pragma clang fp contract(off)
    const size_t N = (size_t)p.X * p.Y * p.Z;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < N; idx += (size_t)gridDim.x * blockDim.x) {
        if (!(link[idx] & LINK_ACTIVE)) continue;
        const Coord c = coord_of(p, idx);
        const bool itf = (link[idx] & LINK_INTERFACE) != 0;
        V3 u;
        if (itf) u = V3(uS[idx], uS[N + idx], uS[2 * N + idx]);
        const double usq = u.norm2();
        fout[idx] = fin[idx];
        for (int j = 1; j < 19; ++j) {
            const int o = opp(j);
            const size_t L = nb_index(p, c, j);
            const uint8_t tl = type[L];
            T v;
            if (tl == T_FLUID || tl == T_INTERFACE) {
                v = fin[(size_t)o * p.PS + L];
            } else if (tl == T_GAS) {
                const double vu = u.dot(V3((double)cx(j), (double)cy(j), (double)cz(j)));
                const double g = wgt(j) * p.initDensity * (2.0 + 9.0 * (vu * vu) - 3.0 * usq);
                const double fsj = (double)fin[(size_t)j * p.PS + idx];
                v = SHIFT ? (T)(-fsj + (g - 2.0 * wgt(j))) : (T)(-fsj + g);
            } else {
                v = fin[(size_t)j * p.PS + idx];
                if (tl == T_CURVED)
                    v = curved_link<T, SHIFT>(p, c, j, v);
                else if (is_slip_wall(tl))
                    v = wall_link<T>(fin, p.PS, p, c, true, j, v,
                                     [&](size_t n) { const uint8_t t = type[n]; return t == T_FLUID || t == T_INTERFACE; });
            }
            fout[(size_t)o * p.PS + idx] = v;
        }
    }
}

}  // namespace

// Extra mass of the curved-wall rule in the streaming of the last LB step
// (LB.cpp:1315-1317): every active node adds m (chi fs_j - chi fStar + BBi) per
// curved link, m its mass (interface mass, or the fluid node's last density).
// Summed exactly into ctr->extra.
template <typename T, bool SHIFT>
__global__ void __launch_bounds__(FS_BLOCK) k_fs_curved_extra(const T* __restrict__ fpost,
                                                              const uint32_t* __restrict__ link,
                                                              const double* __restrict__ mass,
                                                              const T* __restrict__ rhoFluid, FSCounters* __restrict__ ctr,
                                                              double rho0, int massKnown, LBKParams p) {
#pragma clang fp contract(off)
    const size_t N = (size_t)p.X * p.Y * p.Z;
    FS_FOR(i, (uint32_t)N) {
        const uint32_t lw = link[i];
        if (!(lw & LINK_ACTIVE) || !(lw & LINK_CURVED)) continue;
        const uint32_t cw = p.curvw[i];
        const Coord c = coord_of(p, i);
        double m;
        if (lw & LINK_INTERFACE) m = mass[i];
        else m = massKnown ? to_d(rhoFluid[i]) : rho0;
        for (int j = 1; j < 19; ++j) {
            if (!(cw & (1u << j))) continue;
            const CurvedParts q = curved_parts(p, c, j);
            const double fs = SHIFT ? to_d(fpost[(size_t)j * p.PS + i]) + wgt(j) : to_d(fpost[(size_t)j * p.PS + i]);
            exact_add(ctr->extra, m * (q.chi * fs - q.chi * q.fStar + q.BBi));
        }
    }
}

// redistributeMass(extraMass) at the end of the streaming (LB.cpp:1477,
// 1796-1804): every interface node of the current list gains extra / count.
// One workgroup; the accumulator is cleared for the next step.
__global__ void __launch_bounds__(FS_BLOCK) k_fs_add_extra(const uint32_t* __restrict__ list,
                                                           FSCounters* __restrict__ ctr, int cur,
                                                           double* __restrict__ mass) {
#pragma clang fp contract(off)
    const uint32_t nI = ctr->n[cur];
    const double add = exact_value(ctr->extra) / (double)nI;
    for (uint32_t pos = threadIdx.x; pos < nI; pos += blockDim.x) mass[list[pos]] += add;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < 4; ++i) ctr->extra[i] = 0ull;
}

// Mass of the active nodes outside particles after the pass (LB.cpp:1807-1813):
// interface nodes their mass, fluid nodes their last density; summed exactly.
template <typename T>
__global__ void __launch_bounds__(FS_BLOCK) k_fs_total_mass(const uint8_t* __restrict__ type,
                                                            const uint32_t* __restrict__ solid, uint32_t ctag,
                                                            const double* __restrict__ mass, const T* __restrict__ rho,
                                                            FSCounters* __restrict__ ctr, size_t N) {
    long long acc[4] = {0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
        const uint8_t t = type[i];
        if (t != T_FLUID && t != T_INTERFACE) continue;
        if (solid && owner_slot(solid[i], ctag)) continue;
        exact_local(acc, t == T_INTERFACE ? mass[i] : to_d(rho[i]));
    }
    exact_flush(ctr->tot, acc);
}

// redistributeMass(-0.01 * (mass - totalMass)) over the new interface list
// (LB.cpp:1815-1821, 1796-1804), grid-stride; the next pass's first kernel
// clears the sum.
__global__ void __launch_bounds__(FS_BLOCK) k_fs_mass_fix(const uint32_t* __restrict__ next,
                                                          const FSCounters* __restrict__ ctr, int cur,
                                                          double* __restrict__ mass, double totalMass) {
#pragma clang fp contract(off)
    const uint32_t n = ctr->n[cur ^ 1];
    if (!n) return;
    const double massDeficit = exact_value(ctr->tot) - totalMass;
    const double massSurplus = -0.01 * massDeficit;
    const double add = massSurplus / (double)n;
    FS_FOR(q, n) mass[next[q]] += add;
}

void launch_fs_mass_fix(Prec prec, const uint8_t* type, const uint32_t* solid, uint32_t ctag, FSBuffers& b, int cur,
                        int rcur, double totalMass, const LBKParams& p, hipStream_t s) {
    const size_t N = (size_t)p.X * p.Y * p.Z;
    size_t g = (N + FS_BLOCK - 1) / FS_BLOCK;
    if (g > FS_GRID) g = FS_GRID;
    if (prec == Prec::F64)
        k_fs_total_mass<double><<<(unsigned)g, FS_BLOCK, 0, s>>>(type, solid, ctag, b.mass,
                                                                 (const double*)b.rho[rcur], b.ctr, N);
    else
        k_fs_total_mass<float><<<(unsigned)g, FS_BLOCK, 0, s>>>(type, solid, ctag, b.mass, (const float*)b.rho[rcur],
                                                                b.ctr, N);
    k_fs_mass_fix<<<FS_GRID, FS_BLOCK, 0, s>>>(b.list[cur ^ 1], b.ctr, cur, b.mass, totalMass);
}

void launch_fs_curved_extra(Prec prec, const void* fpost, const uint32_t* link, FSBuffers& b, int cur,
                            const void* rhoFluid, bool massKnown, const LBKParams& p, hipStream_t s) {
    if (prec == Prec::F64)
        k_fs_curved_extra<double, false><<<FS_GRID, FS_BLOCK, 0, s>>>((const double*)fpost, link, b.mass,
                                                                      (const double*)rhoFluid, b.ctr, p.initDensity,
                                                                      massKnown ? 1 : 0, p);
    else
        k_fs_curved_extra<float, true><<<FS_GRID, FS_BLOCK, 0, s>>>((const float*)fpost, link, b.mass,
                                                                    (const float*)rhoFluid, b.ctr, p.initDensity,
                                                                    massKnown ? 1 : 0, p);
    k_fs_add_extra<<<1, FS_BLOCK, 0, s>>>(b.list[cur], b.ctr, cur, b.mass);
}

void launch_fs_build_list(const uint8_t* type, FSBuffers& b, int cur, const LBKParams& p, hipStream_t s) {
    const size_t N = (size_t)p.X * p.Y * p.Z;
    size_t g = (N + 255) / 256;
    if (g > 4096) g = 4096;
    k_fs_build_list<<<(unsigned)g, 256, 0, s>>>(type, b.list[cur], b.ctr, cur, N);
}

template <typename T, bool SHIFT>
static void fs_step_t(const T* fpost, T* fout, uint8_t* type, uint32_t* link, uint32_t* slipw, double* visc, FSBuffers& b,
                      int cur, int rcur, uint32_t tag, const double* Fnew, bool pull, const LBKParams& p, hipStream_t s) {
    const T* rho = (const T*)b.rho[rcur];
    k_fs_mass<T, SHIFT><<<FS_GRID, FS_BLOCK, 0, s>>>(fpost, fout, type, b.mass, rho, b.uS, b.list[cur], b.ctr, cur,
                                                    b.mnew, b.state, p, pull ? 1 : 0);
    k_fs_apply<<<FS_GRID, FS_BLOCK, 0, s>>>(b.list[cur], b.ctr, cur, b.mnew, b.state, b.mass, type, link);
    k_fs_spread<T><<<FS_GRID, FS_BLOCK, 0, s>>>(b.list[cur], b.ctr, cur, b.state, type, link, b.claim, b.listA, b.listB,
                                               b.mass, rho, tag, p);
    k_fs_mutate<T, SHIFT><<<FS_GRID, FS_BLOCK, 0, s>>>(b.list[cur], b.ctr, cur, b.state, type, link, b.claim, b.listA,
                                                      b.mass, (T*)b.rho[rcur], b.uS, b.hFi, visc, fout, slipw, tag,
                                                      Fnew[0], Fnew[1], Fnew[2], p);
    k_fs_iso_fluid<T><<<FS_GRID, FS_BLOCK, 0, s>>>(b.list[cur], b.ctr, cur, b.state, b.listA, b.listB, type, link,
                                                  b.mass, rho, p);
    k_fs_iso_gas<<<FS_GRID, FS_BLOCK, 0, s>>>(b.list[cur], b.ctr, cur, b.state, b.listA, b.listB, type, link, b.mass,
                                             b.list[cur ^ 1], p);
    k_fs_redistribute<<<FS_GRID, FS_BLOCK, 0, s>>>(b.list[cur ^ 1], b.ctr, cur, b.mass);
}

void launch_fs_step(Prec prec, const void* fpost, void* fout, uint8_t* type, uint32_t* link, uint32_t* slipw,
                    double* visc, FSBuffers& b, int cur, int rcur, uint32_t tag, const double* Fnew, bool pull,
                    const LBKParams& p, hipStream_t s) {
    if (prec == Prec::F64)
        fs_step_t<double, false>((const double*)fpost, (double*)fout, type, link, slipw, visc, b, cur, rcur, tag, Fnew,
                                 pull, p, s);
    else
        fs_step_t<float, true>((const float*)fpost, (float*)fout, type, link, slipw, visc, b, cur, rcur, tag, Fnew, pull,
                               p, s);
}

void launch_fs_stream_view(Prec prec, const void* fin, void* fout, const uint8_t* type, const uint32_t* link,
                           const double* uS, const LBKParams& p, hipStream_t s) {
    const size_t N = (size_t)p.X * p.Y * p.Z;
    size_t g = (N + 255) / 256;
    if (g > 8192) g = 8192;
    if (prec == Prec::F64)
        k_fs_stream_view<double, false><<<(unsigned)g, 256, 0, s>>>((const double*)fin, (double*)fout, type, link, uS, p);
    else
        k_fs_stream_view<float, true><<<(unsigned)g, 256, 0, s>>>((const float*)fin, (float*)fout, type, link, uS, p);
}

}  // namespace hyb
