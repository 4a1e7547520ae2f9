// Lattice-Boltzmann kernels for gfx950.
//
// Populations live in HBM as structure-of-arrays planes f[q][z][y][x] (two
// copies, ping-pong), so every population load and store of a wavefront is a
// contiguous run along x. One fused kernel does, per active node: pull
// streaming with half-way bounce-back and periodic wrap (LB.cpp:1144-1488),
// moments (node.cpp:63-83), the volumetric penalty force of particle nodes
// (LB.cpp:1851-1919), the Guo velocity shift (node.cpp:85-88), equilibrium
// (node.cpp:90-101), optional Bingham/Smagorinsky viscosity (node.cpp:103-145),
// BGK (node.cpp:158-165) or TRT relaxation and Guo forcing (node.cpp:167-184).
//
// FP64 form: the reference's operation order, compiled without FMA contraction,
// so that results reproduce the CPU reference. FP32 form: populations are stored
// as deviations g = f - w (rho0 = 1) and the collision runs in FP32 with FMA.
#include <hip/hip_runtime.h>

#include "common.h"
#include "hybird_amd.h"
#include "fs_kernels.h"
#include "lattice.h"
#include "lb_kernels.h"

namespace hyb {

namespace {

// Occupancy of the fused kernels: the plain instances (no variable viscosity,
// no curved walls) are compiled for 8 (FP32) and 5 (FP64) wavefronts per SIMD,
// which they reach without spilling (more loads in flight for a bandwidth-bound
// kernel); the others keep the compiler's register choice.
// Workgroup tile of the lattice kernels: a full 256-node x-row per workgroup
// keeps the +-1 x-shifted loads of a row inside one workgroup's cache.
#ifndef HYB_LB_BX
#define HYB_LB_BX 256
#endif
#ifndef HYB_LB_BY
#define HYB_LB_BY 1
#endif
#ifndef HYB_LB_NT
#define HYB_LB_NT 1
#endif
constexpr int BX = HYB_LB_BX;

constexpr int BY = HYB_LB_BY;

#ifndef HYB_LB_NTLOAD
#define HYB_LB_NTLOAD 1
#endif


// Population stores and loads: optionally non-temporal (every population is
// written once and read once per step, so neither benefits from caching).
template <typename T>
__device__ __forceinline__ void st_pop(T* p, T v) {
#if HYB_LB_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}
template <typename T>
__device__ __forceinline__ T ld_pop(const T* p) {
#if HYB_LB_NTLOAD
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

__device__ __forceinline__ bool node_coord(const LBKParams& p, Coord& c) {
    c.x = blockIdx.x * BX + threadIdx.x;
    c.y = blockIdx.y * BY + threadIdx.y;
    c.z = blockIdx.z;
    if (c.x >= p.X || c.y >= p.Y) return false;
    c.idx = (size_t)c.x + (size_t)p.X * ((size_t)c.y + (size_t)p.Y * (size_t)c.z);
    return true;
}

// Whether node n was active when the reference streamed the populations this
// collision pulls, i.e. before the free-surface pass that ran in between.
__device__ __forceinline__ bool active_at_stream(const LBKParams& p, size_t n) {
    if (p.actPrev && p.actTag) {
        const uint32_t w = p.actPrev[n];
        if ((w >> 2) == p.actTag) return (w & 3u) != 0u;
    }
    return (p.linkw[n] & LINK_ACTIVE) != 0u;
}

// Gather the 19 pre-collision populations of node c into f[]: f_i is pulled
// from the neighbour x - c_i (32-bit node index into plane i). Links to static
// walls bounce back the node's own f_opp(i) instead; they only occur next to
// walls, so that fix-up runs as a separate, rarely taken branch and the common
// path issues its 19 loads back to back.
template <typename T, bool CURVED>
__device__ __forceinline__ void wall_fixup(const T* __restrict__ fin, size_t PS, const LBKParams& p, const Coord& c,
                                           uint32_t lw, T f[19]);

template <typename T, bool CURVED = false>
__device__ __forceinline__ void pull19(const T* __restrict__ fin, size_t PS, const LBKParams& p, const Coord& c,
                                       uint32_t lw, T f[19]) {
    const Nb3 nx = nb3(c.x, p.X, p.per[0]);
    const Nb3 ny = nb3(c.y, p.Y, p.per[1]);
    const Nb3 nz = nb3(c.z, p.Z, p.per[2]);
    const uint32_t sX = (uint32_t)p.X, sXY = (uint32_t)p.X * (uint32_t)p.Y;
    const uint32_t self = (uint32_t)c.idx;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int j = opp(i);
        const uint32_t nb = (uint32_t)pick(nx, cx(j)) + sX * (uint32_t)pick(ny, cy(j)) + sXY * (uint32_t)pick(nz, cz(j));
        // a wall neighbour's index is valid memory; its value is replaced below
        f[i] = ld_pop(&fin[(size_t)i * PS + (i == 0 ? self : nb)]);
    }
    wall_fixup<T, CURVED>(fin, PS, p, c, lw, f);
}

// The links of node c that end on a wall: bounce-back of the node's own
// f_opp(i), the slip reflection or the curved-wall rule, in place of the pulled
// value.
template <typename T, bool CURVED>
__device__ __forceinline__ void wall_fixup(const T* __restrict__ fin, size_t PS, const LBKParams& p, const Coord& c,
                                           uint32_t lw, T f[19]) {
    const uint32_t self = (uint32_t)c.idx;
    if (lw & LINK_WALLS) {
        const uint32_t sw = (lw & LINK_SLIP) ? p.slipw[self] : 0u;
        const uint32_t cw = (CURVED && (lw & LINK_CURVED)) ? p.curvw[self] : 0u;
#pragma unroll
        for (int i = 1; i < 19; ++i) {
            const int j = opp(i);
            if (!(lw & (1u << j))) continue;
            const T own = ld_pop(&fin[(size_t)j * PS + self]);
            if (CURVED && (cw & (1u << j)))
                f[i] = curved_link<T, sizeof(T) == 4>(p, c, j, own);
            else if (sw & (1u << j))
                f[i] = wall_link<T>(fin, PS, p, c, true, j, own, [&](size_t n) { return active_at_stream(p, n); });
            else
                f[i] = own;
        }
    }
}

// ------------------------------------------------------------------ FP64 path
// u.c_j with the zero components left out (j is a constant after unrolling);
// equal to the reference's three-term dot product (vector.cpp:134-136) because
// adding +-0 to a nonzero partial sum never changes it, and a zero result only
// enters later terms through squares or a sum with 1.
__device__ __forceinline__ double cdot_rt(int j, const V3& u) {
    double r = 0.0;
    bool first = true;
    if (cx(j) != 0) { r = cx(j) > 0 ? u.x : -u.x; first = false; }
    if (cy(j) != 0) { const double t = cy(j) > 0 ? u.y : -u.y; r = first ? t : r + t; first = false; }
    if (cz(j) != 0) { const double t = cz(j) > 0 ? u.z : -u.z; r = first ? t : r + t; first = false; }
    return r;
}

// node.cpp:63-83: n = sum f_j; u = sum f_j c_j / n (zero terms left out).
__device__ __forceinline__ void moments64(const double f[19], double& n, V3& u) {
    n = 0.0;
#pragma unroll
    for (int j = 0; j < 19; ++j) n += f[j];
    double ux = 0.0, uy = 0.0, uz = 0.0;
#pragma unroll
    for (int j = 1; j < 19; ++j) {
        if (cx(j) > 0) ux += f[j];
        if (cx(j) < 0) ux -= f[j];
        if (cy(j) > 0) uy += f[j];
        if (cy(j) < 0) uy -= f[j];
        if (cz(j) > 0) uz += f[j];
        if (cz(j) < 0) uz -= f[j];
    }
    u = V3(ux / n, uy / n, uz / n);
}

// The penalty force of a particle node (LB.cpp:1865-1890): returns the
// element-side increment diffVel and sets the node force hF = -diffVel.
__device__ __forceinline__ V3 penalty64(const CoupleRec& r, const Coord& c, double mass, double n, const V3& u,
                                        const LBKParams& p, V3& hF) {
    const V3 pos((double)c.x, (double)c.y, (double)c.z);
    const V3 x0(r.x0[0], r.x0[1], r.x0[2]);
    const V3 rv(r.rv[0], r.rv[1], r.rv[2]);
    const V3 v(r.v[0], r.v[1], r.v[2]);
    const V3 w(r.w[0], r.w[1], r.w[2]);
    const V3 radius = pos - x0 / p.unitLength + rv / p.unitLength;
    const V3 localVel = v / p.unitSpeed + (w.cross(radius)) / p.unitAngVel;
    const V3 diff = (mass / n) * (u - localVel);
    hF = V3(0.0 + diff.x * -1.0, 0.0 + diff.y * -1.0, 0.0 + diff.z * -1.0);
    return diff;
}

// node.cpp:103-145 with the node's previous viscosity.
__device__ __forceinline__ double shear_visc64(const double f[19], const double feq[19], double n, double viscOld,
                                               const LBKParams& p) {
    const double minVisc = (0.501 - 0.5) / 3 / 1.0;
    const double maxVisc = (1.8 - 0.5) / 3 / 1.0;
    const double tau = 0.5 + 3.0 * viscOld * 1.0;
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 1; j < 19; ++j) {
        const double d = f[j] - feq[j];
        const int c[3] = {cx(j), cy(j), cz(j)};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int cc = c[a] * c[b];
                if (cc > 0) g[a * 3 + b] += d;
                if (cc < 0) g[a * 3 + b] -= d;
            }
    }
    const double s = 1.5 * 1.0 / (tau * n);
#pragma unroll
    for (int k = 0; k < 9; ++k) g[k] *= s;
    const double sr = sqrt(0.5 * (g[0] * g[0] + g[4] * g[4] + g[8] * g[8] + 2.0 * (g[1] * g[3] + g[6] * g[2] + g[5] * g[7])));
    double nuTurb = 0.0;
    if (p.turbulence) nuTurb = p.turbConst * sr;
    double nuApp = p.nonNewtonian ? p.plasticVisc + p.yieldStress / (n * 2.0 * sr) : viscOld;
    if (p.frictional) {
        // Coulomb-like yield stress from the gauge pressure (n - rho0) c_s^2, none under tension
        const double pg = (n - p.initDensity) / 3.0;
        const double ty = p.frictionFluid * (pg > 0.0 ? pg : 0.0);
        nuApp = ty > 0.0 ? p.plasticVisc + ty / (n * 2.0 * sr) : p.plasticVisc;
    }
    const double a = nuApp + nuTurb;
    // std::min(maxVisc, a) then std::max(minVisc, .): NaN and +inf end at maxVisc
    const double lo = (a < maxVisc) ? a : maxVisc;
    return (minVisc < lo) ? lo : minVisc;
}

// Node mass of the penalty force (LB.cpp:1880): without free surface the
// initial masses (rho0, interface rho0/2) never change; with free surface an
// interface node carries its tracked mass and a fluid node the density of its
// previous collision (updateMass, LB.cpp:1582-1584).
template <bool FS, typename T>
__device__ __forceinline__ double node_mass(uint32_t lw, size_t idx, const FSView& fsv, const LBKParams& p) {
    if (FS) return (lw & LINK_INTERFACE) ? fsv.mass[idx] : (double)((const T*)fsv.rhoPrev)[idx];
    return (lw & LINK_INTERFACE) ? 0.5 * p.initDensity : p.initDensity;
}

// Free-surface mode: nodes the free-surface pass has streamed already (or
// initialised) hold their populations at their own index of the output buffer.
// Each thread reads those before it writes the same addresses, and every store
// depends on the loaded values, so the loads cannot move past the stores.
template <typename T>
__device__ __forceinline__ void load_prepared(const FSView& fsv, size_t PS, size_t idx, uint32_t lw, T f[19]) {
    const T* pre = (const T*)fsv.pre;
#pragma unroll
    for (int j = 0; j < 19; ++j) f[j] = pre[(size_t)j * PS + idx];
    fsv.linkw[idx] = lw & ~LINK_PRESTREAMED;
}

// ------------------------------------------------------------------ hydrodynamic run sums
// Per-slot hydrodynamic sums (LB.cpp:1897-1902) folded into the collision: the
// workgroup's nodes are 256 consecutive x nodes of one row, so the nodes one
// slot owns there form runs. Each particle node stages its penalty increment
// diffVel and node mass in LDS as the collision computes them (every node first
// marks its entry empty); at the end the first node of each run adds the run's
// diffVel, radius x diffVel (radius from the slot's FP64 centre and radius
// vector, LB.cpp:1875) and masses in x order in FP64 and appends one HydroSeg to
// the slot's list. The per-element combine adds the lists exactly (order-free),
// so the result does not depend on the order of the appends. Runs end at
// wavefront boundaries, so only lanes of one wavefront exchange data.
constexpr int RUN_NODES = 256;  // nodes of one workgroup (one row segment)
static_assert(BX == RUN_NODES && BY == 1, "run sums need the workgroup on 256 consecutive x nodes of one row");

template <typename DT>
struct RunLDS {
    uint32_t sid[RUN_NODES];
    DT d[3][RUN_NODES];
    double m[RUN_NODES];
};

// What a collision needs to stage its node: the LDS block and the node's
// position in it (nullptr: no sums this launch).
template <typename DT>
struct RunSlot {
    RunLDS<DT>* lds;
    int n;
    __device__ void clear() const {
        if (lds) lds->sid[n] = 0u;
    }
    __device__ void stage(uint32_t sid, DT dx, DT dy, DT dz, double mass) const {
        if (!lds) return;
        lds->sid[n] = sid;
        lds->d[0][n] = dx;
        lds->d[1][n] = dy;
        lds->d[2][n] = dz;
        lds->m[n] = mass;
    }
};

// The run sums of a workgroup whose lanes hold NPL consecutive nodes each
// (node n = NPL * lane + k, x = blockIdx.x * RUN_NODES + n of row (y, z)).
template <typename DT, int NPL>
__device__ __forceinline__ void hydro_runs(const RunLDS<DT>& L, const HydroSink& hs, const CoupleRec* __restrict__ cp,
                                           const LBKParams& p) {
#pragma clang fp contract(off)
    static_assert(RUN_NODES % (64 * NPL) == 0, "workgroup rows of whole wavefronts");
    constexpr int WAVE_NODES = 64 * NPL;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int y = (int)blockIdx.y, z = (int)blockIdx.z;
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
        const int n = NPL * (int)threadIdx.x + k;
        const uint32_t sid = L.sid[n];
        if (!sid) continue;
        if (n % WAVE_NODES != 0 && L.sid[n - 1] == sid) continue;  // not the first node of its run
        const uint32_t slot = sid - 1;
        const CoupleRec& r = cp[slot];
        const V3 x0(r.x0[0], r.x0[1], r.x0[2]);
        const V3 rv(r.rv[0], r.rv[1], r.rv[2]);
        double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const int end = (n / WAVE_NODES + 1) * WAVE_NODES;
        const int xb = (int)blockIdx.x * RUN_NODES;
        for (int m = n; m < end && L.sid[m] == sid; ++m) {
            const V3 d((double)L.d[0][m], (double)L.d[1][m], (double)L.d[2][m]);
            const V3 radius = V3((double)(xb + m), (double)y, (double)z) - x0 / p.unitLength + rv / p.unitLength;
            const V3 mo = radius.cross(d);
            acc[0] += d.x; acc[1] += d.y; acc[2] += d.z;
            acc[3] += mo.x; acc[4] += mo.y; acc[5] += mo.z;
            acc[6] += L.m[m];
        }
        const uint32_t i = atomicAdd(&hs.count[slot], 1u);
        if (i < hs.cap) {
            HydroSeg* o = hs.segs + (size_t)slot * hs.cap + i;
#pragma unroll
            for (int q = 0; q < 7; ++q) o->s[q] = acc[q];
            o->slot = slot;
            o->pad = 0;
        } else if (p.herr) {
            *p.herr = LB_ERR_SEGMENTS;
        }
    }
}

template <bool PULL, bool VARVISC, int COLL, bool FS, bool CURVED = false>
__device__ __forceinline__ void fused64_node(const double* __restrict__ fin, double* __restrict__ fout,
                                             const uint32_t* __restrict__ link,
                                             const uint32_t* __restrict__ solid,
                                                     const CoupleRec* __restrict__ cp, const RunSlot<double>& rs,
                                                     double* __restrict__ visc, double* __restrict__ mac,
                                                     const LBKParams& p, const FSView& fsv) {
#pragma clang fp contract(off)
    Coord c;
    if (!node_coord(p, c)) return;
    const size_t N = (size_t)p.X * p.Y * p.Z;
    const uint32_t lw = link[c.idx];
    if (!(lw & LINK_ACTIVE)) return;
    double f[19];
    if (FS && (lw & LINK_PRESTREAMED)) load_prepared<double>(fsv, p.PS, c.idx, lw, f);
    else if (PULL) pull19<double, CURVED>(fin, p.PS, p, c, lw, f);
    else {
#pragma unroll
        for (int j = 0; j < 19; ++j) f[j] = ld_pop(&fin[(size_t)j * p.PS + c.idx]);
    }
    double n;
    V3 u;
    moments64(f, n, u);
    if (p.herr && !isfinite(n)) *p.herr = LB_ERR_NUMERIC;
    // penalty coupling force
    V3 hF(0.0, 0.0, 0.0);
    const uint32_t sw = solid ? solid[c.idx] : 0u;
    const uint32_t sid = owner_slot(sw, p.ctag);
    if (sid) {
        const double mass = node_mass<FS, double>(lw, c.idx, fsv, p);
        const V3 d = penalty64(cp[sid - 1], c, mass, n, u, p, hF);
        rs.stage(sid, d.x, d.y, d.z, mass);
    }
    V3 Fb(p.F[0], p.F[1], p.F[2]);
    if (p.ffield) Fb = Fb + V3((double)p.ffield[c.idx], (double)p.ffield[N + c.idx], (double)p.ffield[2 * N + c.idx]);
    const V3 T = Fb + hF;
    u += 0.5 * T / n;  // node.cpp:85-88
    if (CURVED && p.macNext) {
        p.macNext[c.idx] = n;
        p.macNext[N + c.idx] = u.x;
        p.macNext[2 * N + c.idx] = u.y;
        p.macNext[3 * N + c.idx] = u.z;
    }
    if (mac) {
        mac[c.idx] = n;
        mac[N + c.idx] = u.x;
        mac[2 * N + c.idx] = u.y;
        mac[3 * N + c.idx] = u.z;
        mac[4 * N + c.idx] = hF.x;
        mac[5 * N + c.idx] = hF.y;
        mac[6 * N + c.idx] = hF.z;
    }
    if (!fout) return;
    if (FS) {
        ((double*)fsv.rhoCur)[c.idx] = n;
        if (lw & LINK_INTERFACE) {
            fsv.uS[c.idx] = u.x; fsv.uS[N + c.idx] = u.y; fsv.uS[2 * N + c.idx] = u.z;
            fsv.hFi[c.idx] = hF.x; fsv.hFi[N + c.idx] = hF.y; fsv.hFi[2 * N + c.idx] = hF.z;
        }
    }
    const double usq = u.norm2();
    // feq_j = w n (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2), node.cpp:90-101, evaluated where used
    auto feq = [&](int j) {
        const double vu = cdot_rt(j, u);
        return wgt(j) * n * (1.0 + 3.0 * vu + 4.5 * vu * vu - 1.5 * usq);
    };
    double nu = p.visc0;
    if (VARVISC) {
        double fe[19];
#pragma unroll
        for (int j = 0; j < 19; ++j) fe[j] = feq(j);
        nu = shear_visc64(f, fe, n, visc[c.idx], p);
        visc[c.idx] = nu;
    }
    if (COLL == HYB_COLLISION_TRT) {
        const double omP = 1.0 / (0.5 + 3.0 * nu);
        const double omM = 1.0 / (0.5 + p.trtMagic / (1.0 / omP - 0.5));
        const double uT = u.dot(T);
#pragma unroll
        for (int j = 0; j < 19; j += (j == 0 ? 1 : 2)) {
            const int o = opp(j);
            const double fj = f[j], fo = f[o], ej = feq(j), eo = (o == j) ? ej : feq(o);
            const double fp = 0.5 * (fj + fo), fm = 0.5 * (fj - fo);
            const double ep = 0.5 * (ej + eo), em = 0.5 * (ej - eo);
            const double vu = cdot_rt(j, u);
            const double cF = cdot_rt(j, T);
            const double sp = wgt(j) * (9.0 * vu * cF - 3.0 * uT);
            const double sm = wgt(j) * (3.0 * cF);
            st_pop(&fout[(size_t)j * p.PS + c.idx],
                   fj - omP * (fp - ep) - omM * (fm - em) + (1.0 - 0.5 * omP) * sp + (1.0 - 0.5 * omM) * sm);
            if (o != j) {
                // the same expression evaluated for the opposite direction
                const double fp2 = 0.5 * (fo + fj), fm2 = 0.5 * (fo - fj);
                const double ep2 = 0.5 * (eo + ej), em2 = 0.5 * (eo - ej);
                const double vu2 = cdot_rt(o, u);
                const double cF2 = cdot_rt(o, T);
                const double sp2 = wgt(o) * (9.0 * vu2 * cF2 - 3.0 * uT);
                const double sm2 = wgt(o) * (3.0 * cF2);
                st_pop(&fout[(size_t)o * p.PS + c.idx],
                       fo - omP * (fp2 - ep2) - omM * (fm2 - em2) + (1.0 - 0.5 * omP) * sp2 + (1.0 - 0.5 * omM) * sm2);
            }
        }
    } else {
        const double omega = 1.0 / (0.5 + 3.0 * nu * 1.0);
        const double omegaf = 1.0 - 1.0 / (1.0 + 6.0 * nu * 1.0);
#pragma unroll
        for (int j = 0; j < 19; ++j) {
            double fj = f[j] + omega * (feq(j) - f[j]);
            // Guo forcing: forcePart = 9 (c.u) c + 3 (c - u), dotted with the total force
            const double vu = cdot_rt(j, u);
            const double a = 9.0 * vu;
            const V3 fp((double)cx(j) * a + ((double)cx(j) - u.x) * 3.0, (double)cy(j) * a + ((double)cy(j) - u.y) * 3.0,
                        (double)cz(j) * a + ((double)cz(j) - u.z) * 3.0);
            fj += 1.0 * omegaf * wgt(j) * fp.dot(T);
            st_pop(&fout[(size_t)j * p.PS + c.idx], fj);
        }
    }
}

template <bool PULL, bool VARVISC, int COLL, bool FS, bool CURVED = false>
__global__ void __launch_bounds__(BX * BY, (VARVISC || CURVED) ? 1 : 5) k_fused64(const double* __restrict__ fin, double* __restrict__ fout,
                                                     const uint32_t* __restrict__ link,
                                                     const uint32_t* __restrict__ solid,
                                                     const CoupleRec* __restrict__ cp, HydroSink hs,
                                                     double* __restrict__ visc, double* __restrict__ mac,
                                                     LBKParams p, FSView fsv) {
    __shared__ RunLDS<double> lds;
    const bool sums = hs.segs != nullptr;
    const RunSlot<double> rs{sums ? &lds : nullptr, (int)threadIdx.x};
    rs.clear();
    fused64_node<PULL, VARVISC, COLL, FS, CURVED>(fin, fout, link, solid, cp, rs, visc, mac, p, fsv);
    if (sums) hydro_runs<double, 1>(lds, hs, cp, p);
}

// ------------------------------------------------------------------ FP32 deviation path
// g_j = f_j - w_j (rho0 = 1). Moments: rho = 1 + sum g, j = sum g c (sum w c = 0).
// geq_j = w_j (drho + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u^2)).
__device__ __forceinline__ float cdotf(int j, float ux, float uy, float uz) {
    float r = 0.0f;
    if (cx(j) != 0) r += cx(j) > 0 ? ux : -ux;
    if (cy(j) != 0) r += cy(j) > 0 ? uy : -uy;
    if (cz(j) != 0) r += cz(j) > 0 ? uz : -uz;
    return r;
}

// Collision of one node in the FP32 deviation form: moments, penalty force,
// Guo shift, [Bingham/Smagorinsky], BGK/TRT with Guo forcing. sid: the node's
// particle slot + 1 (0: none) with that slot's record r. Writes the side
// outputs (node force, macroscopic fields, viscosity, free-surface state) and,
// when doOut, the post-collision populations into out[] (which may be g);
// returns whether out was written. bad is set for a non-finite density (the
// caller reports it), so that no global store sits between the loads of the
// nodes of one lane.
template <bool VARVISC, int COLL, bool FS, bool CURVED = false>
__device__ __forceinline__ bool collide32(const float g[19], const Coord& c, uint32_t lw, size_t N, uint32_t sid,
                                          const CoupleRecF& r, const RunSlot<float>& rs, double* __restrict__ visc,
                                          double* __restrict__ mac, const LBKParams& p, const FSView& fsv, bool doOut,
                                          float out[19], bool& bad) {
#pragma clang fp contract(fast)
    float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
    for (int j = 0; j < 19; ++j) {
        drho += g[j];
        if (cx(j) > 0) jx += g[j];
        if (cx(j) < 0) jx -= g[j];
        if (cy(j) > 0) jy += g[j];
        if (cy(j) < 0) jy -= g[j];
        if (cz(j) > 0) jz += g[j];
        if (cz(j) < 0) jz -= g[j];
    }
    const float rho = 1.0f + drho;
    bad = !isfinite(rho);
    const float irho = 1.0f / rho;
    float ux = jx * irho, uy = jy * irho, uz = jz * irho;
    float Tx = (float)p.F[0], Ty = (float)p.F[1], Tz = (float)p.F[2];
    if (p.ffield) {
        Tx += p.ffield[c.idx];
        Ty += p.ffield[N + c.idx];
        Tz += p.ffield[2 * N + c.idx];
    }
    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
    if (sid) {
        const double massD = node_mass<FS, float>(lw, c.idx, fsv, p);
        const float mass = (float)massD;
        // rigid-body velocity at the node: v + w x (node - c)
        const float rx = (float)c.x - r.c[0], ry = (float)c.y - r.c[1], rz = (float)c.z - r.c[2];
        const float lvx = r.v[0] + (r.w[1] * rz - r.w[2] * ry);
        const float lvy = r.v[1] + (r.w[2] * rx - r.w[0] * rz);
        const float lvz = r.v[2] + (r.w[0] * ry - r.w[1] * rx);
        const float lf = mass * irho;
        const float dx = lf * (ux - lvx), dy = lf * (uy - lvy), dz = lf * (uz - lvz);
        hx = -dx; hy = -dy; hz = -dz;
        rs.stage(sid, dx, dy, dz, massD);
        Tx += hx; Ty += hy; Tz += hz;
    }
    ux += 0.5f * Tx * irho;
    uy += 0.5f * Ty * irho;
    uz += 0.5f * Tz * irho;
    if (CURVED && p.macNext) {
        p.macNext[c.idx] = (double)rho;
        p.macNext[N + c.idx] = ux;
        p.macNext[2 * N + c.idx] = uy;
        p.macNext[3 * N + c.idx] = uz;
    }
    if (mac) {
        mac[c.idx] = (double)rho;
        mac[N + c.idx] = ux;
        mac[2 * N + c.idx] = uy;
        mac[3 * N + c.idx] = uz;
        mac[4 * N + c.idx] = hx;
        mac[5 * N + c.idx] = hy;
        mac[6 * N + c.idx] = hz;
    }
    if (!doOut) return false;
    if (FS) {
        ((float*)fsv.rhoCur)[c.idx] = rho;
        if (lw & LINK_INTERFACE) {
            fsv.uS[c.idx] = ux; fsv.uS[N + c.idx] = uy; fsv.uS[2 * N + c.idx] = uz;
            fsv.hFi[c.idx] = hx; fsv.hFi[N + c.idx] = hy; fsv.hFi[2 * N + c.idx] = hz;
        }
    }
    const float usq = ux * ux + uy * uy + uz * uz;
    const float base = 1.5f * usq;
    float nu = (float)p.visc0;
    if (VARVISC) {
        // Bingham / Smagorinsky on the non-equilibrium stress (node.cpp:103-145)
        float s[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy yz zx
#pragma unroll
        for (int j = 1; j < 19; ++j) {
            const float vu = cdotf(j, ux, uy, uz);
            const float geq = (float)wgt(j) * (drho + rho * (3.0f * vu + 4.5f * vu * vu - base));
            const float d = g[j] - geq;
            if (cx(j)) s[0] += d;
            if (cy(j)) s[1] += d;
            if (cz(j)) s[2] += d;
            if (cx(j) * cy(j) > 0) s[3] += d;
            if (cx(j) * cy(j) < 0) s[3] -= d;
            if (cy(j) * cz(j) > 0) s[4] += d;
            if (cy(j) * cz(j) < 0) s[4] -= d;
            if (cz(j) * cx(j) > 0) s[5] += d;
            if (cz(j) * cx(j) < 0) s[5] -= d;
        }
        const double viscOld = visc[c.idx];
        const float tau = 0.5f + 3.0f * (float)viscOld;
        const float k = 1.5f / (tau * rho);
        const float g00 = s[0] * k, g11 = s[1] * k, g22 = s[2] * k, g01 = s[3] * k, g12 = s[4] * k, g20 = s[5] * k;
        const float sr = sqrtf(0.5f * (g00 * g00 + g11 * g11 + g22 * g22 + 2.0f * (g01 * g01 + g20 * g20 + g12 * g12)));
        float nuTurb = p.turbulence ? (float)p.turbConst * sr : 0.0f;
        float nuApp = p.nonNewtonian ? (float)p.plasticVisc + (float)p.yieldStress / (rho * 2.0f * sr) : (float)viscOld;
        if (p.frictional) {
            const float ty = (float)p.frictionFluid * (drho > 0.0f ? drho * (1.0f / 3.0f) : 0.0f);  // rho0 = 1
            nuApp = ty > 0.0f ? (float)p.plasticVisc + ty / (rho * 2.0f * sr) : (float)p.plasticVisc;
        }
        const float a = nuApp + nuTurb;
        const float maxV = (float)((1.8 - 0.5) / 3.0), minV = (float)((0.501 - 0.5) / 3.0);
        const float lo = (a < maxV) ? a : maxV;
        nu = (minV < lo) ? lo : minV;
        visc[c.idx] = (double)nu;
    }
    const float omP = 1.0f / (0.5f + 3.0f * nu);
    const float uT = ux * Tx + uy * Ty + uz * Tz;
    const float kp = 1.0f - 0.5f * omP;
    // rest population: even part only
    {
        const float w0 = (float)wgt(0);
        const float geq0 = w0 * (drho - rho * base);
        out[0] = g[0] - omP * (g[0] - geq0) + kp * (w0 * (-3.0f * uT));
    }
    const float omM = (COLL == HYB_COLLISION_TRT) ? 1.0f / (0.5f + (float)p.trtMagic / (1.0f / omP - 0.5f)) : omP;
    const float km = 1.0f - 0.5f * omM;
#pragma unroll
    for (int j = 1; j < 19; j += 2) {
        // j and opp(j) = j + 1 form a pair; with vu = c_j.u, geq_j,o = w (drho + rho (4.5 vu^2 - 1.5 u^2 +- 3 vu))
        const int o = j + 1;
        const float w = (float)wgt(j);
        const float vu = cdotf(j, ux, uy, uz);
        const float cF = cdotf(j, Tx, Ty, Tz);
        const float ep = w * (drho + rho * (4.5f * vu * vu - base));
        const float em = w * rho * 3.0f * vu;
        const float fpl = 0.5f * (g[j] + g[o]), fmi = 0.5f * (g[j] - g[o]);
        const float sp = w * (9.0f * vu * cF - 3.0f * uT);
        const float sm = w * (3.0f * cF);
        const float relP = omP * (fpl - ep) - kp * sp;
        const float relM = omM * (fmi - em) - km * sm;
        out[j] = g[j] - relP - relM;
        out[o] = g[o] - relP + relM;
    }
    return true;
}


template <bool PULL, bool VARVISC, int COLL, bool FS, bool CURVED = false>
__device__ __forceinline__ void fused32_node(const float* __restrict__ fin, float* __restrict__ fout,
                                             const uint32_t* __restrict__ link, const uint32_t* __restrict__ solid,
                                             const CoupleRecF* __restrict__ cp, const RunSlot<float>& rs,
                                             double* __restrict__ visc, double* __restrict__ mac, const LBKParams& p,
                                             const FSView& fsv) {
    Coord c;
    if (!node_coord(p, c)) return;
    const size_t N = (size_t)p.X * p.Y * p.Z;
    const uint32_t lw = link[c.idx];
    if (!(lw & LINK_ACTIVE)) return;
    float g[19];
    if (FS && (lw & LINK_PRESTREAMED)) load_prepared<float>(fsv, p.PS, c.idx, lw, g);
    else if (PULL) pull19<float, CURVED>(fin, p.PS, p, c, lw, g);
    else {
#pragma unroll
        for (int j = 0; j < 19; ++j) g[j] = ld_pop(&fin[(size_t)j * p.PS + c.idx]);
    }
    const uint32_t sid = owner_slot(solid ? solid[c.idx] : 0u, p.ctag);
    CoupleRecF r{};
    if (sid) r = cp[sid - 1];
    float out[19];
    bool bad;
    const bool wrote = collide32<VARVISC, COLL, FS, CURVED>(g, c, lw, N, sid, r, rs, visc, mac, p, fsv,
                                                           fout != nullptr, out, bad);
    if (bad && p.herr) *p.herr = LB_ERR_NUMERIC;
    if (!wrote) return;
#pragma unroll
    for (int j = 0; j < 19; ++j) st_pop(&fout[(size_t)j * p.PS + c.idx], out[j]);
}

template <bool PULL, bool VARVISC, int COLL, bool FS, bool CURVED = false>
__global__ void __launch_bounds__(BX * BY, (VARVISC || CURVED) ? 1 : 8) k_fused32(const float* __restrict__ fin, float* __restrict__ fout,
                                                     const uint32_t* __restrict__ link,
                                                     const uint32_t* __restrict__ solid,
                                                     const CoupleRecF* __restrict__ cp, const CoupleRec* __restrict__ cp64,
                                                     HydroSink hs, double* __restrict__ visc,
                                                     double* __restrict__ mac, LBKParams p, FSView fsv) {
    __shared__ RunLDS<float> lds;
    const bool sums = hs.segs != nullptr;
    const RunSlot<float> rs{sums ? &lds : nullptr, (int)threadIdx.x};
    rs.clear();
    fused32_node<PULL, VARVISC, COLL, FS, CURVED>(fin, fout, link, solid, cp, rs, visc, mac, p, fsv);
    if (sums) hydro_runs<float, 1>(lds, hs, cp64, p);
}

// Standalone pull streaming: turns the post-collision state into the
// reference's step-end (post-stream) populations, for host views only.
template <typename T, bool CURVED = false>
__global__ void __launch_bounds__(BX * BY) k_stream(const T* __restrict__ fin, T* __restrict__ fout,
                                                    const uint32_t* __restrict__ link, LBKParams p) {
    Coord c;
    if (!node_coord(p, c)) return;
    const size_t N = (size_t)p.X * p.Y * p.Z;
    const uint32_t lw = link[c.idx];
    if (!(lw & LINK_ACTIVE)) return;
    T f[19];
    pull19<T, CURVED>(fin, p.PS, p, c, lw, f);
#pragma unroll
    for (int j = 0; j < 19; ++j) fout[(size_t)j * p.PS + c.idx] = f[j];
}

// ------------------------------------------------------------------ initialisation
// Boundary typing (LB.cpp:394-431), plane walls (LB.cpp:497-533) and the
// problem's gas region (LB.cpp:611-784: SPLASH plane, DRUM and NET shapes, or a
// liquid box).
__global__ void k_init_types(uint8_t* __restrict__ type, LBKParams p, int b0, int b1, int b2, int b3, int b4, int b5,
                             int nwalls, const WallDev* __restrict__ walls, int nobj, const ObjDev* __restrict__ objs,
                             CylSet cs, int problem, double lx0, double lx1, double ly0, double ly1, double lz0,
                             double lz1, double avalanchePosit) {
    Coord c;
    if (!node_coord(p, c)) return;
    uint8_t t = T_FLUID;
    auto isWall = [](uint8_t tt) { return tt == 5 || tt == 6 || tt == 7 || tt == 8 || tt == 9; };
    if (c.x == 0 && !isWall(t)) t = (uint8_t)b0;
    else if (c.x == p.X - 1 && !isWall(t)) t = (uint8_t)b1;
    if (c.y == 0 && !isWall(t)) t = (uint8_t)b2;
    else if (c.y == p.Y - 1 && !isWall(t)) t = (uint8_t)b3;
    if (c.z == 0 && !isWall(t)) t = (uint8_t)b4;
    else if (c.z == p.Z - 1 && !isWall(t)) t = (uint8_t)b5;
    const V3 pos((double)c.x, (double)c.y, (double)c.z);
    for (int w = 0; w < nwalls; ++w) {
        const V3 d = pos - walls[w].p / p.unitLength;
        if (walls[w].n.dot(d) < 0.0) {
            if (walls[w].slip) t = walls[w].moving ? 6 : 5;
            else t = walls[w].moving ? 8 : 7;
        }
    }
    // nodes outside a cylinder are curved walls (LB.cpp:552-587)
    for (int k = 0; k < cs.n; ++k)
        if (!CylDev::inside(pos, cs.c[k].p1 / p.unitLength, cs.c[k].naxes, cs.c[k].R / p.unitLength)) t = T_CURVED;
    // objects are static walls (LB.cpp:535-550)
    for (int o = 0; o < nobj; ++o) {
        const V3 d = pos - objs[o].x0 / p.unitLength;
        const double rr = objs[o].r / p.unitLength;
        if (d.norm2() < rr * rr) t = T_WALL;
    }
    if (t == T_FLUID) {
        if (problem == HYB_PROBLEM_SPLASH) {
            if ((pos - V3(0.0, 0.0, 20.0)).dot(V3(0.0, 0.0, 1.0)) > 0.0) t = T_GAS;
        } else if (problem == HYB_PROBLEM_LIQUID_BOX) {
            if (pos.x < lx0 || pos.x > lx1 || pos.y < ly0 || pos.y > ly1 || pos.z < lz0 || pos.z > lz1) t = T_GAS;
        } else if (problem == HYB_PROBLEM_DRUM) {
            // gas outside the cylinder through (lx0, lx1, ly0) along y with radius ly1
            // (all in lattice units, LB.cpp:734-747)
            if (!CylDev::inside(pos, V3(lx0, lx1, ly0), V3(0.0, 1.0, 0.0), ly1)) t = T_GAS;
        } else if (problem == HYB_PROBLEM_NET) {
            // box before the net, then the plane through (xMin, 0, 0) with normal
            // (-1, 0, 1) (LB.cpp:662-690); the bounds are unsigned lattice indices
            const unsigned xMin = (unsigned)int(avalanchePosit / p.unitLength), xMax = (unsigned)p.X;
            const unsigned yMin = 0, yMax = (unsigned)p.Y, zMin = 0, zMax = (unsigned)int(4.0 / p.unitLength);
            const unsigned gx = (unsigned)c.x, gy = (unsigned)c.y, gz = (unsigned)c.z;
            if (gx < xMin || gx > xMax || gy < yMin || gy > yMax || gz < zMin || gz > zMax) t = T_GAS;
            if (t == T_FLUID) {
                const V3 pp((double)int(avalanchePosit / p.unitLength), 0.0, 0.0), nn(-1.0, 0.0, 1.0);
                if ((pos - pp).dot(nn) > 0.0) t = T_GAS;
            }
        }
    }
    type[c.idx] = t;
}

// Interface = fluid with a gas neighbour (LB.cpp:787-797). Each pass reads the
// previous pass's types, which is equivalent to the reference's in-place loops
// because neither loop changes the property it tests on its neighbours.
__global__ void k_tag_interface(const uint8_t* __restrict__ tin, uint8_t* __restrict__ tout, LBKParams p) {
    Coord c;
    if (!node_coord(p, c)) return;
    uint8_t t = tin[c.idx];
    if (t == T_FLUID && interior(p, c)) {
        for (int j = 1; j < 19; ++j)
            if (tin[nb_index(p, c, j)] == T_GAS) t = T_INTERFACE;
    }
    tout[c.idx] = t;
}

// Isolated interface nodes become gas (LB.cpp:799-814).
__global__ void k_remove_isolated(const uint8_t* __restrict__ tin, uint8_t* __restrict__ tout, LBKParams p) {
    Coord c;
    if (!node_coord(p, c)) return;
    uint8_t t = tin[c.idx];
    if (t == T_INTERFACE && interior(p, c)) {
        bool close = false;
        for (int j = 1; j < 19; ++j)
            if (tin[nb_index(p, c, j)] == T_FLUID) { close = true; break; }
        if (!close) t = T_GAS;
    }
    tout[c.idx] = t;
}

__global__ void k_init_links(const uint8_t* __restrict__ type, uint32_t* __restrict__ link, uint32_t* __restrict__ slipw,
                             uint32_t* __restrict__ curvw, LBKParams p) {
    Coord c;
    if (!node_coord(p, c)) return;
    const uint8_t t = type[c.idx];
    uint32_t lw = 0, sw = 0, cw = 0;
    if ((t == T_FLUID || t == T_INTERFACE) && interior(p, c)) {
        lw |= LINK_ACTIVE;
        if (t == T_INTERFACE) lw |= LINK_INTERFACE;
        for (int j = 1; j < 19; ++j) {
            const uint8_t tl = type[nb_index(p, c, j)];
            if (is_plane_wall(tl) || tl == T_CURVED) lw |= (1u << j);
            if (is_slip_wall(tl)) sw |= (1u << j);
            if (tl == T_CURVED) cw |= (1u << j);
        }
        if (sw) lw |= LINK_SLIP;
        if (cw) lw |= LINK_CURVED;
    }
    link[c.idx] = lw;
    if (slipw) slipw[c.idx] = sw;
    if (curvw) curvw[c.idx] = cw;
}

// Largest coordinates of active nodes (LB.cpp:913-921). Wavefront maxima are
// combined per z-plane (mx[3 * z + k]); the host takes the maximum over planes.
__global__ void k_max_active(const uint32_t* __restrict__ link, int* __restrict__ mx, LBKParams p) {
    Coord c;
    int v[3] = {-1, -1, -1};
    if (node_coord(p, c) && (link[c.idx] & LINK_ACTIVE)) { v[0] = c.x; v[1] = c.y; v[2] = c.z; }
    for (int k = 0; k < 3; ++k) {
        int m = v[k];
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(m, off, 64);
            m = o > m ? o : m;
        }
        if ((threadIdx.x & 63) == 0 && m >= 0) atomicMax(&mx[3 * blockIdx.z + k], m);
    }
}

// node::initialize + setEquilibrium (node.cpp:26-61) with the hydrostatic
// density of LB.cpp:917-942. Interface nodes start at rho = 1.
template <typename T, bool SHIFT>
__global__ void k_init_pops(T* __restrict__ f, const uint32_t* __restrict__ link, const int* __restrict__ mx,
                            double* __restrict__ visc, LBKParams p, double u0x, double u0y, double u0z, double Fx,
                            double Fy, double Fz, T* __restrict__ rho, double* __restrict__ mass,
                            double* __restrict__ uS) {
#pragma clang fp contract(off)
    Coord c;
    if (!node_coord(p, c)) return;
    const size_t N = (size_t)p.X * p.Y * p.Z;
    const uint32_t lw = link[c.idx];
    if (!(lw & LINK_ACTIVE)) return;
    // the cfg body force, before collision() zeroes it when forceFieldSolve = 0 (LB.cpp:1074-1076)
    const V3 F(Fx, Fy, Fz);
    double n;
    if (lw & LINK_INTERFACE) {
        n = p.initDensity;
    } else {
        const V3 dh = V3((double)c.x, (double)c.y, (double)c.z) - V3((double)mx[0], (double)mx[1], (double)mx[2]);
        n = p.initDensity + 3.0 * 1.0 * 1.0 * p.initDensity * dh.dot(F);
    }
    const V3 u = (1.0 * F) / 2.0 / n + V3(u0x, u0y, u0z);
    const double usq = u.norm2();
    for (int j = 0; j < 19; ++j) {
        const V3 cj((double)cx(j), (double)cy(j), (double)cz(j));
        const double vu = u.dot(cj);
        const double feq = wgt(j) * n * (1.0 + 3.0 * vu + 4.5 * vu * vu - 1.5 * usq);
        if (SHIFT) f[(size_t)j * p.PS + c.idx] = (T)(feq - wgt(j));
        else f[(size_t)j * p.PS + c.idx] = (T)feq;
    }
    if (visc) visc[c.idx] = p.visc0;
    if (rho) {
        // free surface: density of the "previous collision" and the interface state
        rho[c.idx] = (T)n;
        if (lw & LINK_INTERFACE) {
            mass[c.idx] = 0.5 * p.initDensity;
            uS[c.idx] = u.x; uS[N + c.idx] = u.y; uS[2 * N + c.idx] = u.z;
        }
    }
}

__global__ void k_count_active(const uint32_t* __restrict__ link, unsigned long long* __restrict__ cnt, LBKParams p) {
    Coord c;
    bool a = false;
    if (node_coord(p, c)) a = (link[c.idx] & LINK_ACTIVE) != 0;
    const unsigned long long b = __ballot(a);
    // one counter per z-plane; the host adds the planes
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&cnt[blockIdx.z], (unsigned long long)__popcll(b));
}

// Active nodes outside particles, fluid (cnt[0]) and interface (cnt[1]): the
// initial liquid mass of LB.cpp:209-216 in node units.
__global__ void k_count_mass(const uint32_t* __restrict__ link, const uint32_t* __restrict__ solid,
                             unsigned long long* __restrict__ cnt, LBKParams p) {
    Coord c;
    bool fl = false, in = false;
    if (node_coord(p, c)) {
        const uint32_t lw = link[c.idx];
        const bool owned = solid && owner_slot(solid[c.idx], p.ctag) != 0u;
        fl = (lw & LINK_ACTIVE) && !(lw & LINK_INTERFACE) && !owned;
        in = (lw & LINK_ACTIVE) && (lw & LINK_INTERFACE) && !owned;
    }
    const unsigned long long bf = __ballot(fl), bi = __ballot(in);
    if ((threadIdx.x & 63) == 0) {
        if (bf) atomicAdd(&cnt[0], (unsigned long long)__popcll(bf));
        if (bi) atomicAdd(&cnt[1], (unsigned long long)__popcll(bi));
    }
}

// ------------------------------------------------------------------ particle classification
// Each particle slot (real particles, then ghosts) marks the active nodes
// strictly inside it (vector.cpp:153-158) with the current coupling-step tag
// (link == nullptr: every node inside, active or not).
// Where spheres overlap the highest slot index wins, as in the full re-test of
// LB.cpp:475-495.
__device__ __forceinline__ NodeBox slot_box(const CoupleRec& r, const LBKParams& p) {
    const double c[3] = {r.x0[0] / p.unitLength, r.x0[1] / p.unitLength, r.x0[2] / p.unitLength};
    const double rad = r.r / p.unitLength;
    const int lim[3] = {p.X, p.Y, p.Z};
    NodeBox b;
    for (int k = 0; k < 3; ++k) {
        int lo = (int)floor(c[k] - rad);
        int hi = (int)ceil(c[k] + rad);
        lo = lo < 1 ? 1 : lo;
        hi = hi > lim[k] - 2 ? lim[k] - 2 : hi;
        b.lo[k] = lo;
        b.hi[k] = hi;
    }
    return b;
}

// One wavefront per particle slot in the classification and force-sum kernels;
// each wavefront keeps its row table in LDS (SLOT_WAVES wavefronts per workgroup).
constexpr int SLOT_WAVES = 4;
// nodes per lane whose loads are issued together in the slot kernels
#ifndef HYB_SLOT_BATCH
#define HYB_SLOT_BATCH 4
#endif
constexpr int CH = HYB_SLOT_BATCH;
// wavefronts per particle slot in the classification (a sphere of radius 6 spans 196 rows: four batches)
constexpr int CLASSIFY_PARTS = 4;

struct SlotRows {  // LDS of one wavefront: inclusive node counts and first x of 64 rows
    int scan[64];
    int xa[64];
};

// Hands the nodes of a slot's box strictly inside its sphere (vector.cpp:153-158)
// to the lanes of one wavefront, CH nodes per lane at a time (ok[k] marks the
// valid entries): the caller issues the loads of a batch before using any of
// them. Rows are taken 64 at a time; each row contributes the x interval that
// can hold inside nodes (from the circle radius with one node of margin on each
// side, every candidate is then tested exactly), and a wavefront prefix sum of
// the interval lengths spreads the candidates evenly over the lanes, so that the
// lanes of one access hold consecutive nodes of few rows (few cache lines per
// instruction). A row whose (y, z) distance already reaches the radius holds no
// inside node: rounded sums of non-negative squares are monotone.
// part / parts: this wavefront takes the 64-row batches part, part + parts, ...
template <class F>
__device__ __forceinline__ void for_slot_batches(const NodeBox& b, int lane, SlotRows& sh, const V3& cen, double rad,
                                                 const LBKParams& p, int part, int parts, F&& f) {
#pragma clang fp contract(off)
    const int nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1, nz = b.hi[2] - b.lo[2] + 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    const int rows = ny * nz;
    const double r2 = rad * rad;
    for (int base = 64 * part; base < rows; base += 64 * parts) {
        const int r = base + lane;
        int xa = 0, cnt = 0;
        if (r < rows) {
            const double dy = (double)(b.lo[1] + r % ny) - cen.y, dz = (double)(b.lo[2] + r / ny) - cen.z;
            const double s2 = dy * dy + dz * dz;
            if (s2 < r2) {
                const double w = sqrt(r2 - s2);
                const int a = max(b.lo[0], (int)ceil(cen.x - w) - 1);
                const int c = min(b.hi[0], (int)floor(cen.x + w) + 1);
                if (c >= a) {
                    xa = a;
                    cnt = c - a + 1;
                }
            }
        }
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        const int total = __shfl(incl, 63, 64);
        sh.scan[lane] = incl;
        sh.xa[lane] = xa;
        __builtin_amdgcn_wave_barrier();
        for (int t0 = 0; t0 < total; t0 += 64 * CH) {
            size_t idx[CH];
            int xs[CH], ys[CH], zs[CH];
            bool ok[CH];
#pragma unroll
            for (int k = 0; k < CH; ++k) {
                const int t = t0 + k * 64 + lane;
                ok[k] = false;
                idx[k] = 0;
                xs[k] = ys[k] = zs[k] = 0;
                if (t < total) {
                    int lo = 0, hi = 63;  // first row whose inclusive count exceeds t
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (sh.scan[mid] > t) hi = mid;
                        else lo = mid + 1;
                    }
                    const int row = base + lo;
                    const int x = sh.xa[lo] + (t - (lo ? sh.scan[lo - 1] : 0));
                    const int y = b.lo[1] + row % ny, z = b.lo[2] + row / ny;
                    const V3 d = V3((double)x, (double)y, (double)z) - cen;
                    ok[k] = d.norm2() < r2;
                    xs[k] = x;
                    ys[k] = y;
                    zs[k] = z;
                    idx[k] = (size_t)x + (size_t)p.X * ((size_t)y + (size_t)p.Y * (size_t)z);
                }
            }
            f(idx, xs, ys, zs, ok);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void __launch_bounds__(256) k_classify_set(const int* __restrict__ nslots_p, const CoupleRec* __restrict__ cp,
                               NodeBox* __restrict__ boxes, int* __restrict__ nboxes, uint32_t* __restrict__ solid,
                               const uint32_t* __restrict__ link, LBKParams p) {
#pragma clang fp contract(off)
    __shared__ SlotRows shr[SLOT_WAVES];
    const int ns = *nslots_p;
    if (blockIdx.x == 0 && threadIdx.x == 0) *nboxes = ns;
    const int lane = (int)(threadIdx.x & 63);
    const int unit = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nunits = (int)((gridDim.x * blockDim.x) >> 6);
    SlotRows& sh = shr[threadIdx.x >> 6];
    // CLASSIFY_PARTS wavefronts per slot, each on its own 64-row batches of the slot's box, so that the
    // longest chain of one wavefront is one batch (the marks are an order-free maximum)
    for (int us = unit; us < ns * CLASSIFY_PARTS; us += nunits) {
        const int s = us / CLASSIFY_PARTS, part = us % CLASSIFY_PARTS;
        const CoupleRec r = cp[s];
        const NodeBox b = slot_box(r, p);
        if (lane == 0 && part == 0) boxes[s] = b;
        const V3 cen(r.x0[0] / p.unitLength, r.x0[1] / p.unitLength, r.x0[2] / p.unitLength);
        const double rad = r.r / p.unitLength;
        const uint32_t val = (p.ctag << SOLID_SLOT_BITS) | (uint32_t)(s + 1);
        for_slot_batches(b, lane, sh, cen, rad, p, part, CLASSIFY_PARTS,
                         [&](const size_t* idx, const int*, const int*, const int*, const bool* ok) {
            uint32_t lk[CH];
#pragma unroll
            for (int k = 0; k < CH; ++k) lk[k] = ok[k] ? (link ? link[idx[k]] : LINK_ACTIVE) : 0u;
#pragma unroll
            for (int k = 0; k < CH; ++k)
                if (ok[k] && (lk[k] & LINK_ACTIVE)) atomicMax(&solid[idx[k]], val);
        });
    }
}

// Hydrodynamic force on the moving no-slip walls (LB.cpp:1321-1341): every link
// of an active node to a type-8 wall node adds bounceBackForce (node.cpp:283-291),
// (2 (fs_j - rho0 w_j) - BBi) c_j with BBi = 0 (box walls have zero velocity), to
// the wall whose plane holds that node (the last such wall, LB.cpp:497-533).
// fs is the post-collision state of the last step. Exact fixed-point sums:
// acc[(wall * 3 + k) * 4 ...].
template <typename T, bool SHIFT>
__global__ void __launch_bounds__(BX * BY) k_wall_fhydro(const T* __restrict__ f, const uint8_t* __restrict__ type,
                                                          WallSet ws, unsigned long long* __restrict__ acc,
                                                          LBKParams p) {
#pragma clang fp contract(off)
    Coord c;
    if (!node_coord(p, c)) return;
    const uint8_t t = type[c.idx];
    if (t != T_FLUID && t != T_INTERFACE) return;
    for (int j = 1; j < 19; ++j) {
        const int nx = c.x + cx(j), ny = c.y + cy(j), nz = c.z + cz(j);
        if (nx < 0 || ny < 0 || nz < 0 || nx >= p.X || ny >= p.Y || nz >= p.Z) continue;
        const size_t n = (size_t)nx + (size_t)p.X * ((size_t)ny + (size_t)p.Y * (size_t)nz);
        if (type[n] != T_WALL_MOVING) continue;
        const V3 pos((double)nx, (double)ny, (double)nz);
        int wi = -1;
        for (int k = 0; k < ws.n; ++k) {
            const V3 wp = ws.w[k].p / ws.unitLength;
            if (ws.w[k].n.dot(pos - wp) < 0.0) wi = k;
        }
        if (wi < 0) continue;
        const double fj = (double)f[(size_t)j * p.PS + c.idx];
        const double d = SHIFT ? fj + wgt(j) * (1.0 - p.initDensity) : fj - p.initDensity * wgt(j);
        const double s = (2.0 * d - 0.0) * 1.0;
        if (cx(j)) exact_add(&acc[(wi * 3 + 0) * 4], s * (double)cx(j));
        if (cy(j)) exact_add(&acc[(wi * 3 + 1) * 4], s * (double)cy(j));
        if (cz(j)) exact_add(&acc[(wi * 3 + 2) * 4], s * (double)cz(j));
    }
}

// Status-line reductions over the active nodes (IO.cpp:835-850, 969-985):
// out[0] = max |u|^2 as the bits of a non-negative double (NaN skipped, as
// std::max keeps its first argument), out[1] = active nodes, out[2] = active
// nodes with viscosity > viscMin. Order-free, so exact.
__global__ void __launch_bounds__(256) k_status(const double* __restrict__ mac, const double* __restrict__ visc,
                                                const uint32_t* __restrict__ link, size_t N, double viscMin,
                                                unsigned long long* __restrict__ out) {
#pragma clang fp contract(off)
    double m = 0.0;
    unsigned long long act = 0, hi = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
        if (!(link[i] & LINK_ACTIVE)) continue;
        ++act;
        if (mac) {
            const double ux = mac[N + i], uy = mac[2 * N + i], uz = mac[3 * N + i];
            const double u2 = ux * ux + uy * uy + uz * uz;
            if (u2 > m) m = u2;
        }
        if (visc && visc[i] > viscMin) ++hi;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
        act += __shfl_xor(act, off, 64);
        hi += __shfl_xor(hi, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&out[0], (unsigned long long)__double_as_longlong(m));
        atomicAdd(&out[1], act);
        atomicAdd(&out[2], hi);
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ a, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

dim3 lattice_grid(const LBKParams& p) { return dim3((p.X + BX - 1) / BX, (p.Y + BY - 1) / BY, p.Z); }

}  // namespace

// ------------------------------------------------------------------ launchers
#define HYB_FUSED_LAUNCH(PULL, VV, CO, CUR)                                                                        \
    if (pull == PULL && varVisc == VV && collision == CO) {                                                        \
        if (prec == Prec::F64) {                                                                                   \
            if (fs)                                                                                                \
                k_fused64<PULL, VV, CO, true, CUR><<<g, b, 0, s>>>((const double*)fin, (double*)fout, link, solid, \
                                                                   cp, sink, visc, mac, p, *fs);                   \
            else                                                                                                   \
                k_fused64<PULL, VV, CO, false, CUR><<<g, b, 0, s>>>((const double*)fin, (double*)fout, link,       \
                                                                    solid, cp, sink, visc, mac, p, none);          \
        } else {                                                                                                   \
            if (fs)                                                                                                \
                k_fused32<PULL, VV, CO, true, CUR><<<g, b, 0, s>>>((const float*)fin, (float*)fout, link, solid,   \
                                                                   cpf, cp, sink, visc, mac, p, *fs);              \
            else                                                                                                   \
                k_fused32<PULL, VV, CO, false, CUR><<<g, b, 0, s>>>((const float*)fin, (float*)fout, link, solid,  \
                                                                    cpf, cp, sink, visc, mac, p, none);            \
        }                                                                                                          \
        return;                                                                                                    \
    }
// Cylinders (p.curvw set, CUR = true): the curved-wall pull and the capture of
// every active node's density and velocity for the next step's curved links.
#define HYB_FUSED_CASE(PULL, VV, CO) HYB_FUSED_LAUNCH(PULL, VV, CO, false)
#define HYB_CURVED_CASE(PULL, VV, CO) HYB_FUSED_LAUNCH(PULL, VV, CO, true)

void launch_lb_fused(Prec prec, bool pull, bool varVisc, int collision, const void* fin, void* fout,
                     const uint32_t* link, const uint32_t* solid, const CoupleRec* cp, const CoupleRecF* cpf,
                     const HydroSink* hs, double* visc, double* mac, const FSView* fs, const LBKParams& p,
                     hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    const FSView none{};
    const HydroSink sink = (hs && solid) ? *hs : HydroSink{nullptr, nullptr, 0u};
    if (p.curvw) {
        HYB_CURVED_CASE(true, false, HYB_COLLISION_BGK)
        HYB_CURVED_CASE(true, false, HYB_COLLISION_TRT)
        HYB_CURVED_CASE(true, true, HYB_COLLISION_BGK)
        HYB_CURVED_CASE(true, true, HYB_COLLISION_TRT)
        HYB_CURVED_CASE(false, false, HYB_COLLISION_BGK)
        HYB_CURVED_CASE(false, false, HYB_COLLISION_TRT)
        HYB_CURVED_CASE(false, true, HYB_COLLISION_BGK)
        HYB_CURVED_CASE(false, true, HYB_COLLISION_TRT)
    }
    HYB_FUSED_CASE(true, false, HYB_COLLISION_BGK)
    HYB_FUSED_CASE(true, false, HYB_COLLISION_TRT)
    HYB_FUSED_CASE(true, true, HYB_COLLISION_BGK)
    HYB_FUSED_CASE(true, true, HYB_COLLISION_TRT)
    HYB_FUSED_CASE(false, false, HYB_COLLISION_BGK)
    HYB_FUSED_CASE(false, false, HYB_COLLISION_TRT)
    HYB_FUSED_CASE(false, true, HYB_COLLISION_BGK)
    HYB_FUSED_CASE(false, true, HYB_COLLISION_TRT)
}

void launch_wall_fhydro(Prec prec, const void* f, const uint8_t* type, const WallSet& ws, unsigned long long* acc,
                        const LBKParams& p, hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    if (prec == Prec::F64) k_wall_fhydro<double, false><<<g, b, 0, s>>>((const double*)f, type, ws, acc, p);
    else k_wall_fhydro<float, true><<<g, b, 0, s>>>((const float*)f, type, ws, acc, p);
}

namespace {
typedef float F4 __attribute__((ext_vector_type(4)));
// Plain non-temporal copy with `T`-sized accesses per lane: the reference point
// for what HBM delivers to a streaming kernel at that width. Each lane issues
// COPY_UNROLL loads (coalesced across the workgroup) before their stores, so
// that enough bytes are in flight to cover the memory latency.
constexpr int COPY_UNROLL = 8;
template <typename T>
__global__ void __launch_bounds__(256) k_copy_nt(const T* __restrict__ src, T* __restrict__ dst, size_t n) {
    const size_t tile = (size_t)blockDim.x * COPY_UNROLL;
    for (size_t base = (size_t)blockIdx.x * tile; base < n; base += (size_t)gridDim.x * tile) {
        T v[COPY_UNROLL];
#pragma unroll
        for (int k = 0; k < COPY_UNROLL; ++k) {
            const size_t i = base + (size_t)k * blockDim.x + threadIdx.x;
            if (i < n) v[k] = __builtin_nontemporal_load(&src[i]);
        }
#pragma unroll
        for (int k = 0; k < COPY_UNROLL; ++k) {
            const size_t i = base + (size_t)k * blockDim.x + threadIdx.x;
            if (i < n) __builtin_nontemporal_store(v[k], &dst[i]);
        }
    }
}
}  // namespace

bool launch_copy_probe(int width, const void* src, void* dst, size_t bytes, hipStream_t s) {
    const unsigned grid = 256 * 8;  // 8 workgroups of 256 lanes per CU
    switch (width) {
        case 4: k_copy_nt<float><<<grid, 256, 0, s>>>((const float*)src, (float*)dst, bytes / 4); return true;
        case 8: k_copy_nt<double><<<grid, 256, 0, s>>>((const double*)src, (double*)dst, bytes / 8); return true;
        case 16: k_copy_nt<F4><<<grid, 256, 0, s>>>((const F4*)src, (F4*)dst, bytes / 16); return true;
        default: return false;
    }
}

namespace {
// out[k] = value of the exact sum acc[4k..4k+3] times scale, k < n (acc == nullptr: zeros)
__global__ void k_exact_finalize(const unsigned long long* __restrict__ acc, int n, double scale,
                                 double* __restrict__ out) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < n) out[k] = acc ? exact_value(acc + 4 * (size_t)k) * scale : 0.0;
}
}  // namespace

void launch_exact_finalize(const unsigned long long* acc, int n, double scale, double* out, hipStream_t s) {
    if (n <= 0) return;
    k_exact_finalize<<<(n + 63) / 64, 64, 0, s>>>(acc, n, scale, out);
}

void launch_status(const double* mac, const double* visc, const uint32_t* link, size_t N, double viscMin,
                   unsigned long long* out, hipStream_t s) {
    k_status<<<1024, 256, 0, s>>>(mac, visc, link, N, viscMin, out);
}

void launch_lb_stream(Prec prec, const void* fin, void* fout, const uint32_t* link, const LBKParams& p, hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    if (p.curvw) {
        if (prec == Prec::F64) k_stream<double, true><<<g, b, 0, s>>>((const double*)fin, (double*)fout, link, p);
        else k_stream<float, true><<<g, b, 0, s>>>((const float*)fin, (float*)fout, link, p);
        return;
    }
    if (prec == Prec::F64) k_stream<double><<<g, b, 0, s>>>((const double*)fin, (double*)fout, link, p);
    else k_stream<float><<<g, b, 0, s>>>((const float*)fin, (float*)fout, link, p);
}

void launch_lb_init_types(uint8_t* type, const int* bnd, int nwalls, const WallDev* walls, int nobj,
                          const ObjDev* objs, const CylSet& cs, int problem, const double* lbx, const LBKParams& p,
                          hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    k_init_types<<<g, b, 0, s>>>(type, p, bnd[0], bnd[1], bnd[2], bnd[3], bnd[4], bnd[5], nwalls, walls, nobj, objs,
                                 cs, problem, lbx[0], lbx[1], lbx[2], lbx[3], lbx[4], lbx[5], lbx[6]);
}

void launch_lb_init_interface(uint8_t* type, uint8_t* tmp, const LBKParams& p, hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    k_tag_interface<<<g, b, 0, s>>>(type, tmp, p);
    k_remove_isolated<<<g, b, 0, s>>>(tmp, type, p);
}

void launch_lb_init_links(const uint8_t* type, uint32_t* link, uint32_t* slipw, uint32_t* curvw, const LBKParams& p,
                          hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    k_init_links<<<g, b, 0, s>>>(type, link, slipw, curvw, p);
}

void launch_lb_max_active(const uint32_t* link, int* mx, const LBKParams& p, hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    k_max_active<<<g, b, 0, s>>>(link, mx, p);
}

void launch_lb_init_pops(Prec prec, void* f, const uint32_t* link, const int* mx, double* visc, const double* u0,
                         const double* Finit, void* rho, double* mass, double* uS, const LBKParams& p, hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    if (prec == Prec::F64)
        k_init_pops<double, false><<<g, b, 0, s>>>((double*)f, link, mx, visc, p, u0[0], u0[1], u0[2], Finit[0],
                                                   Finit[1], Finit[2], (double*)rho, mass, uS);
    else
        k_init_pops<float, true><<<g, b, 0, s>>>((float*)f, link, mx, visc, p, u0[0], u0[1], u0[2], Finit[0], Finit[1],
                                                 Finit[2], (float*)rho, mass, uS);
}

void launch_classify_set(const int* nslots, int max_slots, const CoupleRec* cp, NodeBox* boxes, int* nboxes,
                         uint32_t* solid, const uint32_t* link, const LBKParams& p, hipStream_t s) {
    if (max_slots <= 0) return;
    int grid = (max_slots * CLASSIFY_PARTS + SLOT_WAVES - 1) / SLOT_WAVES;  // CLASSIFY_PARTS wavefronts per slot
    if (grid > 16384) grid = 16384;
    k_classify_set<<<grid, 64 * SLOT_WAVES, 0, s>>>(nslots, cp, boxes, nboxes, solid, link, p);
}


void launch_fill_u32(uint32_t* a, uint32_t v, size_t n, hipStream_t s) {
    if (!n) return;
    k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, v, n);
}

void launch_count_mass(const uint32_t* link, const uint32_t* solid, unsigned long long* count, const LBKParams& p,
                       hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    k_count_mass<<<g, b, 0, s>>>(link, solid, count, p);
}

void launch_count_active(const uint32_t* link, unsigned long long* count, const LBKParams& p, hipStream_t s) {
    const dim3 g = lattice_grid(p), b(BX, BY, 1);
    k_count_active<<<g, b, 0, s>>>(link, count, p);
}

}  // namespace hyb
