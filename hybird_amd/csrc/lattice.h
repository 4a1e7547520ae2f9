// Lattice indexing shared by the lattice-Boltzmann and free-surface kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.h"
#include "lb_kernels.h"

namespace hyb {

// Node coordinates and linear index x + X (y + Y z) (LB.cpp:42-83).
struct Coord {
    int x, y, z;
    size_t idx;
};

__device__ __forceinline__ Coord coord_of(const LBKParams& p, size_t idx) {
    Coord c;
    c.idx = idx;
    c.x = (int)(idx % (size_t)p.X);
    const size_t r = idx / (size_t)p.X;
    c.y = (int)(r % (size_t)p.Y);
    c.z = (int)(r / (size_t)p.Y);
    return c;
}

// Neighbour coordinates x-1, x, x+1 along one axis for an interior node, with the
// periodic rewiring of LB.cpp:437-472 (layers 0 and X-1 are ghosts, period X-2).
struct Nb3 {
    int m, c, pl;
};
__device__ __forceinline__ Nb3 nb3(int x, int X, int per) {
    Nb3 r;
    r.c = x;
    r.m = (per && x == 1) ? X - 2 : x - 1;
    r.pl = (per && x == X - 2) ? 1 : x + 1;
    return r;
}
__device__ __forceinline__ int pick(const Nb3& n, int d) { return d > 0 ? n.pl : (d < 0 ? n.m : n.c); }

// Neighbour index of an interior node along direction j with periodic rewiring.
__device__ __forceinline__ size_t nb_index(const LBKParams& p, const Coord& c, int j) {
    const Nb3 nx = nb3(c.x, p.X, p.per[0]);
    const Nb3 ny = nb3(c.y, p.Y, p.per[1]);
    const Nb3 nz = nb3(c.z, p.Z, p.per[2]);
    return (size_t)pick(nx, cx(j)) + (size_t)p.X * ((size_t)pick(ny, cy(j)) + (size_t)p.Y * (size_t)pick(nz, cz(j)));
}

// Particle ownership word of a node: (coupling step tag << 20) | (slot + 1).
// Entries written by an earlier coupling step carry an older tag and read as
// "no particle", so the array never needs clearing between steps.
__device__ __forceinline__ uint32_t owner_slot(uint32_t w, uint32_t tag) {
    return (w >> SOLID_SLOT_BITS) == tag ? (w & SOLID_SLOT_MASK) : 0u;
}

// ---- exact, order-free sums (free-surface mass surplus, wall forces)
// A term v is |v| 2^96 rounded toward zero to an integer q < 2^128 (|v| < 2^31);
// its four 32-bit limbs are added with sign into four 64-bit integer sums
// (integer addition is associative, so the result is the same in any order).
// The four limbs of |v| 2^96 (false: v is zero, NaN or outside the range).
__device__ inline bool exact_split(double v, unsigned long long limb[4]) {
    if (v == 0.0 || !(v == v)) return false;
    int e = 0;
    const double m = frexp(fabs(v), &e);  // |v| = m 2^e, m in [0.5, 1)
    const unsigned long long M = (unsigned long long)ldexp(m, 53);
    const int k = e + 43;  // |v| 2^96 = M 2^(e - 53 + 96)
    unsigned long long lo = 0, hi = 0;
    if (k >= 0) {
        if (k >= 75) return false;  // |v| >= 2^31: outside the accumulator's range
        if (k == 0) lo = M;
        else if (k < 64) { lo = M << k; hi = M >> (64 - k); }
        else hi = M << (k - 64);
    } else if (k > -64) {
        lo = M >> (-k);
    }
    limb[0] = lo & 0xFFFFFFFFull; limb[1] = lo >> 32; limb[2] = hi & 0xFFFFFFFFull; limb[3] = hi >> 32;
    return true;
}

__device__ inline void exact_add(unsigned long long* acc, double v) {
    unsigned long long limb[4];
    if (!exact_split(v, limb)) return;
    for (int i = 0; i < 4; ++i)
        if (limb[i]) atomicAdd(&acc[i], v < 0.0 ? (0ull - limb[i]) : limb[i]);
}

// The same sum kept per thread in signed 64-bit limbs (each below 2^32, so up to
// 2^31 terms), then folded over the wavefront and added once per wavefront.
__device__ inline void exact_local(long long s[4], double v) {
    unsigned long long limb[4];
    if (!exact_split(v, limb)) return;
    for (int i = 0; i < 4; ++i) s[i] += v < 0.0 ? -(long long)limb[i] : (long long)limb[i];
}

__device__ inline void exact_flush(unsigned long long* acc, long long s[4]) {
    for (int i = 0; i < 4; ++i) {
        long long t = s[i];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if ((threadIdx.x & 63) == 0 && t) atomicAdd(&acc[i], (unsigned long long)t);
    }
}

__device__ inline double exact_value(const unsigned long long* acc) {
    long long a[4];
    for (int i = 0; i < 4; ++i) a[i] = (long long)acc[i];
    for (int i = 0; i < 3; ++i) {  // carry into base-2^32 digits; the top digit keeps the sign
        const long long carry = a[i] >> 32;
        a[i] -= carry * 4294967296LL;
        a[i + 1] += carry;
    }
    return (((double)a[0] * 0x1p-96 + (double)a[1] * 0x1p-64) + (double)a[2] * 0x1p-32) + (double)a[3];
}

// Population f_opp(j) streamed back into a node from a wall in direction j
// (LB.cpp:1344-1456): half-way bounce-back of the node's own f_j (static or,
// with zero wall velocity, moving walls); at a slip wall, on a diagonal link,
// the reflection S1 f_slip(neighbour) + S2 f_j when exactly one of the two
// checked neighbours is active. `fs` holds the post-collision planes, `own` is
// the node's f_j and `active(n)` tells whether node n is fluid or interface.
template <typename T, class A>
__device__ __forceinline__ T wall_link(const T* __restrict__ fs, size_t PS, const LBKParams& p, const Coord& c,
                                       bool slip, int j, T own, A&& active) {
    if (!slip || j <= 6) return own;
    const size_t n1 = nb_index(p, c, slip1check(j)), n2 = nb_index(p, c, slip2check(j));
    const bool a1 = active(n1), a2 = active(n2);
    if (a1 && !a2) return (T)(p.slipS1 * (double)fs[(size_t)slip1dir(j) * PS + n1] + p.slipS2 * (double)own);
    if (!a1 && a2) return (T)(p.slipS1 * (double)fs[(size_t)slip2dir(j) * PS + n2] + p.slipS2 * (double)own);
    return own;
}

// Population f_opp(j) streamed back into a node from a curved wall node w in
// direction j (Mei-Luo-Shyy, LB.cpp:1278-1319, node.cpp:458-472): with delta the
// wall distance fraction of w towards the node (1 - the segment intercept from w,
// LB.cpp:589-602; 0 for w in a boundary layer), chi = (2 delta - 1) / tau (delta >= 1/2) or / (tau - 2), the
// reference velocity ubf = m1 u + m2 uw (delta >= 1/2) or the velocity of the
// fluid node behind (else; uw when there is none), fStar = n w_j (1 + 3 c_j.ubf +
// 4.5 (u.c_j)^2 + 1.5 u^2) and f = (1 - chi) fs_j + chi fStar - BBi with
// BBi = 6 n w_j uw.c_j. uw is the velocity the reference gives the wall node at set-up
// (LB.cpp:946-993): a plane wall's (zero for the box walls) for a node in a wall's
// boundary layer, else the cylinder's surface speed at the node (utils.cpp:106-120,
// zero for a static cylinder) in lattice units. n, u, visc: the node's own from its
// last collision. SHIFT: populations hold f - w (rho0 = 1).
struct CurvedParts {
    double chi, fStar, BBi;
};

// Whether node n was a fluid node (not interface) when the reference streamed the
// populations being pulled: a free-surface pass in between records the type class
// of the nodes it may change as (pass tag << 2) | class (0 inactive, 1 fluid,
// 2 interface).
__device__ __forceinline__ bool fluid_at_stream(const LBKParams& p, size_t n) {
    if (p.actPrev && p.actTag) {
        const uint32_t w = p.actPrev[n];
        if ((w >> 2) == p.actTag) return (w & 3u) == 1u;
    }
    return p.typew[n] == T_FLUID;
}

__device__ __forceinline__ CurvedParts curved_parts(const LBKParams& p, const Coord& c, int j) {
#pragma clang fp contract(off)
    const size_t N = (size_t)p.X * p.Y * p.Z;
    const int i = opp(j);
    const size_t w = nb_index(p, c, j), back = nb_index(p, c, i);
    const int wx = (int)(w % (size_t)p.X), wy = (int)((w / (size_t)p.X) % (size_t)p.Y), wz = (int)(w / ((size_t)p.X * p.Y));
    const double L = p.unitLength;
    const V3 vi((double)cx(i), (double)cy(i), (double)cz(i)), vj((double)cx(j), (double)cy(j), (double)cz(j));
    // boundary-layer nodes have themselves as every neighbour (LB.cpp:394-431), a
    // wall, so the reference never computes their distances: delta stays 0
    const bool layer = wx == 0 || wy == 0 || wz == 0 || wx == p.X - 1 || wy == p.Y - 1 || wz == p.Z - 1;
    const double delta =
        layer ? 0.0 : 1.0 - p.cyl.segmentIntercept(L * V3((double)wx, (double)wy, (double)wz), L * vi);
    const double n = p.macPrev[c.idx];
    const V3 u(p.macPrev[N + c.idx], p.macPrev[2 * N + c.idx], p.macPrev[3 * N + c.idx]);
    const double visc = p.viscw ? p.viscw[c.idx] : p.visc0;
    const double tau = 0.5 + 3.0 * visc;
    const double chi = delta >= 0.5 ? (2.0 * delta - 1.0) / tau : (2.0 * delta - 1.0) / (tau - 2.0);
    V3 vel(0.0, 0.0, 0.0);
    if (p.cyl.moving) {
        const bool inWall = (wx == 0 && (p.wallSides & 1)) || (wx == p.X - 1 && (p.wallSides & 2)) ||
                            (wy == 0 && (p.wallSides & 4)) || (wy == p.Y - 1 && (p.wallSides & 8)) ||
                            (wz == 0 && (p.wallSides & 16)) || (wz == p.Z - 1 && (p.wallSides & 32));
        if (!inWall) vel = p.cyl.getSpeed(V3((double)wx, (double)wy, (double)wz) * L) / p.unitSpeed;
    }
    V3 ubf(0.0, 0.0, 0.0);
    if (delta >= 0.5) {
        const double m1 = (delta - 1) / delta, m2 = 1 / delta;
        ubf = m1 * u + m2 * vel;
    } else if (fluid_at_stream(p, back)) {
        ubf = V3(p.macPrev[N + back], p.macPrev[2 * N + back], p.macPrev[3 * N + back]);
    } else {
        ubf = vel;
    }
    const double usq = u.norm2();
    const double vu = u.dot(vj);
    const double wj = wgt(j);
    const double fStar = n * wj * (1 + 3.0 * vj.dot(ubf) + 4.5 * vu * vu + 1.5 * usq);
    const double BBi = 6.0 * n * wj * vel.dot(vj);
    return CurvedParts{chi, fStar, BBi};
}

template <typename T, bool SHIFT>
__device__ __forceinline__ T curved_link(const LBKParams& p, const Coord& c, int j, T own) {
#pragma clang fp contract(off)
    const CurvedParts q = curved_parts(p, c, j);
    const double wj = wgt(j);
    const double fs = SHIFT ? (double)own + wj : (double)own;
    const double f = (1.0 - q.chi) * fs + q.chi * q.fStar - q.BBi;
    return (T)(SHIFT ? f - wj : f);
}

__device__ __forceinline__ bool interior(const LBKParams& p, const Coord& c) {
    return c.x > 0 && c.x < p.X - 1 && c.y > 0 && c.y < p.Y - 1 && c.z > 0 && c.z < p.Z - 1;
}

}  // namespace hyb
