// Discrete-element kernels for gfx950 (single-sphere elements).
//
// One DEM sub-step (DEM.cpp:331-376) is three launches on the handle's stream:
//  1. k_dem_rebuild (one workgroup): evalMaxDisp and, when the displacement
//     trigger fires, the whole neighbour rebuild of DEM.cpp:1369-1513 —
//     periodic shift, ghosts, linked-cell table, the 14-cell pair list and the
//     near-wall list — reproducing the reference's cell index (vector.cpp:147-149)
//     and unsigned stencil wrap, in the reference's pair order. It also builds a
//     per-particle list of the pairs each particle takes part in.
//  2. k_dem_pairs: one thread per pair evaluates the contact (DEM.cpp:1801-1900,
//     2138-2222) and writes each element's signed increments into that
//     element's slot of the pair (and, on saving sub-steps, its statistics).
//  3. k_dem_gather: per element, the increments in pair-list order (the same
//     order as the reference's sequential loop, so no atomics and bit-identical
//     sums), the wall contact, Newton-Euler (DEM.cpp:1149-1181), the Gear
//     corrector (elmt.cpp:179-254) and the predictor (elmt.cpp:139-177) of the
//     next sub-step. An initial predictor runs once after initialisation.
// The corrector keeps the predicted copies it has just used until the new
// prediction overwrites them, so partner reads in k_dem_pairs always see the
// predicted state of the current sub-step.
#include <hip/hip_runtime.h>

#include "common.h"
#include "dem_kernels.h"
#include "hybird_amd.h"
#include "lattice.h"

namespace hyb {

namespace {

struct El {
    double* el;
    int E;
    __device__ double& at(int f, int e) const { return el[(size_t)f * E + e]; }
    __device__ V3 v3(int f, int e) const { return V3(at(f, e), at(f + 1, e), at(f + 2, e)); }
    __device__ void set3(int f, int e, const V3& v) const { at(f, e) = v.x; at(f + 1, e) = v.y; at(f + 2, e) = v.z; }
    __device__ Q4 q4(int f, int e) const { return Q4(at(f, e), at(f + 1, e), at(f + 2, e), at(f + 3, e)); }
    __device__ void set4(int f, int e, const Q4& q) const {
        at(f, e) = q.q0; at(f + 1, e) = q.q1; at(f + 2, e) = q.q2; at(f + 3, e) = q.q3;
    }
};

constexpr int RB = 1024;  // rebuild workgroup size

// Records a device error: the sticky DEMState word (which stops every later DEM
// kernel) and the handle's host-visible error word.
__device__ __forceinline__ void dem_error(const DEMParams& p, const DEMBuffers& b, int code) {
    b.st->error = code;
    if (p.herr) *p.herr = code;
}

// In-place exclusive scan of a[0..n) by the whole (single) workgroup; returns the total.
__device__ int block_exscan(int* a, int n) {
    __shared__ int sh[RB];
    __shared__ int carry;
    const int t = threadIdx.x;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += RB) {
        const int i = base + t;
        const int v = i < n ? a[i] : 0;
        sh[t] = v;
        __syncthreads();
        for (int off = 1; off < RB; off <<= 1) {
            const int add = t >= off ? sh[t - off] : 0;
            __syncthreads();
            sh[t] += add;
            __syncthreads();
        }
        const int incl = sh[t];
        const int c0 = carry;
        if (i < n) a[i] = c0 + incl - v;
        __syncthreads();
        if (t == RB - 1) carry = c0 + incl;
        __syncthreads();
    }
    const int total = carry;
    __syncthreads();
    return total;
}

// vector.cpp:147-149 (offsets +1, +2, 0 through the parenthesisation).
__device__ __forceinline__ int linearize(const V3& x, const double cw[3], const unsigned nC[3]) {
    return (int)(floor(x.x / cw[0]) + 1 +
                 (double)nC[0] * (floor(x.y / cw[1]) + 1 + ((double)nC[1] * (floor(x.z / cw[2])) + 1)));
}

__device__ __forceinline__ V3 slot_pos(const El& E, const DEMBuffers& b, int P, int s) {
    if (s < P) return E.v3(EF_X, s);
    const int g = s - P;
    const int o = (int)b.ghostIndex[g];
    return E.v3(EF_X, o) + V3(b.ghostVec[3 * g], b.ghostVec[3 * g + 1], b.ghostVec[3 * g + 2]);
}

__device__ __forceinline__ int slot_cluster(const DEMBuffers& b, int P, int s) {
    return s < P ? s : (int)b.ghostIndex[s - P];
}

// Walks the chains of the cell pair (cell0, cell1) exactly as DEM.cpp:1455-1481.
// A chain is the cell's run of the slot array sorted by cell with slots
// descending (the order of the linked list), so its loads do not wait on each
// other; positions and clusters come from the packed slot records (spos).
// With out == nullptr only counts. Returns -1 when the stencil leaves the table.
__device__ int walk_pairs(const DEMParams& p, const DEMBuffers& b, const El& E, int S, unsigned w, uint32_t* out) {
    static constexpr int sc[14][3] = {{0, 0, 0},  {1, 0, 0},  {1, 1, 0},  {0, 1, 0},  {-1, 1, 0},
                                      {0, 0, 1},  {1, 0, 1},  {1, 1, 1},  {0, 1, 1},  {-1, 1, 1},
                                      {-1, 0, 1}, {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}};
    const unsigned sh = w % 14u;
    unsigned r = w / 14u;
    unsigned i0[3];
    i0[2] = r % p.nCells[2];
    r /= p.nCells[2];
    i0[1] = r % p.nCells[1];
    i0[0] = r / p.nCells[1];
    unsigned i1[3];
    for (int k = 0; k < 3; ++k) {
        i1[k] = i0[k] + (unsigned)sc[sh][k];
        if (i1[k] == p.nCells[k]) i1[k] = 0;
    }
    const uint32_t Su = (uint32_t)S;
    const uint32_t cell0 = (i0[2] * p.nCells[1] + i0[1]) * p.nCells[0] + i0[0] + Su;
    const uint32_t cell1 = (i1[2] * p.nCells[1] + i1[1]) * p.nCells[0] + i1[0] + Su;
    const uint32_t tsize = Su + (uint32_t)b.ncells;
    const double r2 = p.nebrRange * p.nebrRange;
    int cnt = 0;
    const int a0 = b.cellCount[cell0 - Su], e0 = b.cellCount[cell0 - Su + 1];  // cell0 - Su < ncells
    if (a0 >= e0) return 0;
    if (cell1 >= tsize) return -1;
    // A stencil step of -1 from index 0 wraps through unsigned arithmetic to the
    // previous linear cell, and from linear cell 0 below the cell heads: then
    // cellTable[cell1] is the link of slot cell1 and the reference walks the
    // rest of that slot's chain (DEM.cpp:1373-1390, 1455-1481). Only that case
    // follows the links.
    const bool linked = cell1 < Su;
    const int a1 = linked ? 0 : b.cellCount[cell1 - Su], e1 = linked ? 0 : b.cellCount[cell1 - Su + 1];
    const auto visit = [&](int n0, const V3& x0, int c0, int n1) {
        if (cell0 != cell1 || n1 < n0) {
            const double* s1 = b.spos + 4 * (size_t)n1;
            if (c0 != (int)s1[3]) {
                const V3 r10 = V3(s1[0], s1[1], s1[2]) - x0;
                if (r10.norm2() < r2) {
                    if (out) {
                        out[2 * cnt] = (uint32_t)n0;
                        out[2 * cnt + 1] = (uint32_t)n1;
                    }
                    ++cnt;
                }
            }
        }
    };
    for (int i0 = a0; i0 < e0; ++i0) {
        const int n0 = b.cursor[i0];
        const double* s0 = b.spos + 4 * (size_t)n0;
        const V3 x0(s0[0], s0[1], s0[2]);
        const int c0 = (int)s0[3];
        if (linked) {
            for (int n1 = b.cellTable[cell1]; n1 > -1; n1 = b.cellTable[n1]) visit(n0, x0, c0, n1);
        } else {
            for (int i1 = a1; i1 < e1; ++i1) visit(n0, x0, c0, b.cursor[i1]);
        }
    }
    return cnt;
}

__global__ void __launch_bounds__(RB) k_dem_rebuild(DEMParams p, DEMBuffers b, int force) {
#pragma clang fp contract(off)
    const El E{b.el, p.P};
    const int t = threadIdx.x;
    const int P = p.P;
    __shared__ double red[RB];
    __shared__ int doit;
    __shared__ int sh_err;
    if (b.st->error) return;
    if (t == 0) sh_err = 0;
    __syncthreads();
    // ---- evalMaxDisp (DEM.cpp:1314-1324) and the trigger (DEM.cpp:342-347)
    if (!force) {
        // max |x1|^2 over elements, from the per-wavefront maxima the corrector
        // left; a NaN or infinite maximum stops the DEM (error 6)
        double m = 0.0;
        const int nw = (P + 63) / 64;
        for (int k = t; k < nw; k += RB) {
            const double v = b.vmax[k];
            if (!(v <= 1.79769313486231570e308)) sh_err = DEM_ERR_NUMERIC;
            if (v > m) m = v;
        }
        red[t] = m;
        __syncthreads();
        if (sh_err) { if (t == 0) dem_error(p, b, DEM_ERR_NUMERIC); return; }
        for (int w = RB / 2; w > 0; w >>= 1) {
            if (t < w && red[t + w] > red[t]) red[t] = red[t + w];
            __syncthreads();
        }
        if (t == 0) {
            const double maxVel = sqrt(red[0]);
            b.st->maxDisp += maxVel * p.deltat;
            doit = b.st->maxDisp > p.trigger;
            if (doit) b.st->maxDisp = 0.0;
        }
    } else if (t == 0) {
        doit = 1;
    }
    __syncthreads();
    if (!doit || P == 0) return;
    if (t == 0) b.st->rebuilds += 1;
    int G = 0;
    if (p.npbc) {
        // ---- pbcShift (DEM.cpp:1554-1584)
        // The predictor of this sub-step already ran (fused into the previous
        // corrector), so a shifted element gets its predicted position
        // recomputed from the shifted x0 exactly as the predictor would.
        for (int e = t; e < P; e += RB) {
            V3 x0 = E.v3(EF_X, e);
            bool moved = false;
            for (int k = 0; k < p.npbc; ++k) {
                const PbcDev& pb = p.pbc[k];
                const double left = pb.n1.dot(x0 - pb.p1);
                const double right = pb.n2.dot(x0 - pb.p2);
                if (left < 0.0) { x0 += pb.v; moved = true; }
                if (right < 0.0) { const V3 m = -1.0 * pb.v; x0 += m; moved = true; }
            }
            if (moved) {
                E.set3(EF_X, e, x0);
                const double* c2 = p.c2;
                const V3 xp0 = x0 + E.v3(EF_X + 3, e) * c2[0] + E.v3(EF_X + 6, e) * c2[1] + E.v3(EF_X + 9, e) * c2[2] +
                               E.v3(EF_X + 12, e) * c2[3] + E.v3(EF_X + 15, e) * c2[4];
                E.set3(EF_XP, e, xp0);
            }
        }
        __syncthreads();
        // ---- side ghosts in (pbc, element) order (DEM.cpp:1590-1618)
        int nSide = 0;
        for (int k = 0; k < p.npbc; ++k) {
            const PbcDev& pb = p.pbc[k];
            for (int e = t; e < P; e += RB) {
                const V3 x0 = E.v3(EF_X, e);
                const double left = pb.n1.dot(x0 - pb.p1);
                const double right = pb.n2.dot(x0 - pb.p2);
                const int flag = left < p.nebrRange ? 1 : (right < p.nebrRange ? 2 : 0);
                b.work[e] = flag ? 1 : 0;
                b.side[e * 3 + k] = flag;  // temporarily the flag
            }
            __syncthreads();
            const int tot = block_exscan(b.work, P);
            for (int e = t; e < P; e += RB) {
                const int flag = b.side[e * 3 + k];
                if (flag) {
                    const int g = nSide + b.work[e];
                    if (g < p.Gcap) {
                        b.ghostIndex[g] = (uint32_t)e;
                        const V3 v = flag == 1 ? pb.v : -1.0 * pb.v;
                        b.ghostVec[3 * g] = v.x;
                        b.ghostVec[3 * g + 1] = v.y;
                        b.ghostVec[3 * g + 2] = v.z;
                        b.ghostB[g] = k;
                    }
                    b.side[e * 3 + k] = g;
                } else {
                    b.side[e * 3 + k] = -1;
                }
            }
            nSide += tot;
            __syncthreads();
        }
        if (nSide > p.Gcap) { if (t == 0) sh_err = 1; }
        __syncthreads();
        if (sh_err) { if (t == 0) dem_error(p, b, 1); return; }
        // ---- corner ghosts: every pair g1 < g2 of side ghosts of one particle (DEM.cpp:1620-1635)
        for (int g = t; g < nSide; g += RB) {
            const int e = (int)b.ghostIndex[g];
            int c = 0;
            for (int k2 = b.ghostB[g] + 1; k2 < p.npbc; ++k2) c += b.side[e * 3 + k2] >= 0 ? 1 : 0;
            b.work[g] = c;
        }
        __syncthreads();
        const int nCorner = block_exscan(b.work, nSide);
        G = nSide + nCorner;
        if (G > p.Gcap) { if (t == 0) dem_error(p, b, 1); return; }
        for (int g = t; g < nSide; g += RB) {
            const int e = (int)b.ghostIndex[g];
            int slot = nSide + b.work[g];
            const V3 v1(b.ghostVec[3 * g], b.ghostVec[3 * g + 1], b.ghostVec[3 * g + 2]);
            for (int k2 = b.ghostB[g] + 1; k2 < p.npbc; ++k2) {
                const int g2 = b.side[e * 3 + k2];
                if (g2 < 0) continue;
                const V3 v2(b.ghostVec[3 * g2], b.ghostVec[3 * g2 + 1], b.ghostVec[3 * g2 + 2]);
                const V3 v = v1 + v2;
                b.ghostIndex[slot] = (uint32_t)e;
                b.ghostVec[3 * slot] = v.x;
                b.ghostVec[3 * slot + 1] = v.y;
                b.ghostVec[3 * slot + 2] = v.z;
                ++slot;
            }
        }
        if (t == 0) b.st->newNeighborList = 1;
        __syncthreads();
    }
    const int S = P + G;
    if (t == 0) { b.st->nGhost = G; b.st->nSlots = S; }
    // ---- linked-cell table (DEM.cpp:1326-1367): chains in descending slot order
    const int nc = b.ncells;
    for (int c = t; c <= nc; c += RB) b.cellCount[c] = 0;
    __syncthreads();
    for (int s = t; s < S; s += RB) {
        const int c = linearize(slot_pos(E, b, P, s), p.cellWidth, p.nCells) + S;
        if (c < 0 || c >= S + nc) {
            sh_err = 2;
            b.cellOf[s] = 0;
        } else {
            b.cellOf[s] = c - S;
            atomicAdd(&b.cellCount[c - S], 1);
        }
    }
    __syncthreads();
    if (sh_err) { if (t == 0) dem_error(p, b, 2); return; }
    block_exscan(b.cellCount, nc + 1);
    for (int c = t; c < nc; c += RB) b.work[c] = b.cellCount[c];
    __syncthreads();
    for (int s = t; s < S; s += RB) {
        const int pos = atomicAdd(&b.work[b.cellOf[s]], 1);
        b.cursor[pos] = s;  // cursor holds the slots sorted by cell (order within a cell fixed below)
    }
    __syncthreads();
    for (int c = t; c < nc; c += RB) {
        const int lo = b.cellCount[c], hi = b.cellCount[c + 1];
        for (int i = lo + 1; i < hi; ++i) {  // insertion sort, descending
            const int v = b.cursor[i];
            int j = i - 1;
            while (j >= lo && b.cursor[j] < v) { b.cursor[j + 1] = b.cursor[j]; --j; }
            b.cursor[j + 1] = v;
        }
        b.cellTable[S + c] = hi > lo ? b.cursor[lo] : -1;
        for (int i = lo; i < hi; ++i) b.cellTable[b.cursor[i]] = (i + 1 < hi) ? b.cursor[i + 1] : -1;
    }
    __syncthreads();
    // ---- pair list in the reference's order (DEM.cpp:1425-1484): count, scan, fill
    for (int s = t; s < S; s += RB) {  // packed slot records: position, cluster
        const V3 x = slot_pos(E, b, P, s);
        double* r = b.spos + 4 * (size_t)s;
        r[0] = x.x;
        r[1] = x.y;
        r[2] = x.z;
        r[3] = (double)slot_cluster(b, P, s);
    }
    __syncthreads();
    const int W = b.W;
    for (int w = t; w < W; w += RB) {
        const int c = walk_pairs(p, b, E, S, (unsigned)w, nullptr);
        if (c < 0) { sh_err = 3; b.work[w] = 0; }
        else b.work[w] = c;
    }
    __syncthreads();
    if (sh_err) { if (t == 0) dem_error(p, b, 3); return; }
    const int T = block_exscan(b.work, W);
    if (T > p.pairCap) { if (t == 0) dem_error(p, b, 4); return; }
    for (int w = t; w < W; w += RB) walk_pairs(p, b, E, S, (unsigned)w, b.pairs + 2 * (size_t)b.work[w]);
    if (t == 0) b.st->nPairs = T;
    // ---- near-wall list (DEM.cpp:1496-1513): the first wall within range
    int nw = 0;
    for (int e = t; e < P; e += RB) {
        const V3 x0 = E.v3(EF_X, e);
        uint32_t wi = 0xFFFFFFFFu;
        for (int k = 0; k < p.nwalls; ++k)
            if (p.walls[k].dist(x0) < p.nebrRange) { wi = (uint32_t)k; break; }
        b.nearWall[e] = wi;
        nw += wi != 0xFFFFFFFFu;
        // near-cylinder list (DEM.cpp:1535-1550): the first cylinder within range
        if (p.ncyl) {
            uint32_t ci = 0xFFFFFFFFu;
            for (int k = 0; k < p.ncyl; ++k)
                if (p.cyl[k].dist(x0) < p.nebrRange) { ci = (uint32_t)k; break; }
            b.nearCyl[e] = ci;
        }
    }
    red[t] = (double)nw;
    // ---- near-object lists (DEM.cpp:1515-1532): every object within range, in list order
    if (p.nobj) {
        const double nebrRange2 = p.nebrRange * p.nebrRange;
        for (int e = t; e < P; e += RB) {
            const V3 x0 = E.v3(EF_X, e);
            int c = 0;
            for (int o = 0; o < p.nobj; ++o) {
                const V3 x0ij = x0 - b.objs[o].x0;
                if (x0ij.norm2() < nebrRange2) {
                    if (c < NEAR_OBJ) b.nearObj[(size_t)e * NEAR_OBJ + c] = (uint32_t)o;
                    ++c;
                }
            }
            if (c > NEAR_OBJ) { sh_err = 5; c = NEAR_OBJ; }
            b.nearObjCount[e] = c;
        }
    }
    __syncthreads();
    if (sh_err) { if (t == 0) dem_error(p, b, 5); return; }
    // ---- per-particle pair lists, ascending pair index
    for (int e = t; e < P; e += RB) b.cursor[e] = 0;
    __syncthreads();
    for (int k = t; k < T; k += RB) {
        const uint32_t n0 = b.pairs[2 * k], n1 = b.pairs[2 * k + 1];
        if ((int)n0 < P) atomicAdd(&b.cursor[n0], 1);
        if ((int)n1 < P) atomicAdd(&b.cursor[n1], 1);
    }
    __syncthreads();
    for (int e = t; e < P; e += RB) b.work[e] = b.cursor[e];
    __syncthreads();
    const int tot2 = block_exscan(b.work, P);
    for (int e = t; e < P; e += RB) { b.csrOff[e] = (uint32_t)b.work[e]; b.cursor[e] = b.work[e]; }
    if (t == 0) b.csrOff[P] = (uint32_t)tot2;
    __syncthreads();
    for (int k = t; k < T; k += RB) {
        const uint32_t n0 = b.pairs[2 * k], n1 = b.pairs[2 * k + 1];
        if ((int)n0 < P) b.csr[atomicAdd(&b.cursor[n0], 1)] = (uint32_t)k;
        if ((int)n1 < P) b.csr[atomicAdd(&b.cursor[n1], 1)] = (uint32_t)k;
    }
    __syncthreads();
    for (int e = t; e < P; e += RB) {
        const int lo = (int)b.csrOff[e], hi = (int)b.csrOff[e + 1];
        for (int i = lo + 1; i < hi; ++i) {
            const uint32_t v = b.csr[i];
            int j = i - 1;
            while (j >= lo && b.csr[j] > v) { b.csr[j + 1] = b.csr[j]; --j; }
            b.csr[j + 1] = v;
        }
    }
    // ---- slot of each pair in its particles' lists
    for (int k = t; k < T; k += RB) { b.posI[k] = 0xFFFFFFFFu; b.posJ[k] = 0xFFFFFFFFu; }
    __syncthreads();
    for (int e = t; e < P; e += RB) {
        const int lo = (int)b.csrOff[e], hi = (int)b.csrOff[e + 1];
        for (int i = lo; i < hi; ++i) {
            const uint32_t k = b.csr[i];
            if ((int)b.pairs[2 * k] == e) b.posI[k] = (uint32_t)i;
            else b.posJ[k] = (uint32_t)i;
        }
    }
    // ---- ghost slots of each particle, ascending (for the hydrodynamic sums)
    for (int e = t; e < P; e += RB) b.ghostOfCount[e] = 0;
    __syncthreads();
    for (int g = t; g < G; g += RB) {
        const int e = (int)b.ghostIndex[g];
        const int k = atomicAdd(&b.ghostOfCount[e], 1);
        if (k < 8) b.ghostOf[e * 8 + k] = P + g;
    }
    __syncthreads();
    for (int e = t; e < P; e += RB) {
        const int n = b.ghostOfCount[e] < 8 ? b.ghostOfCount[e] : 8;
        int* l = b.ghostOf + e * 8;
        for (int i = 1; i < n; ++i) {
            const int v = l[i];
            int j = i - 1;
            while (j >= 0 && l[j] > v) { l[j + 1] = l[j]; --j; }
            l[j + 1] = v;
        }
    }
    // near-wall count
    for (int w = RB / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) b.st->nNearWall = (int)red[0];
}

// ------------------------------------------------------------------ contacts, Newton-Euler, Gear
__device__ __forceinline__ double normal_contact(const DEMParams& p, double overlap, double vrel, double effRad,
                                                 double effMass) {  // DEM.cpp:2138-2174
    if (p.hertzian) {
        const double sqrtEffRad = sqrt(effRad);
        const double sqrtOverlap = sqrt(overlap);
        const double kn = p.knConst * sqrtEffRad * sqrtOverlap;
        const double gamman = 2.0 * p.dampCoeff * sqrt(kn * effMass);
        const double dumpfn = -gamman * vrel;
        const double elasticfn = kn * overlap;
        const double fn = elasticfn + dumpfn;
        return (fn < 0.0) ? 0.0 : fn;  // std::max(fn, 0.0)
    }
    const double elasticfn = p.linearStiff * overlap;
    const double gamman = 2.0 * p.dampCoeff * sqrt(p.linearStiff * effMass);
    const double dumpfn = -gamman * vrel;
    const double fn = elasticfn + dumpfn;
    return (fn < 0.0) ? 0.0 : fn;
}

__device__ __forceinline__ double tangential_contact(const DEMParams& p, double vrelt, double fn, double effRad,
                                                     double effMass, double friction) {  // DEM.cpp:2176-2222
    double ks;
    if (p.hertzian) ks = p.ksConst * sqrt(effRad) * pow(fabs(fn), 1.0 / 3.0);
    else ks = p.linearStiff;
    const double fsMax = friction * fn;
    const double gammas = 2.0 * p.viscTang * sqrt(effMass * ks);
    const double dumpfs = gammas * vrelt;
    return (fsMax < dumpfs) ? fsMax : dumpfs;  // std::min(dumpfs, fsMax)
}

// Predicted position / velocity of a particle slot (updateParticlesPredicted, DEM.cpp:1184-1204).
__device__ __forceinline__ void slot_pred(const El& E, const DEMBuffers& b, int P, int s, V3& x, V3& v, int& elem) {
    if (s < P) {
        x = E.v3(EF_XP, s);
        v = E.v3(EF_XP + 3, s);
        elem = s;
    } else {
        const int g = s - P;
        elem = (int)b.ghostIndex[g];
        x = E.v3(EF_XP, elem) + V3(b.ghostVec[3 * g], b.ghostVec[3 * g + 1], b.ghostVec[3 * g + 2]);
        v = E.v3(EF_XP + 3, elem);
    }
}

// Contact statistics and wall/object force records written by the contact code.
__device__ __forceinline__ void set_stat(StatRec& s, double ov, double fn, double ft, const V3& en, const V3& et) {
    s.ov = ov; s.fn = fn; s.ft = ft;
    s.en[0] = en.x; s.en[1] = en.y; s.en[2] = en.z;
    s.et[0] = et.x; s.et[1] = et.y; s.et[2] = et.z;
    s.flag = 1.0;
}

__device__ __forceinline__ void set_force(ContactForce& c, const V3& fn, double flag) {
    c.fn[0] = fn.x; c.fn[1] = fn.y; c.fn[2] = fn.z;
    c.flag = flag;
}

__device__ __forceinline__ void set_force_t(ContactForce& c, const V3& ft) {
    c.ft[0] = ft.x; c.ft[1] = ft.y; c.ft[2] = ft.z;
    c.flag = 3.0;
}

// The contact of neighbour pair k (DEM.cpp:1801-1900) at the predicted state:
// the signed increments for the pair's first (ri) and second (rj) element, and
// its statistics record st.
__device__ __forceinline__ void pair_contact(const DEMParams& p, const DEMBuffers& b, const El& E, int k, PairRec& ri,
                                             PairRec& rj, StatRec& st) {
#pragma clang fp contract(off)
    const int P = p.P;
    const int n0 = (int)b.pairs[2 * k], n1 = (int)b.pairs[2 * k + 1];
    V3 xI, vI, xJ, vJ;
    int eI, eJ;
    slot_pred(E, b, P, n0, xI, vI, eI);
    slot_pred(E, b, P, n1, xJ, vJ, eJ);
    const double rI = E.at(EF_RAD, eI), rJ = E.at(EF_RAD, eJ);
    const double sigij = rI + rJ;
    const double sigij2 = sigij * sigij;
    const V3 vd = xJ - xI;
    ri.flag = rj.flag = 0;
    ri.pad = rj.pad = 0;
    st = StatRec{(double)n0, (double)n1, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0};
    if (!(vd.norm2() < sigij2)) return;
    const double distance = vd.norm();
    const double overlap = rI + rJ - distance;
    const V3 relVel = vJ - vI;
    const V3 en = vd / distance;
    const double normRelVel = relVel.dot(en);
    const V3 normalRelVel = en * normRelVel;
    const double mI = E.at(EF_M, eI), mJ = E.at(EF_M, eJ);
    const double effMass = mI * mJ / (mI + mJ);
    const double effRad = rI * rJ / (rI + rJ);
    const double fn = normal_contact(p, overlap, normRelVel, effRad, effMass);
    const V3 normalForce = en * fn;
    const V3 vecRadI = rI * en;
    const V3 vecRadJ = -rJ * en;
    // radiusVec is zero for single-sphere elements
    const V3 centerDistI = vecRadI + V3(0.0, 0.0, 0.0);
    const V3 centerDistJ = vecRadJ + V3(0.0, 0.0, 0.0);
    const V3 mn = -1.0 * normalForce;
    ri.fn[0] = mn.x; ri.fn[1] = mn.y; ri.fn[2] = mn.z;  // F - n == F + (-n)
    rj.fn[0] = normalForce.x; rj.fn[1] = normalForce.y; rj.fn[2] = normalForce.z;
    ri.flag = rj.flag = 1;
    const V3 wI = E.v3(EF_WPG, eI), wJ = E.v3(EF_WPG, eJ);
    const V3 relVelContact = relVel - wI.cross(vecRadI) + wJ.cross(vecRadJ);
    const V3 tangRelVelContact = relVelContact - normalRelVel;
    const double nt = tangRelVelContact.norm();
    if (nt != 0.0) {
        const double ft = tangential_contact(p, nt, fn, effRad, effMass, p.fricPart);
        const V3 et = tangRelVelContact / nt;
        const V3 tangForce = ft * et;
        const V3 mI3 = centerDistI.cross(tangForce);
        const V3 mJ3 = -1.0 * centerDistJ.cross(tangForce);
        const V3 mt = -1.0 * tangForce;
        ri.ft[0] = tangForce.x; ri.ft[1] = tangForce.y; ri.ft[2] = tangForce.z;
        ri.m[0] = mI3.x; ri.m[1] = mI3.y; ri.m[2] = mI3.z;
        rj.ft[0] = mt.x; rj.ft[1] = mt.y; rj.ft[2] = mt.z;
        rj.m[0] = mJ3.x; rj.m[1] = mJ3.y; rj.m[2] = mJ3.z;
        ri.flag = rj.flag = 3;
        set_stat(st, overlap, fn, ft, en, et);
    }
}

// One thread per neighbour pair (grid-stride): the contact at the predicted
// state; the signed increments for each of the two elements go into that
// element's slot of the pair in its csr list, so that k_dem_gather adds them in
// pair-list order.
__global__ void __launch_bounds__(256) k_dem_pairs(DEMParams p, DEMBuffers b) {
    const El E{b.el, p.P};
    if (b.st->error) return;
    const int T = b.st->nPairs;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < T; k += gridDim.x * blockDim.x) {
        const uint32_t si = b.posI[k], sj = b.posJ[k];
        PairRec ri, rj;
        StatRec st;
        pair_contact(p, b, E, k, ri, rj, st);
        if (b.pstat) b.pstat[k] = st;
        if (si != 0xFFFFFFFFu) b.prec[si] = ri;
        if (sj != 0xFFFFFFFFu) b.prec[sj] = rj;
    }
}

// Maximum over the 64 lanes of a wavefront (order-free, so deterministic).
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Gear predictor of one element from its corrected state (elmt.cpp:139-177).
__device__ __forceinline__ void predict_elem(const El& E, const DEMParams& p, int e, const V3 x[6], const Q4 q[6],
                                             const V3 w[6]) {
    const double* c1 = p.c1;
    const double* c2 = p.c2;
    const V3 xp0 = x[0] + x[1] * c2[0] + x[2] * c2[1] + x[3] * c2[2] + x[4] * c2[3] + x[5] * c2[4];
    const V3 xp1 = x[1] + x[2] * c2[0] + x[3] * c2[1] + x[4] * c2[2] + x[5] * c2[3];
    const V3 xp2 = x[2] + x[3] * c2[0] + x[4] * c2[1] + x[5] * c2[2];
    const V3 xp3 = x[3] + x[4] * c2[0] + x[5] * c2[1];
    const V3 xp4 = x[4] + x[5] * c2[0];
    const V3 xp5 = x[5];
    Q4 qp0 = q[0] + q[1] * c2[0] + q[2] * c2[1] + q[3] * c2[2] + q[4] * c2[3] + q[5] * c2[4];
    const Q4 qp1 = q[1] + q[2] * c2[0] + q[3] * c2[1] + q[4] * c2[2] + q[5] * c2[3];
    const Q4 qp2 = q[2] + q[3] * c2[0] + q[4] * c2[1] + q[5] * c2[2];
    const Q4 qp3 = q[3] + q[4] * c2[0] + q[5] * c2[1];
    const Q4 qp4 = q[4] + q[5] * c2[0];
    const Q4 qp5 = q[5];
    qp0.normalize();
    const V3 wp0 = w[0] + w[1] * c1[0] + w[2] * c1[1] + w[3] * c1[2] + w[4] * c1[3] + w[5] * c1[4];
    const V3 wp1 = w[1] + w[2] * c1[0] + w[3] * c1[1] + w[4] * c1[2] + w[5] * c1[3];
    const V3 wp2 = w[2] + w[3] * c1[0] + w[4] * c1[1] + w[5] * c1[2];
    const V3 wp3 = w[3] + w[4] * c1[0] + w[5] * c1[1];
    const V3 wp4 = w[4] + w[5] * c1[0];
    const V3 wp5 = w[5];
    E.set3(EF_XP, e, xp0); E.set3(EF_XP + 3, e, xp1); E.set3(EF_XP + 6, e, xp2);
    E.set3(EF_XP + 9, e, xp3); E.set3(EF_XP + 12, e, xp4); E.set3(EF_XP + 15, e, xp5);
    E.set4(EF_QP, e, qp0); E.set4(EF_QP + 4, e, qp1); E.set4(EF_QP + 8, e, qp2);
    E.set4(EF_QP + 12, e, qp3); E.set4(EF_QP + 16, e, qp4); E.set4(EF_QP + 20, e, qp5);
    E.set3(EF_WP, e, wp0); E.set3(EF_WP + 3, e, wp1); E.set3(EF_WP + 6, e, wp2);
    E.set3(EF_WP + 9, e, wp3); E.set3(EF_WP + 12, e, wp4); E.set3(EF_WP + 15, e, wp5);
    E.set3(EF_WPG, e, wp0);
    E.set3(EF_WPL, e, project(wp0, qp0.adjoint()));
}

__global__ void __launch_bounds__(64) k_dem_predict(DEMParams p, DEMBuffers b) {
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const El E{b.el, p.P};
    double v2 = 0.0;
    if (e < p.P) {
        V3 x[6], w[6];
        Q4 q[6];
        for (int k = 0; k < 6; ++k) { x[k] = E.v3(EF_X + 3 * k, e); q[k] = E.q4(EF_Q + 4 * k, e); w[k] = E.v3(EF_W + 3 * k, e); }
        predict_elem(E, p, e, x, q, w);
        v2 = x[1].norm2();
    }
    v2 = wave_max(v2);
    if (threadIdx.x == 0) b.vmax[blockIdx.x] = v2;
}

// Per element: particle-contact increments in pair-list order, the wall contact
// (DEM.cpp:1692-1717, 1902-1982), Newton-Euler (DEM.cpp:1149-1181), the Gear
// corrector (elmt.cpp:179-254) and the predictor of the next sub-step.
__device__ __forceinline__ double gather_elem(const DEMParams& p, const DEMBuffers& b, int e);

__global__ void __launch_bounds__(64) k_dem_gather(DEMParams p, DEMBuffers b) {
    if (b.st->error) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double v2 = e < p.P ? gather_elem(p, b, e) : 0.0;
    v2 = wave_max(v2);
    if (threadIdx.x == 0) b.vmax[blockIdx.x] = v2;
}

// Returns |x1|^2 of the corrected element for the next displacement check.
__device__ __forceinline__ double gather_elem(const DEMParams& p, const DEMBuffers& b, int e) {
#pragma clang fp contract(off)
    const El E{b.el, p.P};
    V3 FP(0.0, 0.0, 0.0), MP(0.0, 0.0, 0.0), FW(0.0, 0.0, 0.0), MW(0.0, 0.0, 0.0);
    const uint32_t lo = b.csrOff[e], hi = b.csrOff[e + 1];
    // the element state does not depend on the contact records: its loads are
    // issued first, so that they overlap with the record loads below
    const uint32_t wi = b.nearWall[e];
    const V3 xpe = E.v3(EF_XP, e);
    const V3 vpe = E.v3(EF_XP + 3, e);
    const double re = E.at(EF_RAD, e), me = E.at(EF_M, e);
    const V3 wpg = E.v3(EF_WPG, e);
    const V3 FH = E.v3(EF_FH, e), MH = E.v3(EF_MH, e);
    const V3 Ie = E.v3(EF_I, e), wpl = E.v3(EF_WPL, e);
    const V3 xp[6] = {xpe, vpe, E.v3(EF_XP + 6, e), E.v3(EF_XP + 9, e), E.v3(EF_XP + 12, e), E.v3(EF_XP + 15, e)};
    const V3 wp[6] = {E.v3(EF_WP, e), E.v3(EF_WP + 3, e), E.v3(EF_WP + 6, e), E.v3(EF_WP + 9, e), E.v3(EF_WP + 12, e),
                      E.v3(EF_WP + 15, e)};
    const Q4 q2 = E.q4(EF_Q + 8, e);  // single spheres: q2 is never set by the force routine
    const Q4 qp[6] = {E.q4(EF_QP, e), E.q4(EF_QP + 4, e), E.q4(EF_QP + 8, e), E.q4(EF_QP + 12, e),
                      E.q4(EF_QP + 16, e), E.q4(EF_QP + 20, e)};
    // contact increments in pair-list order; records are loaded GU at a time
    const auto add = [&](const PairRec& r) {
        const int flag = r.flag;
        if (flag & 1) FP = FP + V3(r.fn[0], r.fn[1], r.fn[2]);
        if (flag & 2) {
            MP = MP + V3(r.m[0], r.m[1], r.m[2]);
            FP = FP + V3(r.ft[0], r.ft[1], r.ft[2]);
        }
    };
    constexpr uint32_t GU = 4;
    uint32_t i = lo;
    for (; i + GU <= hi; i += GU) {
        PairRec r[GU];
#pragma unroll
        for (uint32_t k = 0; k < GU; ++k) r[k] = b.prec[i + k];
#pragma unroll
        for (uint32_t k = 0; k < GU; ++k) add(r[k]);
    }
    for (; i < hi; ++i) add(b.prec[i]);
    StatRec wst{wi == 0xFFFFFFFFu ? -1.0 : (double)wi, (double)e, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0},
                0.0};
    ContactForce wcf{{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0, 0.0};
    if (wi != 0xFFFFFFFFu) {
        const WallDev& w = p.walls[wi];
        const double distance = w.dist(xpe);
        const double overlap = re - distance;
        if (overlap > 0.0) {
            const V3 en = w.n;
            const V3 cpv = w.getSpeed(xpe);
            const V3 relVel = vpe - cpv;
            const double normRelVel = relVel.dot(en);
            const V3 normalRelVel = en * normRelVel;
            const double fn = normal_contact(p, 2.0 * overlap, normRelVel, re, me);
            const V3 normalForce = en * fn;
            const V3 vecRadJ = -re * en;
            const V3 centerDistJ = vecRadJ;
            FW = FW + normalForce;
            set_force(wcf, normalForce, 1.0);
            const V3 relVelContact = relVel + wpg.cross(vecRadJ);
            const V3 tangRelVelContact = relVelContact - normalRelVel;
            const double nt = tangRelVelContact.norm();
            if (nt != 0.0) {
                const double ft = tangential_contact(p, nt, fn, re, me, p.fricWall);
                const V3 et = tangRelVelContact / tangRelVelContact.norm();
                const V3 tangForce = ft * et;
                MW = MW - centerDistJ.cross(tangForce);
                FW = FW - tangForce;
                set_force_t(wcf, tangForce);
                if (b.wstat) set_stat(wst, overlap, fn, ft, en, et);
            }
        }
        b.wcf[e] = wcf;
    }
    if (b.wstat) b.wstat[e] = wst;
    // object contacts (DEM.cpp:1747-1770, 2060-2136), in list order after the wall
    if (p.nobj) {
        const int no = b.nearObjCount[e];
        if (b.ostat)
            for (int k = no; k < NEAR_OBJ; ++k)
                b.ostat[(size_t)e * NEAR_OBJ + k] = StatRec{-1.0, (double)e, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0},
                                                            {0.0, 0.0, 0.0}, 0.0};
        for (int k = 0; k < no; ++k) {
            const uint32_t oslot = b.nearObj[(size_t)e * NEAR_OBJ + k];
            const ObjDev& ob = b.objs[oslot];
            const V3 vectorDistance = xpe - ob.x0;
            const double distance = vectorDistance.norm();
            const double overlap = re + ob.r - distance;
            StatRec ost{(double)oslot, (double)e, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0};
            ContactForce ocf{{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0, 0.0};
            if (!(overlap > 0.0)) {
                if (b.ostat) b.ostat[(size_t)e * NEAR_OBJ + k] = ost;
                b.ocf[(size_t)e * NEAR_OBJ + k] = ocf;
                continue;
            }
            const V3 en = (1.0 / distance) * vectorDistance;
            const V3 relVel = vpe - ob.x1;
            const double normRelVel = relVel.dot(en);
            const V3 normalRelVel = en * normRelVel;
            const double fn = normal_contact(p, 2.0 * overlap, normRelVel, re, me);
            const V3 normalForce = en * fn;
            const V3 vecRadJ = -re * en;
            const V3 centerDistJ = vecRadJ + (xpe - xpe);  // particle = element centre (single spheres)
            FW = FW + normalForce;
            set_force(ocf, normalForce, 1.0);
            const V3 relVelContact = relVel + wpg.cross(vecRadJ);
            const V3 tangRelVelContact = relVelContact - normalRelVel;
            const double nt = tangRelVelContact.norm();
            if (nt != 0.0) {
                const double ft = tangential_contact(p, nt, fn, re, me, p.fricPart);
                const V3 et = tangRelVelContact / tangRelVelContact.norm();
                const V3 tangForce = ft * et;
                MW = MW - centerDistJ.cross(tangForce);
                FW = FW - tangForce;
                set_force_t(ocf, tangForce);
                if (b.ostat) set_stat(ost, overlap, fn, ft, en, et);
            }
            if (b.ostat) b.ostat[(size_t)e * NEAR_OBJ + k] = ost;
            b.ocf[(size_t)e * NEAR_OBJ + k] = ocf;
        }
    }
    // cylinder contact (DEM.cpp:1719-1744, 1984-2058), after the objects
    if (p.ncyl) {
        const uint32_t ci = b.nearCyl[e];
        if (ci != 0xFFFFFFFFu) {
            const CylDev& cyl = p.cyl[ci];
            const double distance = cyl.dist(xpe);
            const double overlap = re - distance;
            if (overlap > 0.0) {
                const V3 vecDistance = cyl.vecDist(xpe);
                const V3 contactPoint = xpe - vecDistance;
                const V3 en = vecDistance / (re - overlap);
                const V3 contactPointVelocity = cyl.getSpeed(contactPoint);
                const V3 relVel = vpe - contactPointVelocity;
                const double normRelVel = relVel.dot(en);
                const V3 normalRelVel = en * normRelVel;
                const double fn = normal_contact(p, 2.0 * overlap, normRelVel, re, me);
                const V3 normalForce = en * fn;
                const V3 vecRadJ = -re * en;
                const V3 centerDistJ = vecRadJ + (xpe - xpe);  // particle = element centre (single spheres)
                FW = FW + normalForce;
                const V3 relVelContact = relVel + wpg.cross(vecRadJ);
                const V3 tangRelVelContact = relVelContact - normalRelVel;
                const double nt = tangRelVelContact.norm();
                if (nt != 0.0) {
                    const double ft = tangential_contact(p, nt, fn, re, me, p.fricWall);
                    const V3 et = tangRelVelContact / tangRelVelContact.norm();
                    const V3 tangForce = ft * et;
                    MW = MW - centerDistJ.cross(tangForce);
                    FW = FW - tangForce;
                }
            }
        }
    }
    // Newton-Euler
    const V3 xp1 = vpe;
    const double rad = re;
    const V3 FVisc = -6.0 * M_PI * p.numVisc * rad * xp1;
    const V3 MVisc = -8.0 * M_PI * p.numVisc * rad * rad * rad * wpg;
    const V3 demF(p.demF[0], p.demF[1], p.demF[2]);
    const V3 x2 = (FVisc + FH + FP + FW) / me + demF;
    const V3 moment = MVisc + MH + MP + MW;
    const Q4 qp0 = qp[0];
    const V3 momentBf = project(moment, qp0.adjoint());
    const V3 waBf = newtonAcc(momentBf, Ie, wpl);
    const V3 w1 = project(waBf, qp0);
    E.set3(EF_FP, e, FP); E.set3(EF_MP, e, MP); E.set3(EF_FW, e, FW); E.set3(EF_MW, e, MW);
    // corrector
    const double* co1 = p.co1;
    const double* co2 = p.co2;
    const V3 x2Corr = x2 - xp[2];
    V3 x[6];
    x[0] = xp[0] + x2Corr * co2[0];
    x[1] = xp[1] + x2Corr * co2[1];
    x[2] = x2;
    x[3] = xp[3] + x2Corr * co2[3];
    x[4] = xp[4] + x2Corr * co2[4];
    x[5] = xp[5] + x2Corr * co2[5];
    const V3 w1Corr = w1 - wp[1];
    V3 w[6];
    w[0] = wp[0] + w1Corr * co1[0];
    w[1] = w1;
    w[2] = wp[2] + w1Corr * co1[2];
    w[3] = wp[3] + w1Corr * co1[3];
    w[4] = wp[4] + w1Corr * co1[4];
    w[5] = wp[5] + w1Corr * co1[5];
    const Q4 q2Corr = q2 - qp[2];
    Q4 q[6];
    q[0] = qp[0] + q2Corr * co2[0];
    q[1] = qp[1] + q2Corr * co2[1];
    q[2] = q2;
    q[3] = qp[3] + q2Corr * co2[3];
    q[4] = qp[4] + q2Corr * co2[4];
    q[5] = qp[5] + q2Corr * co2[5];
    q[0].normalize();
    for (int k = 0; k < 6; ++k) {
        E.set3(EF_X + 3 * k, e, x[k]);
        E.set3(EF_W + 3 * k, e, w[k]);
        if (k != 2) E.set4(EF_Q + 4 * k, e, q[k]);
    }
    E.set3(EF_WG, e, w[0]);
    E.set3(EF_WL, e, project(w[0], q[0].adjoint()));
    // predictor of the next sub-step
    predict_elem(E, p, e, x, q, w);
    return x[1].norm2();
}

// Coupling records of every particle slot after the DEM step: corrected
// position (ghosts: origin + period vector, elmt.cpp:292-298), element velocity
// and angular velocity, radius.
__global__ void k_dem_couple_table(DEMParams p, DEMBuffers b, CoupleRec* __restrict__ cp, CoupleRecF* __restrict__ cpf,
                                   int* __restrict__ nslots, Units u) {
#pragma clang fp contract(off)
    const El E{b.el, p.P};
    const int S = b.st->nSlots;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) *nslots = S;
    if (s >= S) return;
    int o = s;
    V3 x0 = E.v3(EF_X, s < p.P ? s : 0);
    if (s >= p.P) {
        const int g = s - p.P;
        o = (int)b.ghostIndex[g];
        x0 = E.v3(EF_X, o) + V3(b.ghostVec[3 * g], b.ghostVec[3 * g + 1], b.ghostVec[3 * g + 2]);
    }
    CoupleRec r;
    r.x0[0] = x0.x; r.x0[1] = x0.y; r.x0[2] = x0.z;
    r.rv[0] = r.rv[1] = r.rv[2] = 0.0;
    const V3 v = E.v3(EF_X + 3, o), w = E.v3(EF_WG, o);
    r.v[0] = v.x; r.v[1] = v.y; r.v[2] = v.z;
    r.w[0] = w.x; r.w[1] = w.y; r.w[2] = w.z;
    r.r = E.at(EF_RAD, o);
    r.elem = o;
    r.pad = 0;
    cp[s] = r;
    CoupleRecF q;
    const V3 cc = x0 / u.Length;  // radiusVec is zero for single spheres
    const V3 vs = v / u.Speed, ws = w / u.AngVel;
    q.c[0] = (float)cc.x; q.c[1] = (float)cc.y; q.c[2] = (float)cc.z;
    q.v[0] = (float)vs.x; q.v[1] = (float)vs.y; q.v[2] = (float)vs.z;
    q.w[0] = (float)ws.x; q.w[1] = (float)ws.y; q.w[2] = (float)ws.z;
    q.pad[0] = q.pad[1] = q.pad[2] = 0.0f;
    cpf[s] = q;
}

// Element hydrodynamic force and torque (LB.cpp:1897-1918): the run sums the
// collision appended for the element's own slot and for its ghosts' slots,
// added exactly (fixed point, so the result does not depend on the order of the
// appends), then scaled to physical units. One wavefront per element; lane l
// accumulates component l % 7 of every ninth entry (l / 7 < 9), so that each
// entry component is split once; the nine partial sums per component are then
// added through LDS. The lists are emptied for the next LB step.
constexpr int COMBINE_BLOCK = 256;
constexpr int COMBINE_ROWS = 9;  // entries per pass (9 x 7 components = 63 lanes)
__global__ void __launch_bounds__(COMBINE_BLOCK) k_dem_hydro_combine(DEMParams p, DEMBuffers b, HydroSink hs, Units u) {
#pragma clang fp contract(off)
    __shared__ long long red[COMBINE_BLOCK / 64][64][4];
    __shared__ double val[COMBINE_BLOCK / 64][7];
    const int lane = (int)(threadIdx.x & 63);
    const int w = (int)(threadIdx.x >> 6);
    const int e = (int)((blockIdx.x * (unsigned)COMBINE_BLOCK + threadIdx.x) >> 6);
    if (e >= p.P) return;  // whole wavefronts
    const El E{b.el, p.P};
    const int ng = b.ghostOfCount[e] < 8 ? b.ghostOfCount[e] : 8;
    const int q = lane % 7, j = lane / 7;
    long long acc[4] = {0, 0, 0, 0};
    for (int k = -1; k < ng; ++k) {
        const uint32_t slot = k < 0 ? (uint32_t)e : (uint32_t)b.ghostOf[e * 8 + k];
        const uint32_t cnt = hs.count[slot];
        const uint32_t n = cnt < hs.cap ? cnt : hs.cap;
        const HydroSeg* L = hs.segs + (size_t)slot * hs.cap;
        if (j < COMBINE_ROWS) {
            constexpr uint32_t AHEAD = 4;  // loads issued before the first split
            for (uint32_t i = (uint32_t)j; i < n; i += AHEAD * COMBINE_ROWS) {
                double v[AHEAD];
#pragma unroll
                for (uint32_t a = 0; a < AHEAD; ++a) {
                    const uint32_t ia = i + a * COMBINE_ROWS;
                    v[a] = ia < n ? L[ia].s[q] : 0.0;
                }
#pragma unroll
                for (uint32_t a = 0; a < AHEAD; ++a) exact_local(acc, v[a]);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) red[w][lane][i] = acc[i];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // every lane has read the counts above; empty the lists for the next step
    if (lane <= ng) hs.count[lane == 0 ? (uint32_t)e : (uint32_t)b.ghostOf[e * 8 + lane - 1]] = 0u;
    if (lane < 7) {
        long long t[4] = {0, 0, 0, 0};
        for (int r = 0; r < COMBINE_ROWS; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i) t[i] += red[w][lane + 7 * r][i];
        val[w][lane] = exact_value((const unsigned long long*)t);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane != 0) return;
    const double* v = val[w];
    E.set3(EF_FH, e, V3(v[0], v[1], v[2]) * u.Force);
    E.set3(EF_MH, e, V3(v[3], v[4], v[5]) * u.Torque);
    E.at(EF_FVOL, e) = v[6] * u.Volume;
}

}  // namespace

void launch_dem_rebuild(const DEMParams& p, const DEMBuffers& b, int force, hipStream_t s) {
    k_dem_rebuild<<<1, RB, 0, s>>>(p, b, force);
}

// particleEnergy after a sub-step (DEM.cpp:2226-2276, utils.cpp:324-327). The
// per-element terms of a chunk are computed in parallel; one lane then adds them
// in element order, so the sums are the reference's sequential sums.
constexpr int EN_BLOCK = 256;
__global__ void __launch_bounds__(EN_BLOCK) k_dem_energy(DEMParams p, DEMBuffers b, double sx, double sy, double sz,
                                                         double* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double tk[EN_BLOCK], rk[EN_BLOCK], gp[EN_BLOCK];
    const El E{b.el, p.P};
    const V3 F(p.demF[0], p.demF[1], p.demF[2]);
    const double gn = sqrt(F.norm2());
    const bool grav = gn != 0.0;
    V3 n, c(0.0, 0.0, 0.0);
    if (grav) {
        n = V3(-1.0 * F.x / gn, -1.0 * F.y / gn, -1.0 * F.z / gn);
        // the plane through the box corner the gravity points to (n . X < 0 etc.)
        if (n.x * 1.0 + n.y * 0.0 + n.z * 0.0 < 0) c.x = c.x + sx;
        if (n.x * 0.0 + n.y * 1.0 + n.z * 0.0 < 0) c.y = c.y + sy;
        if (n.x * 0.0 + n.y * 0.0 + n.z * 1.0 < 0) c.z = c.z + sz;
    }
    double tKin = 0.0, rKin = 0.0, g = 0.0;
    for (int base = 0; base < p.P; base += EN_BLOCK) {
        const int e = base + threadIdx.x;
        if (e < p.P) {
            const double m = E.at(EF_M, e);
            const V3 v = E.v3(EF_X + 3, e);
            tk[threadIdx.x] = 0.5 * m * v.norm2();
            const V3 w = E.v3(EF_WL, e), I = E.v3(EF_I, e);
            rk[threadIdx.x] = I.dot(V3(w.x * w.x, w.y * w.y, w.z * w.z));
            if (grav) {
                const V3 x = E.v3(EF_X, e);
                const double h = n.dot(x - c);
                gp[threadIdx.x] = m * h * gn;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int cnt = min(EN_BLOCK, p.P - base);
            for (int i = 0; i < cnt; ++i) {
                tKin += tk[i];
                rKin += rk[i];
                if (grav) g += gp[i];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double elastic = 0.0;  // not computed by the reference
        out[0] = rKin + tKin;
        out[1] = rKin;
        out[2] = tKin;
        out[3] = g;
        out[4] = elastic;
        out[5] = rKin + tKin + g + elastic;
    }
}

void launch_dem_energy(const DEMParams& p, const DEMBuffers& b, const double demSize[3], double* out, hipStream_t s) {
    k_dem_energy<<<1, EN_BLOCK, 0, s>>>(p, b, demSize[0], demSize[1], demSize[2], out);
}

namespace {
// Particle state of a saving sub-step, after the predictor (DEM.cpp:349-359:
// particles take the predicted element position and velocity; wGlobal is the
// last corrected angular velocity).
__global__ void k_dem_save_elements(DEMParams p, DEMBuffers b, double* __restrict__ out, double* __restrict__ info) {
    const El E{b.el, p.P};
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0) info[0] = (double)b.st->nPairs;
    if (e >= p.P) return;
    double* o = out + (size_t)e * 9;
    for (int k = 0; k < 3; ++k) {
        o[k] = E.at(EF_XP + k, e);
        o[3 + k] = E.at(EF_XP + 3 + k, e);
        o[6 + k] = E.at(EF_WG + k, e);
    }
}
}  // namespace

void launch_dem_save_elements(const DEMParams& p, const DEMBuffers& b, double* out, double* info, hipStream_t s) {
    k_dem_save_elements<<<(p.P > 0 ? p.P + 63 : 64) / 64, 64, 0, s>>>(p, b, out, info);
}

void launch_dem_predict(const DEMParams& p, const DEMBuffers& b, hipStream_t s) {
    if (p.P <= 0) return;
    k_dem_predict<<<(p.P + 63) / 64, 64, 0, s>>>(p, b);
}

void launch_dem_pairs(const DEMParams& p, const DEMBuffers& b, hipStream_t s) {
    if (p.P <= 0) return;
    int blocks = (p.pairCap + 255) / 256;
    if (blocks > 512) blocks = 512;
    k_dem_pairs<<<blocks, 256, 0, s>>>(p, b);
}

void launch_dem_gather(const DEMParams& p, const DEMBuffers& b, hipStream_t s) {
    if (p.P <= 0) return;
    k_dem_gather<<<(p.P + 63) / 64, 64, 0, s>>>(p, b);
}



void launch_dem_couple_table(const DEMParams& p, const DEMBuffers& b, CoupleRec* cp, CoupleRecF* cpf, int* nslots,
                             const Units& u, hipStream_t s) {
    const int cap = p.P + p.Gcap;
    k_dem_couple_table<<<(cap + 255) / 256 > 0 ? (cap + 255) / 256 : 1, 256, 0, s>>>(p, b, cp, cpf, nslots, u);
}

void launch_dem_hydro_combine(const DEMParams& p, const DEMBuffers& b, const HydroSink& hs, const Units& u,
                              hipStream_t s) {
    if (p.P <= 0) return;
    const int waves_per_block = COMBINE_BLOCK / 64;
    k_dem_hydro_combine<<<(p.P + waves_per_block - 1) / waves_per_block, COMBINE_BLOCK, 0, s>>>(p, b, hs, u);
}

}  // namespace hyb
