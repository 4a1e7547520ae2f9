// Discrete-element kernels for gfx950.
//
// One DEM sub-step (DEM.cpp:331-376) is three launches on the handle's stream:
//  1. k_dem_rebuild (one workgroup): evalMaxDisp and, when the displacement
//     trigger fires, the whole neighbour rebuild of DEM.cpp:1369-1513 —
//     periodic shift, ghosts, linked-cell table, the 14-cell pair list and the
//     near-wall list — reproducing the reference's cell index (vector.cpp:147-149)
//     and unsigned stencil wrap, in the reference's pair order. It also builds a
//     per-particle list of the pairs each particle takes part in.
//  2. k_dem_pairs: one thread per pair evaluates the contact (DEM.cpp:1801-1900,
//     2138-2222) and writes each particle's signed increments into that
//     particle's slot of the pair (and, on saving sub-steps, its statistics).
//  3. k_dem_gather: per element, the increments in pair-list order (the same
//     order as the reference's sequential loop, so no atomics and bit-identical
//     sums; an element of several spheres merges its particles' lists by pair
//     index), the wall, object and cylinder contacts of its particles,
//     Newton-Euler (DEM.cpp:1149-1181), the Gear corrector (elmt.cpp:179-254),
//     the particle update (elmt.cpp:261-290) and the predictor (elmt.cpp:139-177)
//     of the next sub-step. An initial predictor runs once after initialisation.
// The corrector keeps the predicted copies it has just used until the new
// prediction overwrites them, so partner reads in k_dem_pairs always see the
// predicted state of the current sub-step.
//
// Single-sphere elements are their own particle (particle i = element i, the
// element arrays serve both). With multi-sphere elements (DEMParams::clumps)
// particles have their own state arrays, refreshed from their element after the
// corrector and after the predictor.
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

// Element (clusterIndex) of standard particle i.
__device__ __forceinline__ int part_elem(const DEMParams& p, const DEMBuffers& b, int i) {
    return p.clumps ? b.pelem[i] : i;
}

// Corrected position of standard particle i (particle::updateCorrected, elmt.cpp:277-290).
__device__ __forceinline__ V3 part_x(const El& E, const DEMParams& p, const DEMBuffers& b, int i) {
    if (p.clumps) {
        const double* c = b.pcor + 9 * (size_t)i;
        return V3(c[0], c[1], c[2]);
    }
    return E.v3(EF_X, i);
}

// Position of a particle slot: standard particles, then ghosts = origin + period vector (elmt.cpp:292-298).
__device__ __forceinline__ V3 slot_pos(const El& E, const DEMParams& p, const DEMBuffers& b, int s) {
    if (s < p.P) return part_x(E, p, b, s);
    const int g = s - p.P;
    const int o = (int)b.ghostIndex[g];
    return part_x(E, p, b, o) + V3(b.ghostVec[3 * g], b.ghostVec[3 * g + 1], b.ghostVec[3 * g + 2]);
}

__device__ __forceinline__ int slot_cluster(const DEMParams& p, const DEMBuffers& b, int s) {
    return part_elem(p, b, s < p.P ? s : (int)b.ghostIndex[s - p.P]);
}

// particle::updateCorrected / updatePredicted (elmt.cpp:261-290) for the
// particles of element e, from the element's position x0, velocity x1,
// orientation q and angular velocity w (the corrected or the predicted set):
// position x0 + r project(prototype, q), radiusVec = position - x0, velocity
// x1 + w x radiusVec; the particle of a size-1 element sits at its centre.
// out: DEMBuffers::pcor or ppred.
__device__ __forceinline__ void place_particles(const DEMBuffers& b, double* out, int e, double r, const V3& x0,
                                                const V3& x1, const Q4& q, const V3& w) {
#pragma clang fp contract(off)
    const int lo = b.efirst[e], hi = b.efirst[e + 1];
    for (int i = lo; i < hi; ++i) {
        V3 px = x0, rv(0.0, 0.0, 0.0), pv = x1;
        if (hi - lo > 1) {
            const V3 proto(b.pproto[3 * (size_t)i], b.pproto[3 * (size_t)i + 1], b.pproto[3 * (size_t)i + 2]);
            px = px + r * project(proto, q);
            rv = px - x0;
            pv = pv + w.cross(rv);
        }
        double* o = out + 9 * (size_t)i;
        o[0] = px.x; o[1] = px.y; o[2] = px.z;
        o[3] = pv.x; o[4] = pv.y; o[5] = pv.z;
        o[6] = rv.x; o[7] = rv.y; o[8] = rv.z;
    }
}

// Both particle states of element e from the element arrays.
__device__ __forceinline__ void place_both(const El& E, const DEMBuffers& b, int e) {
    const double r = E.at(EF_RAD, e);
    place_particles(b, b.pcor, e, r, E.v3(EF_X, e), E.v3(EF_X + 3, e), E.q4(EF_Q, e), E.v3(EF_WG, e));
    place_particles(b, b.ppred, e, r, E.v3(EF_XP, e), E.v3(EF_XP + 3, e), E.q4(EF_QP, e), E.v3(EF_WPG, e));
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
    // the same visit on a record already loaded (s1: position and cluster of n1)
    const auto visit_rec = [&](int n0, const V3& x0, int c0, int n1, const double s1[4]) {
        if (cell0 != cell1 || n1 < n0) {
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
            // cell1's slots WB at a time: their slot numbers, then their records, are loaded together
            constexpr int WB = 8;
            for (int i1 = a1; i1 < e1; i1 += WB) {
                int n1[WB];
                double s1[WB][4];
#pragma unroll
                for (int k = 0; k < WB; ++k) n1[k] = i1 + k < e1 ? b.cursor[i1 + k] : -1;
#pragma unroll
                for (int k = 0; k < WB; ++k)
                    if (n1[k] >= 0) {
                        const double* q = b.spos + 4 * (size_t)n1[k];
                        s1[k][0] = q[0]; s1[k][1] = q[1]; s1[k][2] = q[2]; s1[k][3] = q[3];
                    }
#pragma unroll
                for (int k = 0; k < WB; ++k)
                    if (n1[k] >= 0) visit_rec(n0, x0, c0, n1[k], s1[k]);
            }
        }
    }
    return cnt;
}

__global__ void __launch_bounds__(RB) k_dem_rebuild(DEMParams p, DEMBuffers b, int force) {
#pragma clang fp contract(off)
    const El E{b.el, p.E};
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
        // max |x1|^2 over elements, left by the corrector (and cleared here for the
        // next one); a NaN or infinite maximum stops the DEM (error 6)
        if (t == 0) {
            unsigned long long* vb = reinterpret_cast<unsigned long long*>(b.vmax);
            const double m = __longlong_as_double((long long)*vb);
            *vb = 0ull;
            if (!(m <= 1.79769313486231570e308)) {
                sh_err = DEM_ERR_NUMERIC;
            } else {
                const double maxVel = sqrt(m);
                b.st->maxDisp += maxVel * p.deltat;
                doit = b.st->maxDisp > p.trigger;
                if (doit) b.st->maxDisp = 0.0;
            }
        }
        __syncthreads();
        if (sh_err) { if (t == 0) dem_error(p, b, DEM_ERR_NUMERIC); return; }
    } else if (t == 0) {
        doit = 1;
    }
    __syncthreads();
    if (!doit || P == 0) return;
    if (t == 0) b.st->rebuilds += 1;
    // tangential springs: keep the old pair list, springs, per-particle lists and
    // ghosts; the new pairs take over the springs of the same element pairs below
    __shared__ int oldT, oldG;
    if (p.spring) {
        if (t == 0) { oldT = b.st->nPairs; oldG = b.st->nGhost; }
        __syncthreads();
        for (int i = t; i < 2 * oldT; i += RB) b.pairsOld[i] = b.pairs[i];
        for (int i = t; i < 3 * oldT; i += RB) b.histOld[i] = b.hist[i];
        for (int i = t; i <= P; i += RB) b.csrOffOld[i] = b.csrOff[i];
        const int nc = (int)b.csrOff[P];
        for (int i = t; i < nc; i += RB) b.csrOld[i] = b.csr[i];
        for (int i = t; i < oldG; i += RB) b.ghostIndexOld[i] = b.ghostIndex[i];
        __syncthreads();
    }
    int G = 0;
    if (p.npbc) {
        // ---- pbcShift (DEM.cpp:1554-1584)
        // The predictor of this sub-step already ran (fused into the previous
        // corrector), so a shifted element gets its predicted position
        // recomputed from the shifted x0 exactly as the predictor would, and
        // its particles are placed again (updateParticlesCorrected, DEM.cpp:1398).
        for (int e = t; e < p.E; e += RB) {
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
                if (p.clumps) place_both(E, b, e);
            }
        }
        __syncthreads();
        // ---- side ghosts in (pbc, element) order (DEM.cpp:1590-1618)
        int nSide = 0;
        for (int k = 0; k < p.npbc; ++k) {
            const PbcDev& pb = p.pbc[k];
            for (int e = t; e < P; e += RB) {  // e: standard particle
                const V3 x0 = part_x(E, p, b, e);
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
        const int c = linearize(slot_pos(E, p, b, s), p.cellWidth, p.nCells) + S;
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
        const V3 x = slot_pos(E, p, b, s);
        double* r = b.spos + 4 * (size_t)s;
        r[0] = x.x;
        r[1] = x.y;
        r[2] = x.z;
        r[3] = (double)slot_cluster(p, b, s);
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
    // ---- near-wall list (DEM.cpp:1496-1513): the first wall within range of each standard particle
    // (all walls within range as a bit mask with DEMParams::allWalls)
    int nw = 0;
    for (int e = t; e < P; e += RB) {
        const V3 x0 = part_x(E, p, b, e);
        uint32_t wi = 0xFFFFFFFFu, mask = 0u;
        for (int k = 0; k < p.nwalls; ++k)
            if (p.walls[k].dist(x0) < p.nebrRange) {
                if (wi == 0xFFFFFFFFu) wi = (uint32_t)k;
                mask |= 1u << k;
                if (!p.allWalls) break;
            }
        b.nearWall[e] = wi;
        if (p.allWalls) b.nearWallMask[e] = mask;
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
            const V3 x0 = part_x(E, p, b, e);
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
    if (p.spring) {
        // a new pair takes the spring of the old pair of the same two standard
        // particles (ghosts count as their origin; found in the old lists of
        // either particle), turned to the new pair's orientation; pairs without
        // one start with a released spring
        for (int k = t; k < T; k += RB) {
            const int n0 = (int)b.pairs[2 * k], n1 = (int)b.pairs[2 * k + 1];
            const int e0 = n0 < P ? n0 : (int)b.ghostIndex[n0 - P];
            const int e1 = n1 < P ? n1 : (int)b.ghostIndex[n1 - P];
            double xi[3] = {0.0, 0.0, 0.0};
            bool found = false;
            for (int side = 0; side < 2 && !found; ++side) {
                const int e = side == 0 ? e0 : e1;
                for (uint32_t i = b.csrOffOld[e]; i < b.csrOffOld[e + 1]; ++i) {
                    const uint32_t kk = b.csrOld[i];
                    const int m0 = (int)b.pairsOld[2 * kk], m1 = (int)b.pairsOld[2 * kk + 1];
                    const int f0 = m0 < P ? m0 : (int)b.ghostIndexOld[m0 - P];
                    const int f1 = m1 < P ? m1 : (int)b.ghostIndexOld[m1 - P];
                    double sgn = 0.0;
                    if (f0 == e0 && f1 == e1) sgn = 1.0;
                    else if (f0 == e1 && f1 == e0) sgn = -1.0;
                    if (sgn == 0.0) continue;
                    for (int c = 0; c < 3; ++c) xi[c] = sgn * b.histOld[3 * (size_t)kk + c];
                    found = true;
                    break;
                }
            }
            for (int c = 0; c < 3; ++c) b.hist[3 * (size_t)k + c] = xi[c];
        }
    }
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

// Tangential stiffness and damping coefficient of a contact (DEM.cpp:2176-2222).
__device__ __forceinline__ void tangential_coeffs(const DEMParams& p, double fn, double effRad, double effMass,
                                                  double& ks, double& gammas) {
    if (p.hertzian) ks = p.ksConst * sqrt(effRad) * pow(fabs(fn), 1.0 / 3.0);
    else ks = p.linearStiff;
    gammas = 2.0 * p.viscTang * sqrt(effMass * ks);
}

__device__ __forceinline__ double tangential_contact(const DEMParams& p, double vrelt, double fn, double effRad,
                                                     double effMass, double friction) {  // DEM.cpp:2176-2222
    double ks, gammas;
    tangential_coeffs(p, fn, effRad, effMass, ks, gammas);
    const double fsMax = friction * fn;
    const double dumpfs = gammas * vrelt;
    return (fsMax < dumpfs) ? fsMax : dumpfs;  // std::min(dumpfs, fsMax)
}

// Tangential spring (extension, not in the reference): the spring xi of the
// contact is turned onto the current tangent plane (normal en), advanced by the
// tangential relative velocity vt over one sub-step, and gives F = k_s xi +
// gamma_s vt with the viscous law's k_s and gamma_s; above friction x fn the
// force is scaled to that limit and xi set to the spring length that gives it.
// Returns F (the force on the contact's first body, along vt while sticking).
__device__ __forceinline__ V3 spring_force(const DEMParams& p, double fn, double effRad, double effMass,
                                           double friction, const V3& en, const V3& vt, double xi[3]) {
#pragma clang fp contract(off)
    double ks, gammas;
    tangential_coeffs(p, fn, effRad, effMass, ks, gammas);
    V3 s(xi[0], xi[1], xi[2]);
    s = s - en * s.dot(en);
    s = s + vt * p.deltat;
    V3 F = s * ks + vt * gammas;
    const double fsMax = friction * fn;
    const double fm = F.norm();
    if (fm > fsMax) {
        F = fm > 0.0 ? F * (fsMax / fm) : V3(0.0, 0.0, 0.0);
        s = ks > 0.0 ? (F - vt * gammas) / ks : V3(0.0, 0.0, 0.0);
    }
    xi[0] = s.x; xi[1] = s.y; xi[2] = s.z;
    return F;
}

// Predicted position, velocity and radiusVec of a particle slot and its element
// (updateParticlesPredicted, DEM.cpp:1184-1204; a ghost is its origin moved by
// the period vector, with the origin's velocity and radiusVec, elmt.cpp:292-298).
template <bool CL>
__device__ __forceinline__ void slot_pred(const El& E, const DEMParams& p, const DEMBuffers& b, int s, V3& x, V3& v,
                                          V3& rv, int& elem) {
    const int o = s < p.P ? s : (int)b.ghostIndex[s - p.P];
    if constexpr (CL) {
        const double* c = b.ppred + 9 * (size_t)o;
        x = V3(c[0], c[1], c[2]);
        v = V3(c[3], c[4], c[5]);
        rv = V3(c[6], c[7], c[8]);
        elem = b.pelem[o];
    } else {
        x = E.v3(EF_XP, o);
        v = E.v3(EF_XP + 3, o);
        rv = V3(0.0, 0.0, 0.0);
        elem = o;
    }
    if (s >= p.P) {
        const int g = s - p.P;
        x = x + V3(b.ghostVec[3 * g], b.ghostVec[3 * g + 1], b.ghostVec[3 * g + 2]);
    }
}

// Number of spheres of element e.
template <bool CL>
__device__ __forceinline__ int elem_size(const DEMBuffers& b, int e) {
    if constexpr (CL) return b.efirst[e + 1] - b.efirst[e];
    return 1;
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
// the signed increments for the pair's first (ri) and second (rj) particle's
// element, the normal-force moments of elements of size > 1 (mnI, mnJ; flag
// bit 2), and its statistics record st.
// SP: tangential spring (xi: the pair's spring); CL: multi-sphere elements present.
template <bool SP, bool CL>
__device__ __forceinline__ void pair_contact(const DEMParams& p, const DEMBuffers& b, const El& E, int k, PairRec& ri,
                                             PairRec& rj, StatRec& st, double* xi, V3& mnI, V3& mnJ) {
#pragma clang fp contract(off)
    const int n0 = (int)b.pairs[2 * k], n1 = (int)b.pairs[2 * k + 1];
    V3 xI, vI, rvI, xJ, vJ, rvJ;
    int eI, eJ;
    slot_pred<CL>(E, p, b, n0, xI, vI, rvI, eI);
    slot_pred<CL>(E, p, b, n1, xJ, vJ, rvJ, eJ);
    const double rI = E.at(EF_RAD, eI), rJ = E.at(EF_RAD, eJ);
    const double sigij = rI + rJ;
    const double sigij2 = sigij * sigij;
    const V3 vd = xJ - xI;
    ri.flag = rj.flag = 0;
    ri.pad = rj.pad = 0;
    st = StatRec{(double)n0, (double)n1, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0};
    if (!(vd.norm2() < sigij2)) {
        if constexpr (SP) xi[0] = xi[1] = xi[2] = 0.0;  // out of contact: the spring is released
        return;
    }
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
    const V3 centerDistI = vecRadI + rvI;
    const V3 centerDistJ = vecRadJ + rvJ;
    const V3 mn = -1.0 * normalForce;
    ri.fn[0] = mn.x; ri.fn[1] = mn.y; ri.fn[2] = mn.z;  // F - n == F + (-n)
    rj.fn[0] = normalForce.x; rj.fn[1] = normalForce.y; rj.fn[2] = normalForce.z;
    ri.flag = rj.flag = 1;
    if constexpr (CL) {  // only elements of size > 1 take the normal force's moment (DEM.cpp:1862-1867)
        if (elem_size<CL>(b, eI) > 1) {
            mnI = -1.0 * centerDistI.cross(normalForce);  // M - c x F == M + (-(c x F))
            ri.flag = 5;
        }
        if (elem_size<CL>(b, eJ) > 1) {
            mnJ = centerDistJ.cross(normalForce);
            rj.flag = 5;
        }
    }
    const V3 wI = E.v3(EF_WPG, eI), wJ = E.v3(EF_WPG, eJ);
    const V3 relVelContact = relVel - wI.cross(vecRadI) + wJ.cross(vecRadJ);
    const V3 tangRelVelContact = relVelContact - normalRelVel;
    double nt, ft;
    V3 et, tangForce;
    if constexpr (SP) {
        tangForce = spring_force(p, fn, effRad, effMass, p.fricPart, en, tangRelVelContact, xi);
        nt = tangForce.norm();
        ft = nt;
        if (nt != 0.0) et = tangForce / nt;
    } else {
        nt = tangRelVelContact.norm();
        if (nt != 0.0) {
            ft = tangential_contact(p, nt, fn, effRad, effMass, p.fricPart);
            et = tangRelVelContact / nt;
            tangForce = ft * et;
        }
    }
    if (nt != 0.0) {
        const V3 mI3 = centerDistI.cross(tangForce);
        const V3 mJ3 = -1.0 * centerDistJ.cross(tangForce);
        const V3 mt = -1.0 * tangForce;
        ri.ft[0] = tangForce.x; ri.ft[1] = tangForce.y; ri.ft[2] = tangForce.z;
        ri.m[0] = mI3.x; ri.m[1] = mI3.y; ri.m[2] = mI3.z;
        rj.ft[0] = mt.x; rj.ft[1] = mt.y; rj.ft[2] = mt.z;
        rj.m[0] = mJ3.x; rj.m[1] = mJ3.y; rj.m[2] = mJ3.z;
        ri.flag |= 2;
        rj.flag |= 2;
        set_stat(st, overlap, fn, ft, en, et);
    }
}

// One thread per neighbour pair (grid-stride): the contact at the predicted
// state; the signed increments for each of the two elements go into that
// element's slot of the pair in its csr list, so that k_dem_gather adds them in
// pair-list order.
template <bool SP, bool CL>
__global__ void __launch_bounds__(256) k_dem_pairs(DEMParams p, DEMBuffers b) {
    const El E{b.el, p.E};
    if (b.st->error) return;
    const int T = b.st->nPairs;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < T; k += gridDim.x * blockDim.x) {
        const uint32_t si = b.posI[k], sj = b.posJ[k];
        PairRec ri, rj;
        StatRec st;
        double xi[3];
        if constexpr (SP) {
            xi[0] = b.hist[3 * (size_t)k]; xi[1] = b.hist[3 * (size_t)k + 1]; xi[2] = b.hist[3 * (size_t)k + 2];
        }
        V3 mnI, mnJ;
        pair_contact<SP, CL>(p, b, E, k, ri, rj, st, xi, mnI, mnJ);
        if constexpr (SP) {
            b.hist[3 * (size_t)k] = xi[0]; b.hist[3 * (size_t)k + 1] = xi[1]; b.hist[3 * (size_t)k + 2] = xi[2];
        }
        if (b.pstat) b.pstat[k] = st;
        // a record without contact is read only for its flag: the other fields are not rewritten
        if (si != 0xFFFFFFFFu) {
            if (ri.flag) b.prec[si] = ri;
            else b.prec[si].flag = 0;
        }
        if (sj != 0xFFFFFFFFu) {
            if (rj.flag) b.prec[sj] = rj;
            else b.prec[sj].flag = 0;
        }
        if constexpr (CL) {
            if (si != 0xFFFFFFFFu && (ri.flag & 4)) {
                b.pmn[3 * (size_t)si] = mnI.x; b.pmn[3 * (size_t)si + 1] = mnI.y; b.pmn[3 * (size_t)si + 2] = mnI.z;
            }
            if (sj != 0xFFFFFFFFu && (rj.flag & 4)) {
                b.pmn[3 * (size_t)sj] = mnJ.x; b.pmn[3 * (size_t)sj + 1] = mnJ.y; b.pmn[3 * (size_t)sj + 2] = mnJ.z;
            }
        }
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

// Folds a wavefront's max |x1|^2 (>= 0, or NaN) into DEMBuffers::vmax: the bit
// patterns of non-negative doubles order like the values, and NaN patterns lie
// above every finite value, so an integer atomic max keeps the maximum (order-free).
__device__ __forceinline__ void vmax_add(const DEMBuffers& b, double v2) {
    if ((threadIdx.x & 63) == 0)
        atomicMax(reinterpret_cast<unsigned long long*>(b.vmax), (unsigned long long)__double_as_longlong(v2));
}

// Gear predictor of one quantity's six derivatives (elmt.cpp:139-177), with the
// position (c2) or angular-velocity (c1) coefficients.
template <class T>
__device__ __forceinline__ void gear_predict(const T x[6], const double* c, T o[6]) {
    o[0] = x[0] + x[1] * c[0] + x[2] * c[1] + x[3] * c[2] + x[4] * c[3] + x[5] * c[4];
    o[1] = x[1] + x[2] * c[0] + x[3] * c[1] + x[4] * c[2] + x[5] * c[3];
    o[2] = x[2] + x[3] * c[0] + x[4] * c[1] + x[5] * c[2];
    o[3] = x[3] + x[4] * c[0] + x[5] * c[1];
    o[4] = x[4] + x[5] * c[0];
    o[5] = x[5];
}

// Gear predictor of one element from its corrected state (elmt.cpp:139-177).
template <bool CL>
__device__ __forceinline__ void predict_elem(const El& E, const DEMParams& p, const DEMBuffers& b, int e, const V3 x[6],
                                             const Q4 q[6], const V3 w[6]) {
    V3 xp[6], wp[6];
    Q4 qp[6];
    gear_predict(x, p.c2, xp);
    gear_predict(q, p.c2, qp);
    qp[0].normalize();
    gear_predict(w, p.c1, wp);
    for (int k = 0; k < 6; ++k) {
        E.set3(EF_XP + 3 * k, e, xp[k]);
        E.set4(EF_QP + 4 * k, e, qp[k]);
        E.set3(EF_WP + 3 * k, e, wp[k]);
    }
    E.set3(EF_WPG, e, wp[0]);
    E.set3(EF_WPL, e, project(wp[0], qp[0].adjoint()));
    // particles from the predicted element (updateParticlesPredicted, DEM.cpp:1184-1204)
    if constexpr (CL) place_particles(b, b.ppred, e, E.at(EF_RAD, e), xp[0], xp[1], qp[0], wp[0]);
}

__global__ void __launch_bounds__(64) k_dem_predict(DEMParams p, DEMBuffers b) {
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const El E{b.el, p.E};
    double v2 = 0.0;
    if (e < p.E) {
        V3 x[6], w[6];
        Q4 q[6];
        for (int k = 0; k < 6; ++k) { x[k] = E.v3(EF_X + 3 * k, e); q[k] = E.q4(EF_Q + 4 * k, e); w[k] = E.v3(EF_W + 3 * k, e); }
        if (p.clumps) predict_elem<true>(E, p, b, e, x, q, w);
        else predict_elem<false>(E, p, b, e, x, q, w);
        v2 = x[1].norm2();
    }
    vmax_add(b, wave_max(v2));
}

// Per element: particle-contact increments in pair-list order, the wall contact
// (DEM.cpp:1692-1717, 1902-1982), Newton-Euler (DEM.cpp:1149-1181), the Gear
// corrector (elmt.cpp:179-254) and the predictor of the next sub-step.
template <bool SP, bool CL>
__device__ __forceinline__ double gather_elem(const DEMParams& p, const DEMBuffers& b, int e);

template <bool SP, bool CL>
__global__ void __launch_bounds__(64) k_dem_gather(DEMParams p, DEMBuffers b) {
    if (b.st->error) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const double v2 = e < p.E ? gather_elem<SP, CL>(p, b, e) : 0.0;
    vmax_add(b, wave_max(v2));
}

// Returns |x1|^2 of the corrected element for the next displacement check.
// SP: tangential spring; CL: multi-sphere elements present.
template <bool SP, bool CL>
__device__ __forceinline__ double gather_elem(const DEMParams& p, const DEMBuffers& b, int e) {
#pragma clang fp contract(off)
    const El E{b.el, p.E};
    V3 FP(0.0, 0.0, 0.0), MP(0.0, 0.0, 0.0), FW(0.0, 0.0, 0.0), MW(0.0, 0.0, 0.0);
    // the element's standard particles (consecutive; a single sphere is its own particle)
    const int ilo = CL ? b.efirst[e] : e, ihi = CL ? b.efirst[e + 1] : e + 1;
    const bool multi = CL && ihi - ilo > 1;
    // the state the contacts and Newton-Euler need; it does not depend on the
    // contact records, so its loads are issued first (the rest of the Gear state is
    // loaded group by group in the corrector, which keeps the registers in use low)
    const V3 xpe = E.v3(EF_XP, e);
    const V3 vpe = E.v3(EF_XP + 3, e);
    const double re = E.at(EF_RAD, e), me = E.at(EF_M, e);
    const V3 wpg = E.v3(EF_WPG, e);
    const V3 FH = E.v3(EF_FH, e), MH = E.v3(EF_MH, e);
    const V3 Ie = E.v3(EF_I, e), wpl = E.v3(EF_WPL, e);
    Q4 q2 = E.q4(EF_Q + 8, e);  // single spheres: q2 is never set by the force routine
    const Q4 qp0 = E.q4(EF_QP, e), qp1 = E.q4(EF_QP + 4, e);
    // contact increments in pair-list order
    const auto add = [&](const PairRec& r) {
        const int flag = r.flag;
        if (flag & 1) FP = FP + V3(r.fn[0], r.fn[1], r.fn[2]);
        if (flag & 2) {
            MP = MP + V3(r.m[0], r.m[1], r.m[2]);
            FP = FP + V3(r.ft[0], r.ft[1], r.ft[2]);
        }
    };
    if (!multi) {  // one list; records are loaded GU at a time
        const uint32_t lo = b.csrOff[ilo], hi = b.csrOff[ilo + 1];
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
    } else {
        // the lists of the element's particles merged by pair index: every pair
        // of the element appears in exactly one of them (pairs within one
        // element are never listed), so this is the reference's pair order
        constexpr int MAXS = 4;
        uint32_t cur[MAXS], end[MAXS];
        const int nl = ihi - ilo;
        for (int l = 0; l < MAXS; ++l) {
            cur[l] = l < nl ? b.csrOff[ilo + l] : 0u;
            end[l] = l < nl ? b.csrOff[ilo + l + 1] : 0u;
        }
        for (;;) {
            int best = -1;
            uint32_t bk = 0xFFFFFFFFu;
            for (int l = 0; l < nl; ++l)
                if (cur[l] < end[l]) {
                    const uint32_t k = b.csr[cur[l]];
                    if (k < bk) { bk = k; best = l; }
                }
            if (best < 0) break;
            const uint32_t i = cur[best]++;
            const PairRec r = b.prec[i];
            if (r.flag & 1) FP = FP + V3(r.fn[0], r.fn[1], r.fn[2]);
            if (r.flag & 4) MP = MP + V3(b.pmn[3 * (size_t)i], b.pmn[3 * (size_t)i + 1], b.pmn[3 * (size_t)i + 2]);
            if (r.flag & 2) {
                MP = MP + V3(r.m[0], r.m[1], r.m[2]);
                FP = FP + V3(r.ft[0], r.ft[1], r.ft[2]);
            }
        }
    }
    // predicted position and velocity of standard particle i of this element
    const auto part_pred = [&](int i, V3& px, V3& pv) {
        if constexpr (CL) {
            const double* c = b.ppred + 9 * (size_t)i;
            px = V3(c[0], c[1], c[2]);
            pv = V3(c[3], c[4], c[5]);
        } else {
            px = xpe;
            pv = vpe;
        }
    };
    // the contact of standard particle i (predicted px, pv) with wall wi (DEM.cpp:1902-1982): adds to FW, MW
    // and fills its force record and statistics; wxi: the spring (spring model)
    const auto wall_contact = [&](int i, uint32_t wi, const V3& px, const V3& pv, ContactForce& wcf, StatRec& wst,
                                  double wxi[3]) -> bool {
        const WallDev& w = p.walls[wi];
        const double distance = w.dist(px);
        const double overlap = re - distance;
        if (!(overlap > 0.0)) return false;
        const V3 en = w.n;
        const V3 cpv = w.getSpeed(px);
        const V3 relVel = pv - cpv;
        const double normRelVel = relVel.dot(en);
        const V3 normalRelVel = en * normRelVel;
        const double fn = normal_contact(p, 2.0 * overlap, normRelVel, re, me);
        const V3 normalForce = en * fn;
        const V3 vecRadJ = -re * en;
        V3 centerDistJ = vecRadJ;
        if (multi) centerDistJ = centerDistJ + (px - xpe);
        FW = FW + normalForce;
        if (multi) MW = MW + centerDistJ.cross(normalForce);
        set_force(wcf, normalForce, 1.0);
        const V3 relVelContact = relVel + wpg.cross(vecRadJ);
        const V3 tangRelVelContact = relVelContact - normalRelVel;
        if constexpr (SP) {
            const V3 tangForce = spring_force(p, fn, re, me, p.fricWall, en, tangRelVelContact, wxi);
            const double ft = tangForce.norm();
            if (ft != 0.0) {
                MW = MW - centerDistJ.cross(tangForce);
                FW = FW - tangForce;
                set_force_t(wcf, tangForce);
                if (b.wstat) set_stat(wst, overlap, fn, ft, en, tangForce / ft);
            }
        } else {
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
        return true;
    };
    // wall contacts of the element's particles (DEM.cpp:1692-1717): the first wall within range of each
    // particle, or every wall within range in wall order (DEMParams::allWalls; record i * nwalls + w)
    for (int i = ilo; i < ihi; ++i) {
        V3 px, pv;
        part_pred(i, px, pv);
        if (p.allWalls) {
            const uint32_t mask = b.nearWallMask[i];
            const int ws = p.nwalls;
            for (int wi = 0; wi < ws; ++wi) {
                const size_t slot = (size_t)i * ws + wi;
                StatRec wst{(mask >> wi & 1u) ? (double)wi : -1.0, (double)i, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0},
                            {0.0, 0.0, 0.0}, 0.0};
                ContactForce wcf{{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0, 0.0};
                // spring model: one spring per (particle, wall), kept while that wall stays in contact
                double wxi[3] = {0.0, 0.0, 0.0};
                bool wcontact = false;
                if (mask >> wi & 1u) {
                    if (SP && b.whist[4 * slot + 3] == (double)wi) {
                        wxi[0] = b.whist[4 * slot]; wxi[1] = b.whist[4 * slot + 1]; wxi[2] = b.whist[4 * slot + 2];
                    }
                    wcontact = wall_contact(i, (uint32_t)wi, px, pv, wcf, wst, wxi);
                }
                b.wcf[slot] = wcf;
                if (b.wstat) b.wstat[slot] = wst;
                if constexpr (SP) {
                    double* wh = b.whist + 4 * slot;
                    wh[0] = wcontact ? wxi[0] : 0.0;
                    wh[1] = wcontact ? wxi[1] : 0.0;
                    wh[2] = wcontact ? wxi[2] : 0.0;
                    wh[3] = wcontact ? (double)wi : -1.0;
                }
            }
            continue;
        }
        const uint32_t wi = b.nearWall[i];
        StatRec wst{wi == 0xFFFFFFFFu ? -1.0 : (double)wi, (double)i, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0},
                    {0.0, 0.0, 0.0}, 0.0};
        ContactForce wcf{{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0, 0.0};
        // spring against the near wall (spring model): kept while the same wall stays in contact
        double wxi[3] = {0.0, 0.0, 0.0};
        bool wcontact = false;
        if (SP && wi != 0xFFFFFFFFu && b.whist[4 * (size_t)i + 3] == (double)wi) {
            wxi[0] = b.whist[4 * (size_t)i]; wxi[1] = b.whist[4 * (size_t)i + 1]; wxi[2] = b.whist[4 * (size_t)i + 2];
        }
        if (wi != 0xFFFFFFFFu) {
            wcontact = wall_contact(i, wi, px, pv, wcf, wst, wxi);
            b.wcf[i] = wcf;
        }
        if constexpr (SP) {
            double* wh = b.whist + 4 * (size_t)i;
            wh[0] = wcontact ? wxi[0] : 0.0;
            wh[1] = wcontact ? wxi[1] : 0.0;
            wh[2] = wcontact ? wxi[2] : 0.0;
            wh[3] = wcontact ? (double)wi : -1.0;
        }
        if (b.wstat) b.wstat[i] = wst;
    }
    // object contacts (DEM.cpp:1747-1770, 2060-2136), in list order after the walls
    if (p.nobj) {
        for (int i = ilo; i < ihi; ++i) {
            V3 px, pv;
            part_pred(i, px, pv);
            const int no = b.nearObjCount[i];
            if (b.ostat)
                for (int k = no; k < NEAR_OBJ; ++k)
                    b.ostat[(size_t)i * NEAR_OBJ + k] = StatRec{-1.0, (double)i, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0},
                                                                {0.0, 0.0, 0.0}, 0.0};
            for (int k = 0; k < no; ++k) {
                const uint32_t oslot = b.nearObj[(size_t)i * NEAR_OBJ + k];
                const ObjDev& ob = b.objs[oslot];
                const V3 vectorDistance = px - ob.x0;
                const double distance = vectorDistance.norm();
                const double overlap = re + ob.r - distance;
                StatRec ost{(double)oslot, (double)i, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0};
                ContactForce ocf{{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0, 0.0};
                if (!(overlap > 0.0)) {
                    if (b.ostat) b.ostat[(size_t)i * NEAR_OBJ + k] = ost;
                    b.ocf[(size_t)i * NEAR_OBJ + k] = ocf;
                    continue;
                }
                const V3 en = (1.0 / distance) * vectorDistance;
                const V3 relVel = pv - ob.x1;
                const double normRelVel = relVel.dot(en);
                const V3 normalRelVel = en * normRelVel;
                const double fn = normal_contact(p, 2.0 * overlap, normRelVel, re, me);
                const V3 normalForce = en * fn;
                const V3 vecRadJ = -re * en;
                const V3 centerDistJ = vecRadJ + (px - xpe);  // zero offset for a single sphere
                FW = FW + normalForce;
                if (multi) MW = MW + centerDistJ.cross(normalForce);
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
                if (b.ostat) b.ostat[(size_t)i * NEAR_OBJ + k] = ost;
                b.ocf[(size_t)i * NEAR_OBJ + k] = ocf;
            }
        }
    }
    // cylinder contacts (DEM.cpp:1719-1744, 1984-2058), after the objects
    if (p.ncyl) {
        for (int i = ilo; i < ihi; ++i) {
            const uint32_t ci = b.nearCyl[i];
            if (ci == 0xFFFFFFFFu) continue;
            V3 px, pv;
            part_pred(i, px, pv);
            const CylDev& cyl = p.cyl[ci];
            const double distance = cyl.dist(px);
            const double overlap = re - distance;
            if (!(overlap > 0.0)) continue;
            const V3 vecDistance = cyl.vecDist(px);
            const V3 contactPoint = px - vecDistance;
            const V3 en = vecDistance / (re - overlap);
            const V3 contactPointVelocity = cyl.getSpeed(contactPoint);
            const V3 relVel = pv - contactPointVelocity;
            const double normRelVel = relVel.dot(en);
            const V3 normalRelVel = en * normRelVel;
            const double fn = normal_contact(p, 2.0 * overlap, normRelVel, re, me);
            const V3 normalForce = en * fn;
            const V3 vecRadJ = -re * en;
            const V3 centerDistJ = vecRadJ + (px - xpe);  // zero offset for a single sphere
            FW = FW + normalForce;
            if (multi) MW = MW + centerDistJ.cross(normalForce);
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
    // Newton-Euler
    const V3 xp1 = vpe;
    const double rad = re;
    const V3 FVisc = -6.0 * M_PI * p.numVisc * rad * xp1;
    const V3 MVisc = -8.0 * M_PI * p.numVisc * rad * rad * rad * wpg;
    const V3 demF(p.demF[0], p.demF[1], p.demF[2]);
    const V3 x2 = (FVisc + FH + FP + FW) / me + demF;
    const V3 moment = MVisc + MH + MP + MW;
    const V3 momentBf = project(moment, qp0.adjoint());
    const V3 waBf = newtonAcc(momentBf, Ie, wpl);
    const V3 w1 = project(waBf, qp0);
    if (multi) {  // rotational acceleration as a quaternion (DEM.cpp:1176-1179)
        const Q4 waQuat = quatAcc(waBf, qp1);
        q2 = 0.5 * qp0.multiply(waQuat);
    }
    E.set3(EF_FP, e, FP); E.set3(EF_MP, e, MP); E.set3(EF_FW, e, FW); E.set3(EF_MW, e, MW);
    // corrector (elmt.cpp:179-254) and the next sub-step's predictor, one quantity at a time
    const double* co1 = p.co1;
    const double* co2 = p.co2;
    V3 xc[2], xn[2];  // corrected and predicted position and velocity
    {
        const V3 xp[6] = {xpe, vpe, E.v3(EF_XP + 6, e), E.v3(EF_XP + 9, e), E.v3(EF_XP + 12, e), E.v3(EF_XP + 15, e)};
        const V3 x2Corr = x2 - xp[2];
        V3 x[6], o[6];
        x[0] = xp[0] + x2Corr * co2[0];
        x[1] = xp[1] + x2Corr * co2[1];
        x[2] = x2;
        x[3] = xp[3] + x2Corr * co2[3];
        x[4] = xp[4] + x2Corr * co2[4];
        x[5] = xp[5] + x2Corr * co2[5];
        gear_predict(x, p.c2, o);
        for (int k = 0; k < 6; ++k) {
            E.set3(EF_X + 3 * k, e, x[k]);
            E.set3(EF_XP + 3 * k, e, o[k]);
        }
        xc[0] = x[0]; xc[1] = x[1];
        xn[0] = o[0]; xn[1] = o[1];
    }
    Q4 qc0, qn0;  // corrected and predicted orientation
    {
        const Q4 qp[6] = {qp0, qp1, E.q4(EF_QP + 8, e), E.q4(EF_QP + 12, e), E.q4(EF_QP + 16, e), E.q4(EF_QP + 20, e)};
        const Q4 q2Corr = q2 - qp[2];
        Q4 q[6], o[6];
        q[0] = qp[0] + q2Corr * co2[0];
        q[1] = qp[1] + q2Corr * co2[1];
        q[2] = q2;
        q[3] = qp[3] + q2Corr * co2[3];
        q[4] = qp[4] + q2Corr * co2[4];
        q[5] = qp[5] + q2Corr * co2[5];
        q[0].normalize();
        gear_predict(q, p.c2, o);
        o[0].normalize();
        for (int k = 0; k < 6; ++k) {
            if (k != 2 || multi) E.set4(EF_Q + 4 * k, e, q[k]);
            E.set4(EF_QP + 4 * k, e, o[k]);
        }
        qc0 = q[0];
        qn0 = o[0];
    }
    V3 wc0, wn0;  // corrected and predicted angular velocity
    {
        const V3 wp[6] = {E.v3(EF_WP, e), E.v3(EF_WP + 3, e), E.v3(EF_WP + 6, e), E.v3(EF_WP + 9, e),
                          E.v3(EF_WP + 12, e), E.v3(EF_WP + 15, e)};
        const V3 w1Corr = w1 - wp[1];
        V3 w[6], o[6];
        w[0] = wp[0] + w1Corr * co1[0];
        w[1] = w1;
        w[2] = wp[2] + w1Corr * co1[2];
        w[3] = wp[3] + w1Corr * co1[3];
        w[4] = wp[4] + w1Corr * co1[4];
        w[5] = wp[5] + w1Corr * co1[5];
        gear_predict(w, p.c1, o);
        for (int k = 0; k < 6; ++k) {
            E.set3(EF_W + 3 * k, e, w[k]);
            E.set3(EF_WP + 3 * k, e, o[k]);
        }
        E.set3(EF_WG, e, w[0]);
        E.set3(EF_WL, e, project(w[0], qc0.adjoint()));
        E.set3(EF_WPG, e, o[0]);
        E.set3(EF_WPL, e, project(o[0], qn0.adjoint()));
        wc0 = w[0];
        wn0 = o[0];
    }
    // particles from the corrected and from the predicted element (updateParticlesCorrected and
    // updateParticlesPredicted, DEM.cpp:1184-1227)
    if constexpr (CL) {
        place_particles(b, b.pcor, e, re, xc[0], xc[1], qc0, wc0);
        place_particles(b, b.ppred, e, re, xn[0], xn[1], qn0, wn0);
    }
    return xc[1].norm2();
}

// Coupling records of every particle slot after the DEM step: corrected
// position and radiusVec (ghosts: origin + period vector with the origin's
// radiusVec, elmt.cpp:292-298), the element's velocity and angular velocity,
// radius.
__global__ void k_dem_couple_table(DEMParams p, DEMBuffers b, CoupleRec* __restrict__ cp, CoupleRecF* __restrict__ cpf,
                                   int* __restrict__ nslots, Units u) {
#pragma clang fp contract(off)
    const El E{b.el, p.E};
    const int S = b.st->nSlots;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) *nslots = S;
    if (s >= S) return;
    const int o = s < p.P ? s : (int)b.ghostIndex[s - p.P];  // standard particle
    const int e = part_elem(p, b, o);
    V3 x0, rv(0.0, 0.0, 0.0);
    if (p.clumps) {
        const double* c = b.pcor + 9 * (size_t)o;
        x0 = V3(c[0], c[1], c[2]);
        rv = V3(c[6], c[7], c[8]);
    } else {
        x0 = E.v3(EF_X, o);
    }
    if (s >= p.P) {
        const int g = s - p.P;
        x0 = x0 + V3(b.ghostVec[3 * g], b.ghostVec[3 * g + 1], b.ghostVec[3 * g + 2]);
    }
    CoupleRec r;
    r.x0[0] = x0.x; r.x0[1] = x0.y; r.x0[2] = x0.z;
    r.rv[0] = rv.x; r.rv[1] = rv.y; r.rv[2] = rv.z;
    const V3 v = E.v3(EF_X + 3, e), w = E.v3(EF_WG, e);
    r.v[0] = v.x; r.v[1] = v.y; r.v[2] = v.z;
    r.w[0] = w.x; r.w[1] = w.y; r.w[2] = w.z;
    r.r = E.at(EF_RAD, e);
    r.elem = e;
    r.pad = 0;
    cp[s] = r;
    CoupleRecF q;
    const V3 cc = x0 / u.Length - rv / u.Length;  // radius = node - cc (LB.cpp:1875)
    const V3 vs = v / u.Speed, ws = w / u.AngVel;
    q.c[0] = (float)cc.x; q.c[1] = (float)cc.y; q.c[2] = (float)cc.z;
    q.v[0] = (float)vs.x; q.v[1] = (float)vs.y; q.v[2] = (float)vs.z;
    q.w[0] = (float)ws.x; q.w[1] = (float)ws.y; q.w[2] = (float)ws.z;
    q.pad[0] = q.pad[1] = q.pad[2] = 0.0f;
    cpf[s] = q;
}

// Element hydrodynamic force and torque (LB.cpp:1897-1918): the run sums the
// collision appended for the slots of the element's particles and their ghosts,
// added exactly (fixed point, so the result does not depend on the order of the
// appends), then scaled to physical units. One wavefront per element, one
// element per workgroup; lane l accumulates component l % 7 of every ninth
// entry (l / 7 < 9), so that each entry component is split once; the nine
// partial sums per component are then added through LDS. The lists are emptied
// for the next LB step. (Four elements per workgroup, each wavefront on its own
// LDS rows, gave results that varied from run to run for a few elements past
// the first 1024 on this hardware; one wavefront per workgroup does not.)
constexpr int COMBINE_BLOCK = 64;
constexpr int COMBINE_ROWS = 9;  // entries per pass (9 x 7 components = 63 lanes)
__global__ void __launch_bounds__(COMBINE_BLOCK) k_dem_hydro_combine(DEMParams p, DEMBuffers b, HydroSink hs, Units u) {
#pragma clang fp contract(off)
    __shared__ long long red[COMBINE_BLOCK / 64][64][4];
    __shared__ double val[COMBINE_BLOCK / 64][7];
    const int lane = (int)(threadIdx.x & 63);
    const int w = (int)(threadIdx.x >> 6);
    const int e = (int)((blockIdx.x * (unsigned)COMBINE_BLOCK + threadIdx.x) >> 6);
    if (e >= p.E) return;  // whole wavefronts
    const El E{b.el, p.E};
    const int ilo = p.clumps ? b.efirst[e] : e, ihi = p.clumps ? b.efirst[e + 1] : e + 1;
    const int q = lane % 7, j = lane / 7;
    long long acc[4] = {0, 0, 0, 0};
    for (int i = ilo; i < ihi; ++i)
    for (int k = -1, ng = b.ghostOfCount[i] < 8 ? b.ghostOfCount[i] : 8; k < ng; ++k) {
        const uint32_t slot = k < 0 ? (uint32_t)i : (uint32_t)b.ghostOf[i * 8 + k];
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
    for (int i = ilo; i < ihi; ++i) {
        const int ng = b.ghostOfCount[i] < 8 ? b.ghostOfCount[i] : 8;
        if (lane <= ng) hs.count[lane == 0 ? (uint32_t)i : (uint32_t)b.ghostOf[i * 8 + lane - 1]] = 0u;
    }
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
    const El E{b.el, p.E};
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
    for (int base = 0; base < p.E; base += EN_BLOCK) {
        const int e = base + threadIdx.x;
        if (e < p.E) {
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
            const int cnt = min(EN_BLOCK, p.E - base);
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
// particles hold the predicted position and velocity; wGlobal is the element's
// last corrected angular velocity).
__global__ void k_dem_save_elements(DEMParams p, DEMBuffers b, double* __restrict__ out, double* __restrict__ info) {
    const El E{b.el, p.E};
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) info[0] = (double)b.st->nPairs;
    if (i >= p.P) return;
    const int e = part_elem(p, b, i);
    double* o = out + (size_t)i * 9;
    for (int k = 0; k < 3; ++k) {
        o[k] = p.clumps ? b.ppred[9 * (size_t)i + k] : E.at(EF_XP + k, e);
        o[3 + k] = p.clumps ? b.ppred[9 * (size_t)i + 3 + k] : E.at(EF_XP + 3 + k, e);
        o[6 + k] = E.at(EF_WG + k, e);
    }
}

// Both particle states of every element (set-up and state load).
__global__ void k_dem_particles(DEMParams p, DEMBuffers b) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p.E) return;
    place_both(El{b.el, p.E}, b, e);
}
}  // namespace

void launch_dem_save_elements(const DEMParams& p, const DEMBuffers& b, double* out, double* info, hipStream_t s) {
    k_dem_save_elements<<<(p.P > 0 ? p.P + 63 : 64) / 64, 64, 0, s>>>(p, b, out, info);
}

void launch_dem_predict(const DEMParams& p, const DEMBuffers& b, hipStream_t s) {
    if (p.E <= 0) return;
    k_dem_predict<<<(p.E + 63) / 64, 64, 0, s>>>(p, b);
}

void launch_dem_particles(const DEMParams& p, const DEMBuffers& b, hipStream_t s) {
    if (p.E <= 0 || !p.clumps) return;
    k_dem_particles<<<(p.E + 63) / 64, 64, 0, s>>>(p, b);
}

void launch_dem_pairs(const DEMParams& p, const DEMBuffers& b, hipStream_t s) {
    if (p.P <= 0) return;
    int blocks = (p.pairCap + 255) / 256;
    if (blocks > 512) blocks = 512;
    if (p.spring) {
        if (p.clumps) k_dem_pairs<true, true><<<blocks, 256, 0, s>>>(p, b);
        else k_dem_pairs<true, false><<<blocks, 256, 0, s>>>(p, b);
    } else {
        if (p.clumps) k_dem_pairs<false, true><<<blocks, 256, 0, s>>>(p, b);
        else k_dem_pairs<false, false><<<blocks, 256, 0, s>>>(p, b);
    }
}

void launch_dem_gather(const DEMParams& p, const DEMBuffers& b, hipStream_t s) {
    if (p.E <= 0) return;
    const int g = (p.E + 63) / 64;
    if (p.spring) {
        if (p.clumps) k_dem_gather<true, true><<<g, 64, 0, s>>>(p, b);
        else k_dem_gather<true, false><<<g, 64, 0, s>>>(p, b);
    } else {
        if (p.clumps) k_dem_gather<false, true><<<g, 64, 0, s>>>(p, b);
        else k_dem_gather<false, false><<<g, 64, 0, s>>>(p, b);
    }
}



void launch_dem_couple_table(const DEMParams& p, const DEMBuffers& b, CoupleRec* cp, CoupleRecF* cpf, int* nslots,
                             const Units& u, hipStream_t s) {
    const int cap = p.P + p.Gcap;
    k_dem_couple_table<<<(cap + 255) / 256 > 0 ? (cap + 255) / 256 : 1, 256, 0, s>>>(p, b, cp, cpf, nslots, u);
}

void launch_dem_hydro_combine(const DEMParams& p, const DEMBuffers& b, const HydroSink& hs, const Units& u,
                              hipStream_t s) {
    if (p.E <= 0) return;
    const int waves_per_block = COMBINE_BLOCK / 64;
    k_dem_hydro_combine<<<(p.E + waves_per_block - 1) / waves_per_block, COMBINE_BLOCK, 0, s>>>(p, b, hs, u);
}

}  // namespace hyb
