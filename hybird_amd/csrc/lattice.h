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

__device__ __forceinline__ bool interior(const LBKParams& p, const Coord& c) {
    return c.x > 0 && c.x < p.X - 1 && c.y > 0 && c.y < p.Y - 1 && c.z > 0 && c.z < p.Z - 1;
}

}  // namespace hyb
