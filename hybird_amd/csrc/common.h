// Shared device/host helpers for the hybird_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace hyb {

// D3Q19 velocity set in the reference's order O, x, y, z, xy, yz, zx (lattice.h:36-60),
// opposite directions (lattice.h:85) and weights (lattice.h:98-101).
__host__ __device__ constexpr int cx(int j) {
    constexpr int v[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    return v[j];
}
__host__ __device__ constexpr int cy(int j) {
    constexpr int v[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0};
    return v[j];
}
__host__ __device__ constexpr int cz(int j) {
    constexpr int v[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    return v[j];
}
__host__ __device__ constexpr int opp(int j) {
    constexpr int v[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    return v[j];
}
__host__ __device__ constexpr double wgt(int j) {
    return j == 0 ? 12.0 / 36.0 : (j < 7 ? 2.0 / 36.0 : 1.0 / 36.0);
}

// Slip-wall reflection (lattice.h:87-95): for a diagonal link j, the axis
// directions whose neighbours decide the reflection and the populations
// reflected from them.
__host__ __device__ constexpr int slip1check(int j) {
    constexpr int v[19] = {0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 4, 3, 4, 5, 6, 1, 2, 6, 5};
    return v[j];
}
__host__ __device__ constexpr int slip1dir(int j) {
    constexpr int v[19] = {0, 0, 0, 0, 0, 0, 0, 9, 10, 8, 7, 13, 14, 12, 11, 18, 17, 15, 16};
    return v[j];
}
__host__ __device__ constexpr int slip2check(int j) {
    constexpr int v[19] = {0, 0, 0, 0, 0, 0, 0, 3, 4, 2, 1, 5, 6, 4, 3, 5, 6, 1, 2};
    return v[j];
}
__host__ __device__ constexpr int slip2dir(int j) {
    constexpr int v[19] = {0, 0, 0, 0, 0, 0, 0, 10, 9, 7, 8, 14, 13, 11, 12, 17, 18, 16, 15};
    return v[j];
}

// Node type codes (node.h:73-83).
enum NodeType : uint8_t { T_FLUID = 0, T_GAS = 2, T_INTERFACE = 3, T_PERIODIC = 4, T_SLIP = 5, T_SLIP_MOVING = 6,
                          T_WALL = 7, T_WALL_MOVING = 8, T_CURVED = 9 };

// Wall types the lattice streams against: slip (5), moving slip (6), no-slip (7),
// moving no-slip (8). Moving walls have the generic zero velocity
// (DEM.cpp:435-638), so they bounce back like static walls.
__host__ __device__ constexpr bool is_plane_wall(uint8_t t) { return t >= 5 && t <= 8; }
__host__ __device__ constexpr bool is_slip_wall(uint8_t t) { return t == 5 || t == 6; }

// Per-node link word: bit j (1..18) = the link in direction j is a wall;
// bit 0 = some of those walls are slip walls (their directions are in a
// separate per-node slip word);
// bit 29 = the free-surface pass has already streamed this node's populations
// for the next collision (they sit at the node's own index of the output
// buffer); bit 30 = node is interface; bit 31 = node is active (fluid or interface).
constexpr uint32_t LINK_ACTIVE = 1u << 31;
constexpr uint32_t LINK_INTERFACE = 1u << 30;
constexpr uint32_t LINK_PRESTREAMED = 1u << 29;
constexpr uint32_t LINK_CURVED = 1u << 28;  // some wall links end on curved (cylinder) wall nodes
constexpr uint32_t LINK_SLIP = 1u;
constexpr uint32_t LINK_WALLS = 0x7FFFEu;

// 3-vector with the reference's operation order (vector.cpp:38-180): every
// operation is component-wise, dot = x*x' + y*y' + z*z' evaluated left to right.
struct V3 {
    double x, y, z;
    __host__ __device__ V3() : x(0.0), y(0.0), z(0.0) {}
    __host__ __device__ V3(double a, double b, double c) : x(a), y(b), z(c) {}
    __host__ __device__ V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    __host__ __device__ V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    __host__ __device__ V3 operator*(double s) const { return V3(x * s, y * s, z * s); }
    __host__ __device__ V3 operator/(double s) const { return V3(x / s, y / s, z / s); }
    __host__ __device__ V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    __host__ __device__ V3& operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
    __host__ __device__ double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    __host__ __device__ V3 cross(const V3& o) const { return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
    __host__ __device__ double norm2() const { return x * x + y * y + z * z; }
    __host__ __device__ double norm() const { return sqrt(x * x + y * y + z * z); }
};
__host__ __device__ inline V3 operator*(double s, const V3& v) { return V3(v.x * s, v.y * s, v.z * s); }

struct Q4 {
    double q0, q1, q2, q3;
    __host__ __device__ Q4() : q0(1.0), q1(0.0), q2(0.0), q3(0.0) {}
    __host__ __device__ Q4(double a, double b, double c, double d) : q0(a), q1(b), q2(c), q3(d) {}
    __host__ __device__ Q4 operator+(const Q4& o) const { return Q4(q0 + o.q0, q1 + o.q1, q2 + o.q2, q3 + o.q3); }
    __host__ __device__ Q4 operator-(const Q4& o) const { return Q4(q0 - o.q0, q1 - o.q1, q2 - o.q2, q3 - o.q3); }
    __host__ __device__ Q4 operator*(double s) const { return Q4(q0 * s, q1 * s, q2 * s, q3 * s); }
    __host__ __device__ double norm() const { return sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3); }
    __host__ __device__ double norm2() const { return q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3; }
    __host__ __device__ void normalize() { const double n = norm(); q0 /= n; q1 /= n; q2 /= n; q3 /= n; }
    __host__ __device__ Q4 adjoint() const { return Q4(q0, -q1, -q2, -q3); }
    // this * r, vector.cpp:296-305
    __host__ __device__ Q4 multiply(const Q4& r) const {
        return Q4(r.q0 * q0 - r.q1 * q1 - r.q2 * q2 - r.q3 * q3, r.q0 * q1 + r.q1 * q0 - r.q2 * q3 + r.q3 * q2,
                  r.q0 * q2 + r.q1 * q3 + r.q2 * q0 - r.q3 * q1, r.q0 * q3 - r.q1 * q2 + r.q2 * q1 + r.q3 * q0);
    }
};
__host__ __device__ inline Q4 operator*(double s, const Q4& q) { return Q4(q.q0 * s, q.q1 * s, q.q2 * s, q.q3 * s); }

__host__ __device__ inline V3 project(const V3& v, const Q4& q) {  // vector.cpp:481-491
    const Q4 qa = q.adjoint();
    const Q4 vq(0.0, v.x, v.y, v.z);
    Q4 r = q.multiply(vq);
    r = r.multiply(qa);
    return V3(r.q1, r.q2, r.q3);
}
__host__ __device__ inline V3 newtonAcc(const V3& m, const V3& I, const V3& w) {  // vector.cpp:493-498
    return V3((m.x + (I.y - I.z) * w.y * w.z) / I.x, (m.y + (I.z - I.x) * w.z * w.x) / I.y,
              (m.z + (I.x - I.y) * w.x * w.y) / I.z);
}
__host__ __device__ inline Q4 quatAcc(const V3& wa, const Q4& Q1) {  // vector.cpp:500-503
    Q4 w(0.0, wa.x, wa.y, wa.z);
    w.q0 = -2.0 * Q1.norm2();
    return w;
}

// Unit conversions (node.cpp:476-488), all in double.
struct Units {
    double Length, Time, Density, Volume, Speed, Accel, AngVel, Force, Torque, Mass, KinVisc, DynVisc, Stress;
};

// Plane wall (utils.h:18-80) and periodic pair (utils.h:144-163), in device-friendly form.
struct WallDev {
    V3 n, p, rotCenter, omega, vel;
    int moving, slip;
    __host__ __device__ double dist(const V3& x) const { return n.dot(x - p); }
    __host__ __device__ V3 getSpeed(const V3& pt) const {
        if (moving) {
            const V3 dc = pt - rotCenter;
            const V3 da = dc - (dc.dot(n)) * n;
            return vel + omega.cross(da);
        }
        return V3(0.0, 0.0, 0.0);
    }
};

// Static sphere of the object file (DEM.cpp:131-173): position, velocity,
// radius and the object's index (originals, then periodic copies).
struct ObjDev {
    V3 x0, x1;
    double r;
    int index;
};

struct PbcDev {
    V3 v;              // period vector
    V3 p1, n1, p2, n2; // the two planes
};

// Infinite cylinder through p1 and p2 with radius R, fluid inside
// (utils.h:84-130, utils.cpp:56-232); physical units. The functions keep the
// reference's operation order.
struct CylDev {
    V3 p1, p2, naxes, omega;
    double R;
    int moving, slip;
    // R minus the distance of pt from the axis (utils.cpp:56-67)
    __host__ __device__ double dist(const V3& pt) const {
        const V3 p1dist = pt - p1;
        const V3 p1distax = (p1dist.dot(naxes)) * naxes;
        const V3 p1distcylinder = p1distax - p1dist;
        return R - p1distcylinder.norm();
    }
    // pt minus its nearest point on the surface (utils.cpp:69-82)
    __host__ __device__ V3 vecDist(const V3& pt) const {
        const V3 pt1 = pt - p1;
        const V3 projectedPoint = p1 + (naxes.dot(pt1)) * naxes;
        const V3 distFromAxes = pt - projectedPoint;
        return pt - (projectedPoint + distFromAxes / distFromAxes.norm() * R);
    }
    // surface velocity of a rotating cylinder at pt (utils.cpp:106-120)
    __host__ __device__ V3 getSpeed(const V3& pt) const {
        if (moving) {
            const V3 pt1 = pt - p1;
            const V3 projectedPoint = p1 + (naxes.dot(pt1)) * naxes;
            const V3 distFromAxes = pt - projectedPoint;
            return omega.cross(distFromAxes);
        }
        return V3(0.0, 0.0, 0.0);
    }
    // lattice position inside, with axis point and radius in lattice units (vector.cpp:160-170)
    __host__ __device__ static bool inside(const V3& pos, const V3& p1L, const V3& naxes, double RL) {
        const V3 p1dist = pos - p1L;
        const V3 p1distax = (p1dist.dot(naxes)) * naxes;
        const V3 p1distcylinder = p1dist - p1distax;
        return p1distcylinder.norm2() < RL * RL;
    }
    // time at which start + t dir meets the surface, 0 when the quadratic has no
    // admissible root (utils.cpp:177-232; both roots use the minus sign there)
    __host__ __device__ double segmentIntercept(const V3& start, const V3& dir) const {
        const V3 AB = p2 - p1;
        const V3 AO = start - p1;
        const V3 AOxAB = AO.cross(AB);
        const V3 VxAB = dir.cross(AB);
        const double ab2 = AB.dot(AB);
        const double a = VxAB.dot(VxAB);
        const double b = 2 * VxAB.dot(AOxAB);
        const double c = AOxAB.dot(AOxAB) - (R * R * ab2);
        const double delta = b * b - 4 * a * c;
        if (delta < 0) return 0.0;
        const double t = (-b - sqrt(delta)) / (2 * a);
        if (t < 0) return 0.0;
        return t;
    }
};

}  // namespace hyb
