// ORACLE — test infrastructure, not the product.
//
// FP64, single-thread CPU restatement of the reference's coupled LBM-DEM cycle
// (goCycle, hybird.cpp:35-66), written from the reference's behaviour with the
// reference's operation order so that it can be compared bit-for-bit with the
// compiled reference (oracle/_ref) and used as the checker of the HIP path.
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
//
// Every function cites the reference file:line it restates. Pinning (what each
// part is checked against) is listed in oracle_api.cpp.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "hybird_amd.h"

namespace orc {

// ---- value types with the reference's operation semantics (vector.cpp:38-180) ----
struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    void reset() { x = y = z = 0.0; }
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(double s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(double s) const { return V3(x / s, y / s, z / s); }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3& operator-=(const V3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    V3& operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
    V3& operator/=(double s) { x /= s; y /= s; z /= s; return *this; }
    double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 cross(const V3& o) const { return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
    double norm2() const { return x * x + y * y + z * z; }
    double norm() const { return std::sqrt(x * x + y * y + z * z); }
    V3 transport() const { return V3(y * y + z * z, z * z + x * x, x * x + y * y); }
    bool insideSphere(const V3& c, double r) const { V3 d = *this - c; return d.norm2() < r * r; }
    bool insidePlane(const V3& p, const V3& n) const { V3 d = *this - p; return n.dot(d) < 0.0; }
};
// scalar * vector: vector.cpp:108-113 multiplies each component by the scalar
inline V3 operator*(double s, const V3& v) { return V3(v.x * s, v.y * s, v.z * s); }

struct Q4 {
    double q0 = 1.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    Q4() = default;
    Q4(double a, double b, double c, double d) : q0(a), q1(b), q2(c), q3(d) {}
    Q4 operator+(const Q4& o) const { return Q4(q0 + o.q0, q1 + o.q1, q2 + o.q2, q3 + o.q3); }
    Q4 operator-(const Q4& o) const { return Q4(q0 - o.q0, q1 - o.q1, q2 - o.q2, q3 - o.q3); }
    Q4 operator*(double s) const { return Q4(q0 * s, q1 * s, q2 * s, q3 * s); }
    double norm() const { return std::sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3); }
    double norm2() const { return q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3; }
    void normalize() { double n = norm(); q0 /= n; q1 /= n; q2 /= n; q3 /= n; }
    Q4 adjoint() const { return Q4(q0, -q1, -q2, -q3); }
    // vector.cpp:296-305: product this * r
    Q4 multiply(const Q4& r) const {
        return Q4(r.q0 * q0 - r.q1 * q1 - r.q2 * q2 - r.q3 * q3,
                  r.q0 * q1 + r.q1 * q0 - r.q2 * q3 + r.q3 * q2,
                  r.q0 * q2 + r.q1 * q3 + r.q2 * q0 - r.q3 * q1,
                  r.q0 * q3 - r.q1 * q2 + r.q2 * q1 + r.q3 * q0);
    }
};
inline Q4 operator*(double s, const Q4& q) { return Q4(q.q0 * s, q.q1 * s, q.q2 * s, q.q3 * s); }

// vector.cpp:481-491 v' = q v q*
inline V3 project(const V3& v, const Q4& q) {
    const Q4 qa = q.adjoint();
    const Q4 vq(0.0, v.x, v.y, v.z);
    Q4 r = q.multiply(vq);
    r = r.multiply(qa);
    return V3(r.q1, r.q2, r.q3);
}
// vector.cpp:493-498 Euler equations in the body frame
inline V3 newtonAcc(const V3& m, const V3& I, const V3& w) {
    const double xA = (m.x + (I.y - I.z) * w.y * w.z) / I.x;
    const double yA = (m.y + (I.z - I.x) * w.z * w.x) / I.y;
    const double zA = (m.z + (I.x - I.y) * w.x * w.y) / I.z;
    return V3(xA, yA, zA);
}
// vector.cpp:500-503
inline Q4 quatAcc(const V3& wa, const Q4& Q1) { Q4 w(0.0, wa.x, wa.y, wa.z); w.q0 = -2.0 * Q1.norm2(); return w; }

// ---- D3Q19 (lattice.h:36-101) ----
extern const int CX[19], CY[19], CZ[19];
extern const int OPP[19];
extern const double W[19];
extern const V3 VEC[19];

// ---- units (node.cpp:476-488) ----
struct Units {
    double Length = 1, Time = 1, Density = 1;
    double Volume = 1, Speed = 1, Accel = 1, AngVel = 1, Force = 1, Torque = 1, Mass = 1, KinVisc = 1, DynVisc = 1,
           Stress = 1, Pressure = 1;
    void setComposite();
};

// ---- entities (elmt.h, utils.h) ----
struct Particle {
    uint32_t particleIndex = 0, clusterIndex = 0, protoIndex = 0, tableCell = 0;
    double r = 0.0, D = 0.0;
    V3 x0, x1, radiusVec;
};

struct Element {
    uint32_t index = 0;
    std::vector<int> components;
    uint32_t size = 1;
    double radius = 0.0, m = 1.0;
    V3 I = V3(1, 1, 1);
    V3 x0, x1, x2, x3, x4, x5, xp0, xp1, xp2, xp3, xp4, xp5;
    Q4 q0, q1, q2, q3, q4, q5, qp0, qp1, qp2, qp3, qp4, qp5;
    V3 wGlobal, wLocal, wpGlobal, wpLocal;
    V3 w0, w1, w2, w3, w4, w5, wp0, wp1, wp2, wp3, wp4, wp5;
    V3 FHydro, FParticle, FWall, FGrav, MHydro, MParticle, MWall;
    double fluidVolume = 0.0;
    Element() {
        q0 = Q4(1, 0, 0, 0);
        q1 = q2 = q3 = q4 = q5 = Q4(0, 0, 0, 0);
        qp0 = Q4(1, 0, 0, 0);
        qp1 = qp2 = qp3 = qp4 = qp5 = Q4(0, 0, 0, 0);
        components.resize(1);
    }
};

struct Wall {
    uint32_t index = 0;
    V3 n, p, rotCenter, omega, vel, FHydro, FParticle, trans;
    bool moving = false, slip = false, flag = false, translating = false;
    double dist(const V3& x) const { return n.dot(x - p); }  // utils.cpp:14-17
    V3 getSpeed(const V3& pt) const;                          // utils.cpp:37-47
};

struct Object {  // static sphere of the object file (DEM.cpp:131-173; elmt.h object)
    uint32_t index = 0, originalIndex = 0;
    double r = 0.0;
    V3 x0, x1, FParticle;
};

struct Pbc {
    V3 p, v;
    Wall pl1, pl2;
    void setPlanes();  // utils.cpp:240-245
};

struct Ghost {
    uint32_t ghostIndex = 0;
    V3 pbcVector;
};

struct Material {
    double density = 1, youngMod = 1, poisson = 0.3, knConst = 1, ksConst = 1, linearStiff = 0, restitution = 0.5,
           dampCoeff = 0.5, viscTang = 0.5, frictionCoefPart = 0.3, frictionCoefWall = 0.3;
    int contactModel = HYB_CONTACT_HERTZIAN;
};

struct Node {
    double f[19], fs[19];
    double n = 0.0;
    V3 u, hydroForce;
    double mass = 0.0, newMass = 0.0, visc = 1.0, shearRate = 0.0;
};

struct Sim;

struct DEM {
    std::vector<double> demSize = std::vector<double>(3, 1.0);
    double maxDisp = 0.0, nebrRange = 0.0;
    double cellWidth[3] = {0, 0, 0};
    unsigned nCells[3] = {0, 0, 0};
    std::vector<int> cellTable;
    std::vector<uint32_t> neighborTable, nearWallTable;
    double numVisc = 0.0;
    std::vector<std::vector<V3>> prototypes;
    double demInitialRepeat = 0.0, deltat = 1.0, demTime = 0.0;
    uint32_t demTimeStep = 0;
    V3 demF;
    Material mat;
    uint32_t multiStep = 1;
    double criticalRatio = 0.1;
    std::vector<Wall> walls;
    std::vector<Pbc> pbcs;
    std::vector<Object> objects;  // originals, then periodic copies
    uint32_t stdObjects = 0;
    std::vector<uint32_t> nearObjectTable;
    std::vector<Element> elmts;
    std::vector<Particle> particles;
    uint32_t stdParticles = 0;
    std::vector<Ghost> ghosts;
    bool newNeighborList = false;
    bool allWalls = false;  // HYB_WALL_CONTACTS_ALL (not in the reference): no `break` after the first near wall
    // particleEnergy (utils.h:165-190): kin, rotKin, trKin, grav, elastic, total
    double energy[6] = {0, 0, 0, 0, 0, 0};
    // values the reference freezes in function-local statics at the first call
    double neighListTrigger = 0.0;
    double c1[5], c2[5], coeff1ord[6], coeff2ord[6];
    bool first_step_done = false;
    int error = 0;
    std::string errmsg;

    void get(const hyb_config& cfg, int n, const double* rec);
    void init(const std::vector<unsigned>& boundary, const std::vector<unsigned>& lbSize, const Units& unit, const V3& lbF);
    void step();
    // internals
    void compositeProperties();
    void initializeWalls(const std::vector<unsigned>& boundary, const std::vector<unsigned>& lbSize, const Units& unit);
    void initializePbcs(const std::vector<unsigned>& boundary, const std::vector<unsigned>& lbSize, const Units& unit);
    void initNeighborParameters();
    void evalMaxDisp();
    void evalCellTable();
    void evalNeighborTable();
    void evalNearWallTable();
    void evalNearObjectTable();
    void periodicObjects();
    void objectParticleCollision(Object& o, const Particle& pJ, const V3& vectorDistance);
    void pbcShift();
    void createGhosts();
    void predictor();
    void corrector();
    void evaluateForces();
    void updateParticlesPredicted();
    void updateParticlesCorrected();
    void updateEnergy();
    void particleParticleCollision(const Particle& pi, const Particle& pj, const V3& vectorDistance);
    void wallParticleCollision(Wall& w, const Particle& pj, double overlap);
    double normalContact(double overlap, double vrel, double effRad, double effMass) const;
    double tangentialContact(double vrelt, double fn, double effRad, double effMass, double friction) const;
};

struct LB {
    // parameters (LB.h:36-101)
    double initDynVisc = (1.0 - 0.5) / 3.0, yieldStress = 1.0, plasticVisc = 1.0, turbConst = 0.0, slipCoefficient = 0.0;
    double initDensity = 0.0;
    V3 initVelocity, lbF;
    uint32_t totNodes = 1;
    double totalMass = 0.0;
    int ne[19];
    unsigned shift[3];
    int domain[3];
    std::vector<uint8_t> t;        // node type code
    std::vector<uint8_t> p;        // inside-particle flag
    std::vector<uint32_t> solidIndex;
    std::vector<V3> positions;
    std::vector<uint32_t> nb;      // [N][19] neighbour table
    std::vector<unsigned> boundary = std::vector<unsigned>(6, 1);
    std::vector<Node> nodes;
    std::vector<uint8_t> alloc;    // node allocated (nodes[it] != nullptr in the reference)
    std::vector<uint32_t> activeNodes, fluidNodes, particleNodes, interfaceNodes;
    bool freeSurface = false, forceField = false, nonNewtonian = false, turbulenceOn = false;
    uint32_t time = 0;
    std::vector<unsigned> lbSize = std::vector<unsigned>(3, 1);
    Units unit;
    int problem = HYB_PROBLEM_NONE;
    double liquidBox[6] = {0, 0, 0, 0, 0, 0};
    // extensions
    int collision = HYB_COLLISION_BGK;
    double trtMagic = 0.25;
    // streaming constants frozen at the first call (LB.cpp:1152-1155)
    bool streamConstsSet = false;
    double S1 = 0.0, S2 = 1.0;
    int error = 0;
    std::string errmsg;

    void get(const hyb_config& cfg);
    void latticeDefinition();
    void init(DEM& dem);
    void step(DEM& dem);
    void couplingStep(DEM& dem);
    void freeSurfaceStep();

    unsigned getX(uint32_t i) const { return i % lbSize[0]; }
    unsigned getY(uint32_t i) const { return (i % (lbSize[0] * lbSize[1])) / lbSize[0]; }
    unsigned getZ(uint32_t i) const { return i / (lbSize[0] * lbSize[1]); }
    bool isActive(uint32_t i) const { return t[i] == 0 || t[i] == 3; }
    bool isWall(uint32_t i) const { return t[i] == 5 || t[i] == 6 || t[i] == 7 || t[i] == 8 || t[i] == 9; }

    // internals
    void initializeLatticeBoundaries();
    void initializeParticleBoundaries(const std::vector<Particle>& particles);
    void initializeWallBoundaries(const std::vector<Wall>& walls);
    void initializeObjectBoundaries(const std::vector<Object>& objects);
    void initializeInterface();
    void initializeLists();
    void resetLists();
    void initializeVariables();
    void initializeWalls(const std::vector<Wall>& walls);
    void cleanLists();
    void reconstruction();
    void computeHydroForces(DEM& dem);
    void collision_();
    void streaming(DEM& dem);
    void updateMass();
    void updateInterface();
    void redistributeMass(double massSurplus);
    void findNewActive(DEM& dem);
    void findNewSolid(DEM& dem);
    void nodeInitialize(uint32_t it, double density, const V3& velocity, double massFunction, double viscosity, const V3& F);
};

struct Sim {
    hyb_config cfg;
    LB lb;
    DEM dem;
    int demSolve = 0, lbmSolve = 0;
    uint32_t currentTimeStep = 0;
    void cycle();
};

}  // namespace orc
