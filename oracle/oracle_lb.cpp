// ORACLE — test infrastructure, not the product. FP64 CPU restatement of the
// reference's lattice-Boltzmann side (LB.cpp, node.cpp, lattice.h).
#include <algorithm>
#include <cstdio>

#include "oracle.h"

namespace orc {

// lattice.h:36-60 (order O, x, y, z, xy, yz, zx), lattice.h:85, lattice.h:98-101
const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0};
const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
const double W[19] = {12.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0,
                      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
const V3 VEC[19] = {V3(0, 0, 0), V3(1, 0, 0), V3(-1, 0, 0), V3(0, 1, 0), V3(0, -1, 0), V3(0, 0, 1), V3(0, 0, -1),
                    V3(1, 1, 0), V3(-1, -1, 0), V3(-1, 1, 0), V3(1, -1, 0), V3(0, 1, 1), V3(0, -1, -1),
                    V3(0, -1, 1), V3(0, 1, -1), V3(1, 0, 1), V3(-1, 0, -1), V3(1, 0, -1), V3(-1, 0, 1)};

// slip-wall reflection tables (lattice.h:87-95): for a diagonal link j, the two
// axis directions whose neighbours decide the reflection, and the populations
// that are reflected from them
const int SLIP1CHECK[19] = {0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 4, 3, 4, 5, 6, 1, 2, 6, 5};
const int SLIP1[19] = {0, 0, 0, 0, 0, 0, 0, 9, 10, 8, 7, 13, 14, 12, 11, 18, 17, 15, 16};
const int SLIP2CHECK[19] = {0, 0, 0, 0, 0, 0, 0, 3, 4, 2, 1, 5, 6, 4, 3, 5, 6, 1, 2};
const int SLIP2[19] = {0, 0, 0, 0, 0, 0, 0, 10, 9, 7, 8, 14, 13, 11, 12, 17, 18, 16, 15};

static const double C1 = 3.0, C2 = 4.5, C3 = 1.5;   // node.cpp:90-93
static const double minTau = 0.501, maxTau = 1.8;    // lattice.h:29-30

void Units::setComposite() {  // node.cpp:476-488
    Volume = Length * Length * Length;
    Speed = Length / Time;
    Accel = Length / Time / Time;
    AngVel = 1.0 / Time;
    KinVisc = Length * Length / Time;
    DynVisc = Density * Length * Length / Time;
    Force = Density * Length * Length * Length * Length / Time / Time;
    Torque = Density * Length * Length * Length * Length * Length / Time / Time;
    Mass = Density * Length * Length * Length;
    Stress = Density * Length * Length / Time / Time;
    Pressure = Density * Length * Length / Time / Time;
}

// ---------------------------------------------------------------- node math
// node.cpp:26-61 initialize + setEquilibrium
void LB::nodeInitialize(uint32_t it, double density, const V3& velocity, double massFunction, double viscosity,
                        const V3& F) {
    Node& nd = nodes[it];
    nd.shearRate = 0.0;
    nd.n = density;
    nd.mass = massFunction;
    nd.u = (1.0 * F) / 2.0 / nd.n + velocity;
    nd.visc = viscosity;
    const double usq = nd.u.norm2();
    for (int j = 0; j < 19; ++j) {
        const double vu = nd.u.dot(VEC[j]);
        nd.f[j] = nd.fs[j] = W[j] * nd.n * (1.0 + C1 * vu + C2 * vu * vu - C3 * usq);
    }
}

static void reconstruct(Node& nd) {  // node.cpp:63-83
    nd.n = 0.0;
    for (int j = 0; j < 19; ++j) nd.n += nd.f[j];
    nd.u.reset();
    for (int j = 0; j < 19; ++j) nd.u += nd.f[j] * VEC[j];
    nd.u /= nd.n;
}

static void shiftVelocity(Node& nd, const V3& F) {  // node.cpp:85-88
    const V3 total = F + nd.hydroForce;
    nd.u += 0.5 * total / nd.n;
}

static void computeEquilibrium(const Node& nd, double feq[19]) {  // node.cpp:90-101
    const double usq = nd.u.norm2();
    for (int j = 0; j < 19; ++j) {
        const double vu = nd.u.dot(VEC[j]);
        feq[j] = W[j] * nd.n * (1.0 + C1 * vu + C2 * vu * vu - C3 * usq);
    }
}

// node.cpp:103-145. The 3x3 tensor is accumulated component by component as
// (f - feq) * (c c^T), scaled by 1.5/(tau n); magnitude as vector.cpp:453-457.
static void computeShearRate(Node& nd, const double feq[19], double turbConst, double plasticVisc, double yieldStress,
                             bool nonNewtonian, bool turbulenceOn) {
    const double minVisc = (minTau - 0.5) / 3 / 1.0;
    const double maxVisc = (maxTau - 0.5) / 3 / 1.0;
    const double tau = 0.5 + 3.0 * nd.visc * 1.0;
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 19; ++j) {
        const double c[3] = {VEC[j].x, VEC[j].y, VEC[j].z};
        const double d = nd.f[j] - feq[j];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) g[a * 3 + b] += (c[a] * c[b]) * d;
    }
    const double s = 1.5 * 1.0 / (tau * nd.n);
    for (int k = 0; k < 9; ++k) g[k] *= s;
    nd.shearRate = std::sqrt(0.5 * (g[0] * g[0] + g[4] * g[4] + g[8] * g[8] + 2.0 * (g[1] * g[3] + g[6] * g[2] + g[5] * g[7])));
    double nuTurb = 0.0;
    if (turbulenceOn) nuTurb = turbConst * nd.shearRate;
    double nuApp = 0.0;
    if (nonNewtonian) nuApp = plasticVisc + yieldStress / (nd.n * 2.0 * nd.shearRate);
    else nuApp = nd.visc;
    nd.visc = std::max(minVisc, std::min(maxVisc, nuApp + nuTurb));
}

static void solveCollision(Node& nd, const double feq[19]) {  // node.cpp:158-165
    const double omega = 1.0 / (0.5 + 3.0 * nd.visc * 1.0);
    for (int j = 0; j < 19; ++j) nd.f[j] += omega * (feq[j] - nd.f[j]);
}

static void addForce(Node& nd, const V3& F) {  // node.cpp:167-184 (Guo forcing)
    const double F1 = 3.0, F2 = 9.0;
    const double omegaf = 1.0 - 1.0 / (1.0 + 6.0 * nd.visc * 1.0);
    const V3 total = F + nd.hydroForce;
    for (int j = 0; j < 19; ++j) {
        const double vu = nd.u.dot(VEC[j]);
        const V3 vmu = VEC[j] - nd.u;
        const V3 fp = F2 * vu * VEC[j] + F1 * vmu;
        nd.f[j] += 1.0 * omegaf * W[j] * fp.dot(total);
    }
}

// Two-relaxation-time collision with Guo forcing (extension; not in the
// reference). omega+ follows the viscosity, omega- = 1/(1/2 + Lambda/(1/omega+ - 1/2)).
// The forcing term of node.cpp:167-184 is split into its even and odd parts,
// relaxed with (1 - omega+/2) and (1 - omega-/2). With Lambda = 1/4 and tau = 1
// both rates are 1 and the update equals solveCollision + addForce.
static void trtCollide(Node& nd, const double feq[19], const V3& F, double magic) {
    const double omP = 1.0 / (0.5 + 3.0 * nd.visc);
    const double omM = 1.0 / (0.5 + magic / (1.0 / omP - 0.5));
    const V3 total = F + nd.hydroForce;
    double fo[19];
    for (int j = 0; j < 19; ++j) fo[j] = nd.f[j];
    for (int j = 0; j < 19; ++j) {
        const int o = OPP[j];
        const double fp = 0.5 * (fo[j] + fo[o]), fm = 0.5 * (fo[j] - fo[o]);
        const double ep = 0.5 * (feq[j] + feq[o]), em = 0.5 * (feq[j] - feq[o]);
        const double vu = nd.u.dot(VEC[j]);
        const double cF = VEC[j].dot(total);
        const double sp = W[j] * (9.0 * vu * cF - 3.0 * nd.u.dot(total));
        const double sm = W[j] * (3.0 * cF);
        nd.f[j] = fo[j] - omP * (fp - ep) - omM * (fm - em) + (1.0 - 0.5 * omP) * sp + (1.0 - 0.5 * omM) * sm;
    }
}

// ---------------------------------------------------------------- parsing
void LB::get(const hyb_config& c) {  // LB.cpp:85-188
    unit.Length = c.unitLength;
    unit.Time = c.unitTime;
    unit.Density = c.unitDensity;
    double lbSizeX = c.lbSize[0], lbSizeY = c.lbSize[1], lbSizeZ = c.lbSize[2];
    lbSize.resize(3);
    lbSize[0] = int(std::floor(lbSizeX / unit.Length + 0.5));
    lbSize[1] = int(std::floor(lbSizeY / unit.Length + 0.5));
    lbSize[2] = int(std::floor(lbSizeZ / unit.Length + 0.5));
    plasticVisc = c.plasticVisc;
    yieldStress = c.yieldStress;
    initDynVisc = c.initVisc;
    freeSurface = c.freeSurface != 0;
    forceField = c.forceField != 0;
    nonNewtonian = c.nonNewtonian != 0;
    turbulenceOn = c.turbulence != 0;
    if (unit.Time == 0.0) {
        const double tauOpt = 1.0;
        if (nonNewtonian) unit.Time = unit.Density * unit.Length * unit.Length / plasticVisc * (1.0 / 3.0) * (tauOpt - 0.5);
        else unit.Time = unit.Density * unit.Length * unit.Length / initDynVisc * (1.0 / 3.0) * (tauOpt - 0.5);
    }
    unit.setComposite();
    initDynVisc /= unit.DynVisc;
    plasticVisc /= unit.DynVisc;
    yieldStress /= unit.Stress;
    initDensity = 1.0;
    initVelocity = V3(c.initVelocity[0], c.initVelocity[1], c.initVelocity[2]);
    initVelocity /= unit.Speed;
    lbF = V3(c.lbF[0], c.lbF[1], c.lbF[2]);
    lbF /= unit.Accel;
    boundary.resize(6);
    for (int k = 0; k < 6; ++k) boundary[k] = (unsigned)c.boundary[k];
    slipCoefficient = c.slipCoefficient;
    turbConst = c.turbConst;
    problem = c.problem;
    for (int k = 0; k < 6; ++k) liquidBox[k] = c.liquidBox[k];
    collision = c.collision;
    trtMagic = c.trtMagic;
}

void LB::latticeDefinition() {  // LB.cpp:42-83
    shift[2] = lbSize[0] * lbSize[1];
    shift[1] = lbSize[0];
    shift[0] = 1;
    domain[2] = lbSize[2] * lbSize[1] * lbSize[0] - 2 * shift[2];
    domain[1] = lbSize[1] * lbSize[0] - 2 * shift[1];
    domain[0] = lbSize[0] - 2 * shift[0];
    for (int j = 0; j < 19; ++j) ne[j] = CX[j] * (int)shift[0] + CY[j] * (int)shift[1] + CZ[j] * (int)shift[2];
}

// ---------------------------------------------------------------- init
void LB::initializeLatticeBoundaries() {  // LB.cpp:370-473
    const uint32_t N = totNodes;
    nb.assign((size_t)N * 19, 0u);  // d[0] is value-initialised to 0 and never set (LB.cpp:377-386)
    for (uint32_t it = 0; it < N; ++it)
        for (int j = 1; j < 19; ++j) nb[(size_t)it * 19 + j] = it + ne[j];
    auto setType = [&](uint32_t it, unsigned code) {
        if (code == 1 || code > 9) { error = HYB_E_ARG; errmsg = "invalid boundary type"; return; }
        t[it] = (uint8_t)code;
    };
    auto self = [&](uint32_t it) { for (int j = 0; j < 19; ++j) nb[(size_t)it * 19 + j] = it; };
    for (uint32_t it = 0; it < N; ++it) {
        if (getX(it) == 0 && !isWall(it)) { setType(it, boundary[0]); self(it); }
        else if (getX(it) == lbSize[0] - 1 && !isWall(it)) { setType(it, boundary[1]); self(it); }
        if (getY(it) == 0 && !isWall(it)) { setType(it, boundary[2]); self(it); }
        else if (getY(it) == lbSize[1] - 1 && !isWall(it)) { setType(it, boundary[3]); self(it); }
        if (getZ(it) == 0 && !isWall(it)) { setType(it, boundary[4]); self(it); }
        else if (getZ(it) == lbSize[2] - 1 && !isWall(it)) { setType(it, boundary[5]); self(it); }
    }
    for (uint32_t it = 0; it < N; ++it) {
        if (t[it] != 0) continue;
        int pbc[19] = {0};
        const uint32_t* d = &nb[(size_t)it * 19];
        for (int j = 1; j < 19; ++j) {
            if (t[d[1]] == 4 && CX[j] > 0) pbc[j] -= domain[0];
            if (t[d[2]] == 4 && CX[j] < 0) pbc[j] += domain[0];
            if (t[d[3]] == 4 && CY[j] > 0) pbc[j] -= domain[1];
            if (t[d[4]] == 4 && CY[j] < 0) pbc[j] += domain[1];
            if (t[d[5]] == 4 && CZ[j] > 0) pbc[j] -= domain[2];
            if (t[d[6]] == 4 && CZ[j] < 0) pbc[j] += domain[2];
        }
        for (int j = 1; j < 19; ++j) nb[(size_t)it * 19 + j] += pbc[j];
    }
}

void LB::initializeParticleBoundaries(const std::vector<Particle>& particles) {  // LB.cpp:475-495
    for (uint32_t it = 0; it < totNodes; ++it) {
        p[it] = 0;
        if (isActive(it)) {
            for (size_t n = 0; n < particles.size(); ++n) {
                if (positions[it].insideSphere(particles[n].x0 / unit.Length, particles[n].r / unit.Length)) {
                    p[it] = 1;
                    solidIndex[it] = particles[n].particleIndex;
                }
            }
        }
    }
}

void LB::initializeWallBoundaries(const std::vector<Wall>& walls) {  // LB.cpp:497-533
    for (size_t iw = 0; iw < walls.size(); ++iw) {
        const Wall& w = walls[iw];
        if (w.flag) continue;
        for (uint32_t it = 0; it < totNodes; ++it) {
            if (positions[it].insidePlane(w.p / unit.Length, w.n)) {
                solidIndex[it] = w.index;
                if (w.slip) t[it] = w.moving ? 6 : 5;
                else t[it] = w.moving ? 8 : 7;
            }
        }
    }
}

void LB::initializeObjectBoundaries(const std::vector<Object>& objects) {  // LB.cpp:535-550
    for (const Object& o : objects)
        for (uint32_t it = 0; it < totNodes; ++it)
            if (positions[it].insideSphere(o.x0 / unit.Length, o.r / unit.Length)) {
                solidIndex[it] = o.index;
                t[it] = 7;
            }
}

void LB::initializeInterface() {  // LB.cpp:605-814 (generic tail + the SPLASH plane + a liquid-box extension)
    if (problem == HYB_PROBLEM_SPLASH) {
        const V3 pp(0.0, 0.0, 20.0), nn(0.0, 0.0, 1.0);
        for (uint32_t it = 0; it < totNodes; ++it)
            if (t[it] == 0 && (positions[it] - pp).dot(nn) > 0.0) t[it] = 2;
    } else if (problem == HYB_PROBLEM_LIQUID_BOX) {
        for (uint32_t it = 0; it < totNodes; ++it) {
            if (t[it] != 0) continue;
            const double x = getX(it), y = getY(it), z = getZ(it);
            if (x < liquidBox[0] || x > liquidBox[1] || y < liquidBox[2] || y > liquidBox[3] || z < liquidBox[4] ||
                z > liquidBox[5])
                t[it] = 2;
        }
    } else if (problem != HYB_PROBLEM_NONE) {
        error = HYB_E_UNSUPPORTED;
        errmsg = "problemName not supported";
    }
    for (uint32_t it = 0; it < totNodes; ++it) {
        if (t[it] == 0) {
            for (int j = 1; j < 19; ++j)
                if (t[nb[(size_t)it * 19 + j]] == 2) t[it] = 3;
        }
    }
    for (uint32_t it = 0; it < totNodes; ++it) {
        if (t[it] == 3) {
            bool closeToFluid = false;
            for (int j = 1; j < 19; ++j)
                if (t[nb[(size_t)it * 19 + j]] == 0) { closeToFluid = true; break; }
            if (!closeToFluid) t[it] = 2;
        }
    }
}

void LB::initializeLists() {  // LB.cpp:817-876
    fluidNodes.clear();
    particleNodes.clear();
    interfaceNodes.clear();
    for (uint32_t it = 0; it < totNodes; ++it) {
        if (t[it] == 0) { alloc[it] = 1; nodes[it] = Node(); fluidNodes.push_back(it); }
        else if (t[it] == 3) { alloc[it] = 1; nodes[it] = Node(); interfaceNodes.push_back(it); }
        if (p[it]) particleNodes.push_back(it);
    }
}

void LB::resetLists() {  // LB.cpp:878-907
    fluidNodes.clear();
    particleNodes.clear();
    interfaceNodes.clear();
    for (uint32_t it = 0; it < totNodes; ++it) {
        if (t[it] == 0) fluidNodes.push_back(it);
        else if (t[it] == 3) interfaceNodes.push_back(it);
        if (p[it]) particleNodes.push_back(it);
    }
}

void LB::initializeVariables() {  // LB.cpp:909-944
    unsigned maxX = 0, maxY = 0, maxZ = 0;
    for (uint32_t it = 0; it < totNodes; ++it) {
        if (isActive(it)) {
            maxX = std::max(maxX, getX(it));
            maxY = std::max(maxY, getY(it));
            maxZ = std::max(maxZ, getZ(it));
        }
    }
    const V3 maxP(maxX, maxY, maxZ);
    for (uint32_t it = 0; it < totNodes; ++it) {
        if (t[it] == 0) {
            const V3 dh = positions[it] - maxP;
            nodeInitialize(it, initDensity + 3.0 * 1.0 * 1.0 * initDensity * dh.dot(lbF), initVelocity, initDensity,
                           initDynVisc, lbF);
        } else if (t[it] == 3) {
            nodeInitialize(it, initDensity, initVelocity, 0.5 * initDensity, initDynVisc, lbF);
        }
    }
}

void LB::initializeWalls(const std::vector<Wall>& walls) {  // LB.cpp:946-996 (plane walls; no cylinders/objects)
    for (uint32_t it = 0; it < totNodes; ++it) {
        if (isWall(it)) continue;
        for (int j = 0; j < 19; ++j) {
            const uint32_t link = nb[(size_t)it * 19 + j];
            if (!alloc[link]) {
                if (t[link] == 7 || t[link] == 5) {
                    alloc[link] = 1;
                    nodes[link] = Node();
                    nodeInitialize(link, initDensity, V3(), 0.0, 0.0, V3());
                } else if (t[link] == 8 || t[link] == 6) {
                    // LB.cpp:969-990: node velocity from its wall (the node lies inside the wall's plane)
                    const Wall& w = walls[solidIndex[link]];
                    if (!positions[link].insidePlane(w.p / unit.Length, w.n)) {
                        error = HYB_E_UNSUPPORTED;
                        errmsg = "moving wall node outside its wall plane (cylinders/objects are not supported)";
                        continue;
                    }
                    const V3 solidVelocity = w.getSpeed(positions[link] * unit.Length) / unit.Speed;
                    alloc[link] = 1;
                    nodes[link] = Node();
                    nodeInitialize(link, initDensity, solidVelocity, 0.0, 0.0, V3());
                } else if (t[link] == 9) {
                    error = HYB_E_UNSUPPORTED;
                    errmsg = "curved walls are not supported";
                }
            }
        }
    }
}

void LB::init(DEM& dem) {  // LB.cpp:190-219
    totNodes = lbSize[0] * lbSize[1] * lbSize[2];
    nodes.assign(totNodes, Node());
    alloc.assign(totNodes, 0);
    t.assign(totNodes, 0);
    p.assign(totNodes, 0);
    solidIndex.assign(totNodes, 0);
    positions.resize(totNodes);
    for (uint32_t it = 0; it < totNodes; ++it) positions[it] = V3(getX(it), getY(it), getZ(it));
    initializeLatticeBoundaries();
    initializeParticleBoundaries(dem.particles);
    initializeWallBoundaries(dem.walls);
    initializeObjectBoundaries(dem.objects);
    initializeInterface();
    initializeLists();
    initializeVariables();
    initializeWalls(dem.walls);
    cleanLists();
    totalMass = 0.0;
    for (uint32_t it = 0; it < totNodes; ++it)
        if (isActive(it) && !p[it]) totalMass += nodes[it].mass;
}

// ---------------------------------------------------------------- step
void LB::cleanLists() {  // LB.cpp:1000-1060
    auto dedupe = [](std::vector<uint32_t>& v) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    };
    dedupe(fluidNodes);
    dedupe(interfaceNodes);
    dedupe(particleNodes);
    activeNodes.clear();
    activeNodes.insert(activeNodes.end(), fluidNodes.begin(), fluidNodes.end());
    activeNodes.insert(activeNodes.end(), interfaceNodes.begin(), interfaceNodes.end());
    std::sort(activeNodes.begin(), activeNodes.end());
    for (int ind = (int)activeNodes.size() - 1; ind > 0; --ind) {
        const uint32_t i = activeNodes[ind], j = activeNodes[ind - 1];
        if (t[i] == 2) activeNodes.erase(activeNodes.begin() + ind);
        else if (i == j) activeNodes.erase(activeNodes.begin() + ind);
    }
    for (uint32_t a : activeNodes)
        if (t[a] == 2) { error = HYB_E_STATE; errmsg = "GAS ERROR"; }
}

void LB::reconstruction() {  // LB.cpp:1062-1070
    for (uint32_t a : activeNodes) reconstruct(nodes[a]);
}

void LB::computeHydroForces(DEM& dem) {  // LB.cpp:1851-1919
    for (auto& e : dem.elmts) { e.FHydro = V3(); e.MHydro = V3(); e.fluidVolume = 0.0; }
    for (uint32_t index : activeNodes) {
        Node& nd = nodes[index];
        nd.hydroForce.reset();
        if (p[index]) {
            const uint32_t pi = solidIndex[index];
            const Particle& pa = dem.particles[pi];
            Element& el = dem.elmts[pa.clusterIndex];
            const V3 radius = positions[index] - pa.x0 / unit.Length + pa.radiusVec / unit.Length;
            const V3 localVel = el.x1 / unit.Speed + (el.wGlobal.cross(radius)) / unit.AngVel;
            const V3 diffVel = (nd.mass / nd.n) * (nd.u - localVel);
            nd.hydroForce += -1.0 * diffVel;
            el.fluidVolume += nd.mass;
            el.FHydro += 1.0 * diffVel;
            el.MHydro += 1.0 * radius.cross(diffVel);
        }
    }
    for (auto& e : dem.elmts) {
        e.FHydro *= unit.Force;
        e.MHydro *= unit.Torque;
        e.fluidVolume *= unit.Volume;
    }
}

void LB::collision_() {  // LB.cpp:1072-1142
    if (!forceField) lbF.reset();
    for (uint32_t index : activeNodes) {
        Node& nd = nodes[index];
        double feq[19];
        shiftVelocity(nd, lbF);
        computeEquilibrium(nd, feq);
        if (nonNewtonian || turbulenceOn)
            computeShearRate(nd, feq, turbConst, plasticVisc, yieldStress, nonNewtonian, turbulenceOn);
        if (collision == HYB_COLLISION_TRT) {
            trtCollide(nd, feq, lbF, trtMagic);
        } else {
            solveCollision(nd, feq);
            addForce(nd, lbF);
        }
    }
}

void LB::streaming(DEM& dem) {  // LB.cpp:1144-1488 (active, gas, plane-wall links; no curved walls)
    const double C2x2 = 9.0, C3x2 = 3.0;
    if (!streamConstsSet) { S1 = slipCoefficient; S2 = 1.0 - slipCoefficient; streamConstsSet = true; }
    double extraMass = 0.0;
    double staticPres[19];
    for (int j = 0; j < 19; ++j) staticPres[j] = initDensity * W[j];
    const double BBCoeff = 2.0 * 3.0 * 1.0 * 1.0;
    for (auto& w : dem.walls) w.FHydro.reset();
    for (uint32_t a : activeNodes)
        for (int j = 0; j < 19; ++j) nodes[a].fs[j] = nodes[a].f[j];
    for (uint32_t it : activeNodes) {
        Node& nd = nodes[it];
        for (int j = 1; j < 19; ++j) {
            const uint32_t link = nb[(size_t)it * 19 + j];
            const uint8_t lt = t[link];
            if (lt == 0 || lt == 3) {
                nd.f[OPP[j]] = nodes[link].fs[OPP[j]];
            } else if (lt == 2) {
                const double usq = nd.u.norm2();
                const double vuj = nd.u.dot(VEC[j]);
                nd.f[OPP[j]] = -nd.fs[j] + W[j] * initDensity * (2.0 + C2x2 * (vuj * vuj) - C3x2 * usq);
            } else if (lt == 8) {  // moving wall (LB.cpp:1321-1341)
                const V3 vel = nodes[link].u;
                const double BBi = BBCoeff * nd.n * W[j] * vel.dot(VEC[j]);
                const V3 BBforce = ((2.0 * (nd.fs[j] - staticPres[j]) - BBi) * 1.0) * VEC[j];  // node.cpp:283-291
                dem.walls[solidIndex[link]].FHydro += BBforce;
                nd.f[OPP[j]] = nd.fs[j] - BBi;
                extraMass += BBi * nd.mass;
            } else if (lt == 7) {
                nd.f[OPP[j]] = nd.fs[j];
            } else if (lt == 5) {  // static slip wall (LB.cpp:1358-1400)
                nd.f[OPP[j]] = nd.fs[j];
                if (j > 6) {
                    const uint32_t c1 = nb[(size_t)it * 19 + SLIP1CHECK[j]], c2 = nb[(size_t)it * 19 + SLIP2CHECK[j]];
                    const bool a1 = isActive(c1), a2 = isActive(c2);
                    if (a1 && !a2) nd.f[OPP[j]] = S1 * nodes[c1].fs[SLIP1[j]] + S2 * nd.fs[j];
                    else if (!a1 && a2) nd.f[OPP[j]] = S1 * nodes[c2].fs[SLIP2[j]] + S2 * nd.fs[j];
                }
            } else if (lt == 6) {  // moving slip wall (LB.cpp:1402-1456)
                const V3 vel = nodes[link].u;
                const double BBi = BBCoeff * nd.n * W[j] * vel.dot(VEC[j]);
                bool done = false;
                if (j > 6) {
                    const uint32_t c1 = nb[(size_t)it * 19 + SLIP1CHECK[j]], c2 = nb[(size_t)it * 19 + SLIP2CHECK[j]];
                    const bool a1 = isActive(c1), a2 = isActive(c2);
                    if (a1 && !a2) {
                        nd.f[OPP[j]] = S1 * nodes[c1].fs[SLIP1[j]] + S2 * (nd.fs[j] - BBi);
                        extraMass += S2 * nd.mass * BBi;
                        done = true;
                    } else if (!a1 && a2) {
                        nd.f[OPP[j]] = S1 * nodes[c2].fs[SLIP2[j]] + S2 * (nd.fs[j] - BBi);
                        extraMass += S2 * nd.mass * BBi;
                        done = true;
                    }
                }
                if (!done) {
                    nd.f[OPP[j]] = nd.fs[j] - BBi;
                    extraMass += nd.mass * BBi;
                }
            } else {
                error = HYB_E_UNSUPPORTED;
                errmsg = "TYPE ERROR: unsupported link type in streaming";
                return;
            }
        }
    }
    redistributeMass(extraMass);
    for (auto& w : dem.walls) w.FHydro *= unit.Force;
}

void LB::redistributeMass(double massSurplus) {  // LB.cpp:1796-1804
    const double addMass = massSurplus / interfaceNodes.size();
    for (uint32_t i : interfaceNodes) nodes[i].mass += addMass;
}

void LB::step(DEM& dem) {  // LB.cpp:221-233
    cleanLists();
    reconstruction();
    computeHydroForces(dem);
    collision_();
    streaming(dem);
}

// ---------------------------------------------------------------- coupling
void LB::findNewActive(DEM& dem) {  // LB.cpp:1921-1967
    std::vector<uint32_t> popUp;
    for (size_t ip = 0; ip < particleNodes.size(); ++ip) {
        const uint32_t index = particleNodes[ip];
        const V3 pos = positions[index];
        const uint32_t cl = dem.particles[solidIndex[index]].clusterIndex;
        bool newActive = true;
        for (int comp : dem.elmts[cl].components) {
            if (pos.insideSphere(dem.particles[comp].x0 / unit.Length, dem.particles[comp].r / unit.Length)) {
                newActive = false;
                break;
            }
        }
        if (newActive) { popUp.push_back(index); p[index] = 0; }
    }
    for (uint32_t nn : popUp)
        for (int ind = (int)particleNodes.size() - 1; ind >= 0; --ind)
            if (nn == particleNodes[ind]) particleNodes.erase(particleNodes.begin() + ind);
}

void LB::findNewSolid(DEM& dem) {  // LB.cpp:1969-2033 (the list grows while it is scanned)
    std::vector<uint32_t> newSolid;
    for (size_t it = 0; it < particleNodes.size(); ++it) {
        const uint32_t index = particleNodes[it];
        const uint32_t cl = dem.particles[solidIndex[index]].clusterIndex;
        for (int k = 1; k < 7; ++k) {
            const uint32_t link = nb[(size_t)index * 19 + k];
            if (p[link]) continue;
            const V3 lp = positions[link];
            bool isNew = false;
            unsigned comp = 0;
            for (int c : dem.elmts[cl].components) {
                comp = (unsigned)c;
                if (lp.insideSphere(dem.particles[comp].x0 / unit.Length, dem.particles[comp].r / unit.Length)) {
                    isNew = true;
                    break;
                }
            }
            if (isNew) {
                bool already = false;
                for (uint32_t q : newSolid) if (q == link) { already = true; break; }
                if (!already) {
                    solidIndex[link] = comp;
                    p[link] = 1;
                    newSolid.push_back(link);
                    particleNodes.push_back(link);
                }
            }
        }
    }
}

void LB::couplingStep(DEM& dem) {  // LB.cpp:247-280
    if (dem.newNeighborList) {
        initializeParticleBoundaries(dem.particles);  // updateIndices, LB.cpp:1826-1849
        resetLists();
        dem.newNeighborList = false;
    }
    findNewActive(dem);
    findNewSolid(dem);
}

// ---------------------------------------------------------------- free surface
void LB::updateMass() {  // LB.cpp:1492-1590 (active, gas, plane-wall links; no curved walls)
    for (uint32_t i : interfaceNodes) nodes[i].newMass = nodes[i].mass;
    for (uint32_t index : interfaceNodes) {
        double deltaMass = 0.0;
        Node& nd = nodes[index];
        for (int j = 1; j < 19; ++j) {
            const uint32_t link = nb[(size_t)index * 19 + j];
            double averageMass = 0.0;
            if (t[link] == 3) averageMass = 0.5 * (nodes[link].mass + nd.mass);
            else if (t[link] == 0) averageMass = 1.0;
            else if (t[link] == 2) averageMass = 0.0;
            else if (t[link] == 8) averageMass = 1.0 * nd.mass;
            else if (t[link] == 6) {  // LB.cpp:1545-1575
                if (j > 6) {
                    const uint32_t c1 = nb[(size_t)index * 19 + SLIP1CHECK[j]], c2 = nb[(size_t)index * 19 + SLIP2CHECK[j]];
                    const bool a1 = isActive(c1), a2 = isActive(c2);
                    if (a1 && !a2) averageMass += 1.0 * (1.0 - slipCoefficient) * nd.mass;
                    else if (!a1 && a2) averageMass += 1.0 * (1.0 - slipCoefficient) * nd.mass;
                    else averageMass += 1.0 * nd.mass;
                } else {
                    averageMass += 1.0 * nd.mass;
                }
            }
            deltaMass += 1.0 * averageMass * (nd.f[OPP[j]] - nd.fs[j]);
        }
        nd.newMass += deltaMass;
    }
    for (uint32_t i : fluidNodes) nodes[i].mass = nodes[i].n;
    for (uint32_t i : interfaceNodes) nodes[i].mass = nodes[i].newMass;
}

void LB::updateInterface() {  // LB.cpp:1592-1794
    double massSurplus = 0.0;
    std::vector<uint32_t> filled, emptied, newInterface;
    // findInterfaceMutants (LB.cpp:1620-1650)
    for (int ind = (int)interfaceNodes.size() - 1; ind >= 0; --ind) {
        const uint32_t i = interfaceNodes[ind];
        if (nodes[i].mass > nodes[i].n) {
            filled.push_back(i);
            t[i] = 0;
            fluidNodes.push_back(i);
            interfaceNodes.erase(interfaceNodes.begin() + ind);
        } else if (nodes[i].mass < 0.0) {
            emptied.push_back(i);
            t[i] = 2;
            interfaceNodes.erase(interfaceNodes.begin() + ind);
        }
    }
    // smoothenInterface (LB.cpp:1652-1701)
    for (uint32_t index : filled) {
        for (int j = 1; j < 19; ++j) {
            const uint32_t link = nb[(size_t)index * 19 + j];
            if (t[link] == 2) {
                t[link] = 3;
                newInterface.push_back(link);
                alloc[link] = 1;
                nodes[link] = Node();
                nodeInitialize(link, initDensity, nodes[index].u, 0.01 * initDensity, nodes[index].visc,
                               nodes[index].hydroForce + lbF);
                massSurplus -= 0.01 * initDensity;
            }
        }
    }
    for (uint32_t index : emptied) {
        for (int j = 0; j < 19; ++j) {
            const uint32_t link = nb[(size_t)index * 19 + j];
            if (t[link] == 0) {
                t[link] = 3;
                newInterface.push_back(link);
                nodes[link].mass = 0.99 * nodes[link].n;
                massSurplus += 0.01 * nodes[link].n;
                for (int ind = (int)fluidNodes.size() - 1; ind >= 0; --ind)
                    if (fluidNodes[ind] == link) fluidNodes.erase(fluidNodes.begin() + ind);
            }
        }
    }
    // updateMutants (LB.cpp:1703-1742)
    for (uint32_t e : emptied) {
        bool notInterface = true;
        for (uint32_t q : newInterface) if (e == q) notInterface = false;
        if (notInterface) {
            massSurplus += nodes[e].mass;
            alloc[e] = 0;
            nodes[e] = Node();
        }
    }
    for (uint32_t f : filled) {
        bool notInterface = true;
        for (uint32_t q : newInterface) if (f == q) notInterface = false;
        if (notInterface) {
            massSurplus += nodes[f].mass - nodes[f].n;
            nodes[f].mass = nodes[f].n;
        }
    }
    for (uint32_t q : newInterface) interfaceNodes.push_back(q);
    // removeIsolated (LB.cpp:1744-1794)
    for (int ind = (int)interfaceNodes.size() - 1; ind >= 0; --ind) {
        const uint32_t i = interfaceNodes[ind];
        bool surroundedFluid = true;
        for (int j = 1; j < 19; ++j)
            if (t[nb[(size_t)i * 19 + j]] == 2) { surroundedFluid = false; break; }
        if (surroundedFluid) {
            massSurplus += nodes[i].mass - nodes[i].n;
            nodes[i].mass = nodes[i].n;
            t[i] = 0;
            interfaceNodes.erase(interfaceNodes.begin() + ind);
            fluidNodes.push_back(i);
        }
    }
    for (int ind = (int)interfaceNodes.size() - 1; ind >= 0; --ind) {
        const uint32_t i = interfaceNodes[ind];
        bool surroundedGas = true;
        for (int j = 1; j < 19; ++j)
            if (t[nb[(size_t)i * 19 + j]] == 0) { surroundedGas = false; break; }
        if (surroundedGas) {
            massSurplus += nodes[i].mass;
            interfaceNodes.erase(interfaceNodes.begin() + ind);
            t[i] = 2;
            alloc[i] = 0;
            nodes[i] = Node();
        }
    }
    redistributeMass(massSurplus);
}

void LB::freeSurfaceStep() {  // LB.cpp:235-245
    updateMass();
    updateInterface();
}

}  // namespace orc
