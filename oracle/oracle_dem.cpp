// ORACLE — test infrastructure, not the product. FP64 CPU restatement of the
// reference's discrete-element side (DEM.cpp, elmt.cpp, utils.cpp).
#include <algorithm>
#include <cstdio>

#include "oracle.h"

namespace orc {

V3 Wall::getSpeed(const V3& pt) const {  // utils.cpp:37-47
    if (moving) {
        const V3 dc = pt - rotCenter;
        const V3 da = dc - (dc.dot(n)) * n;
        return vel + omega.cross(da);
    }
    return V3(0.0, 0.0, 0.0);
}

void Pbc::setPlanes() {  // utils.cpp:240-245
    pl1.p = p;
    pl2.p = p + v;
    pl1.n = v / v.norm();
    pl2.n = -1.0 * pl1.n;
}

// ---------------------------------------------------------------- parsing / init
void DEM::get(const hyb_config& c, int n, const double* rec) {  // DEM.cpp:13-183
    allWalls = c.wallContacts == HYB_WALL_CONTACTS_ALL;  // extension: every wall in range, in wall order
    mat.density = c.density;
    mat.contactModel = c.contactModel;
    mat.youngMod = c.youngMod;
    mat.poisson = c.poisson;
    mat.knConst = 2.0 / 3.0 * mat.youngMod / (1 - mat.poisson * mat.poisson);
    mat.ksConst = 2.0 * mat.youngMod / (2.0 - mat.poisson) / (1.0 + mat.poisson);
    mat.linearStiff = c.linearStiff;
    mat.restitution = c.restitution;
    // DEM.cpp:43-53: the HERTZIAN case falls through into LINEAR, whose formula
    // (with +pi where pi^2 was meant) is the one that survives
    mat.dampCoeff = -1.0 * std::sqrt(2.0) * std::log(mat.restitution) /
                    std::sqrt((std::log(mat.restitution) * std::log(mat.restitution) + M_PI));
    mat.viscTang = c.viscTang;
    mat.frictionCoefPart = c.frictionCoefPart;
    mat.frictionCoefWall = c.frictionCoefWall;
    const V3 translate(c.translate[0], c.translate[1], c.translate[2]);
    const double scale = c.scale;
    elmts.clear();
    for (int e = 0; e < n; ++e) {
        const double* r = rec + (size_t)e * HYB_ELEMENT_RECORD;
        Element el;
        el.index = (uint32_t)r[0];
        el.size = (uint32_t)r[1];
        el.radius = r[2] * scale;
        el.x0 = V3(r[3] * scale, r[4] * scale, r[5] * scale) + translate;
        el.x1 = V3(r[6], r[7], r[8]);
        el.w0 = V3(r[9], r[10], r[11]);
        el.q0 = Q4(r[12], r[13], r[14], r[15]);
        el.q1 = Q4(r[16], r[17], r[18], r[19]);
        elmts.push_back(el);
    }
    numVisc = c.numVisc;
    multiStep = (uint32_t)c.multiStep;
    criticalRatio = c.criticalRatio;
    demInitialRepeat = c.demInitialRepeat;
}

void DEM::compositeProperties() {  // DEM.cpp:404-433
    prototypes.assign(5, {});
    prototypes[1] = {V3()};
    prototypes[2] = {V3(0.5, 0.0, 0.0), V3(-0.5, 0.0, 0.0)};
    prototypes[3] = {V3(0.0, 1.0, 0.0), V3(-std::sqrt(3) / 2, -1 / 2, 0.0), V3(std::sqrt(3) / 2, -1 / 2, 0.0)};
    prototypes[4] = {V3(0.0, 0.0, 1.0), V3(0.0, 2.0 * std::sqrt(2) / 3.0, -1.0 / 3.0),
                     V3(2.0 * std::sqrt(6) / 6.0, -2.0 * std::sqrt(2) / 6.0, -1.0 / 3.0),
                     V3(-2.0 * std::sqrt(6) / 6.0, -2.0 * std::sqrt(2) / 6.0, -1.0 / 3.0)};
}

static void elmtInitialize(Element& e, double partDensity, const std::vector<std::vector<V3>>& proto,
                           const V3& demF) {  // elmt.cpp:17-83 (wSolver = true)
    e.xp0 = e.x0;
    e.xp1 = e.x1;
    e.xp2.reset(); e.x2.reset(); e.xp3.reset(); e.x3.reset(); e.xp4.reset(); e.x4.reset(); e.xp5.reset(); e.x5.reset();
    e.qp0 = e.q0;
    e.qp1 = e.q1;
    e.qp2 = e.q2 = Q4(0, 0, 0, 0);
    e.qp3 = e.q3 = Q4(0, 0, 0, 0);
    e.qp4 = e.q4 = Q4(0, 0, 0, 0);
    e.qp5 = e.q5 = Q4(0, 0, 0, 0);
    e.wpGlobal = e.wGlobal = e.wpLocal = e.wLocal = e.w0;
    e.wp1.reset(); e.w1.reset(); e.wp2.reset(); e.w2.reset(); e.wp3.reset(); e.w3.reset();
    e.wp4.reset(); e.w4.reset(); e.wp5.reset(); e.w5.reset();
    const double singleMass = 4.0 / 3.0 * partDensity * M_PI * e.radius * e.radius * e.radius;
    e.m = e.size * singleMass;
    e.I = e.size * 2.0 / 5.0 * singleMass * e.radius * e.radius * V3(1.0, 1.0, 1.0);
    for (uint32_t n = 0; n < e.size; ++n) e.I += singleMass * e.radius * e.radius * proto[e.size][n].transport();
    e.FHydro.reset(); e.FParticle.reset(); e.FWall.reset();
    e.FGrav = demF * e.m;
    e.MHydro.reset(); e.MParticle.reset(); e.MWall.reset();
}

static void particleUpdateCorrected(Particle& p, const Element& e, const std::vector<std::vector<V3>>& proto) {
    // elmt.cpp:277-290
    p.x0 = e.x0;
    p.radiusVec = V3();
    p.x1 = e.x1;
    if (e.size > 1) {
        p.x0 = p.x0 + p.r * project(proto[e.size][p.protoIndex], e.q0);
        p.radiusVec = p.x0 - e.x0;
        p.x1 = p.x1 + e.wGlobal.cross(p.radiusVec);
    }
}

static void particleUpdatePredicted(Particle& p, const Element& e, const std::vector<std::vector<V3>>& proto) {
    // elmt.cpp:261-275
    p.x0 = e.xp0;
    p.radiusVec = V3();
    p.x1 = e.xp1;
    if (e.size > 1) {
        p.x0 = p.x0 + p.r * project(proto[e.size][p.protoIndex], e.qp0);
        p.radiusVec = p.x0 - e.xp0;
        p.x1 = p.x1 + e.wpGlobal.cross(p.radiusVec);
    }
}

static void ghostUpdate(Particle& g, const Particle& o, const V3& v) {  // elmt.cpp:292-298
    g.x0 = o.x0 + v;
    g.radiusVec = o.radiusVec;
    g.x1 = o.x1;
}

void DEM::initializeWalls(const std::vector<unsigned>& boundary, const std::vector<unsigned>& lbSize, const Units& unit) {
    // DEM.cpp:435-638 (the six box walls; problem-specific extra walls are not built)
    walls.clear();
    uint32_t index = 0;
    const V3 ps[6] = {V3(0.5 * unit.Length, 0, 0), V3((double(lbSize[0]) - 1.5) * unit.Length, 0, 0),
                      V3(0, 0.5 * unit.Length, 0), V3(0, (double(lbSize[1]) - 1.5) * unit.Length, 0),
                      V3(0, 0, 0.5 * unit.Length), V3(0, 0, (double(lbSize[2]) - 1.5) * unit.Length)};
    const V3 ns[6] = {V3(1, 0, 0), V3(-1, 0, 0), V3(0, 1, 0), V3(0, -1, 0), V3(0, 0, 1), V3(0, 0, -1)};
    for (int k = 0; k < 6; ++k) {
        const unsigned b = boundary[k];
        if (b == 5 || b == 6 || b == 7 || b == 8) {
            Wall w;
            w.p = ps[k];
            w.n = ns[k];
            w.index = index++;
            w.flag = false;
            w.translating = (k == 5);   // DEM.cpp:612-613 (only used by evolveBoundaries)
            w.moving = (b == 6 || b == 8);
            w.slip = (b == 5 || b == 6);
            walls.push_back(w);
        }
    }
}

void DEM::initializePbcs(const std::vector<unsigned>& boundary, const std::vector<unsigned>& lbSize, const Units& unit) {
    // DEM.cpp:937-988
    pbcs.clear();
    const V3 ps[3] = {V3(0.5 * unit.Length, 0, 0), V3(0, 0.5 * unit.Length, 0), V3(0, 0, 0.5 * unit.Length)};
    for (int k = 0; k < 3; ++k) {
        if (boundary[2 * k] == 4) {
            if (boundary[2 * k + 1] != 4) { error = HYB_E_ARG; errmsg = "Periodic boundaries error!"; return; }
            Pbc pb;
            pb.p = ps[k];
            V3 v;
            if (k == 0) v = V3((double(lbSize[0]) - 2.0) * unit.Length, 0, 0);
            if (k == 1) v = V3(0, (double(lbSize[1]) - 2.0) * unit.Length, 0);
            if (k == 2) v = V3(0, 0, (double(lbSize[2]) - 2.0) * unit.Length);
            pb.v = v;
            pb.setPlanes();
            pbcs.push_back(pb);
        }
    }
}

void DEM::init(const std::vector<unsigned>& boundary, const std::vector<unsigned>& lbSize, const Units& unit,
               const V3& lbF) {  // DEM.cpp:185-297
    demSize[0] = double(lbSize[0]) * unit.Length;
    demSize[1] = double(lbSize[1]) * unit.Length;
    demSize[2] = double(lbSize[2]) * unit.Length;
    demF = lbF * unit.Accel;
    if (elmts.size()) {
        if (multiStep == 0) { error = HYB_E_UNSUPPORTED; errmsg = "multiStep = 0 (automatic) is not supported"; return; }
        deltat = unit.Time / (double(multiStep));
    } else {
        deltat = unit.Time;
        multiStep = 1;
    }
    compositeProperties();
    particles.clear();
    uint32_t globalIndex = 0;
    for (auto& e : elmts) {
        elmtInitialize(e, mat.density, prototypes, demF);
        e.components.resize(e.size);
        for (uint32_t i = 0; i < e.size; ++i) {  // elmt::generateParticles, elmt.cpp:122-137
            Particle p;
            p.particleIndex = globalIndex;
            e.components[i] = (int)globalIndex;
            p.clusterIndex = e.index;
            p.r = e.radius;
            p.D = p.r * 2.0;
            p.protoIndex = i;
            particleUpdateCorrected(p, e, prototypes);
            particles.push_back(p);
            ++globalIndex;
        }
    }
    stdParticles = (uint32_t)particles.size();
    initializeWalls(boundary, lbSize, unit);
    initializePbcs(boundary, lbSize, unit);
    stdObjects = (uint32_t)objects.size();
    if (elmts.size()) {
        initNeighborParameters();
        periodicObjects();
        evalNeighborTable();
    }
    // determineTimeStep(unit.Time) repeats the same choice (DEM.cpp:299-329)
    if (elmts.size()) deltat = unit.Time / (double(multiStep));
    else { deltat = unit.Time; multiStep = 1; }
    // coefficients the reference captures in statics on the first step (DEM.cpp:1068-1069, 1089-1096)
    const double dt = deltat;
    const double c[5] = {dt, dt * dt / 2.0, dt * dt * dt / 6.0, dt * dt * dt * dt / 24.0, dt * dt * dt * dt * dt / 120.0};
    for (int k = 0; k < 5; ++k) { c1[k] = c[k]; c2[k] = c[k]; }
    const double g1[6] = {95.0 / 288.0, 1.0, 25.0 / 24.0, 35.0 / 72.0, 5.0 / 48.0, 1.0 / 120.0};
    const double g2[6] = {3.0 / 16.0, 251.0 / 360.0, 1.0, 11.0 / 18.0, 1.0 / 6.0, 1.0 / 60.0};
    const double co1[6] = {g1[0] * c[0], g1[1] * c[0] / c[0], g1[2] * c[0] / c[1], g1[3] * c[0] / c[2], g1[4] * c[0] / c[3],
                           g1[5] * c[0] / c[4]};
    const double co2[6] = {g2[0] * c[1], g2[1] * c[1] / c[0], g2[2] * c[1] / c[1], g2[3] * c[1] / c[2], g2[4] * c[1] / c[3],
                           g2[5] * c[1] / c[4]};
    for (int k = 0; k < 6; ++k) { coeff1ord[k] = co1[k]; coeff2ord[k] = co2[k]; }
}

// ---------------------------------------------------------------- neighbour search
void DEM::initNeighborParameters() {  // DEM.cpp:1270-1312
    double maxRad = 0.0;
    for (auto& e : elmts) if (maxRad < e.radius) maxRad = e.radius;
    if (elmts.size() == 0) {
        for (int k = 0; k < 3; ++k) cellWidth[k] = demSize[k];
    } else {
        for (int k = 0; k < 3; ++k) cellWidth[k] = std::min(maxRad * 5.0, demSize[k]);
    }
    for (int k = 0; k < 3; ++k) {
        nCells[k] = int(std::ceil(demSize[k] / cellWidth[k]));
        cellWidth[k] = demSize[k] / double(nCells[k]);
        nCells[k] += 2;
    }
    nebrRange = std::max(maxRad * 3.0, 0.5 * std::min(cellWidth[0], std::min(cellWidth[1], cellWidth[2])));
    maxDisp = 0.5 * nebrRange;
}

void DEM::evalMaxDisp() {  // DEM.cpp:1314-1324
    double maxVel = 0.0;
    for (auto& e : elmts) if (e.x1.norm2() > maxVel) maxVel = e.x1.norm2();
    maxVel = std::sqrt(maxVel);
    maxDisp += maxVel * deltat;
}

// vector.cpp:147-149: the parenthesisation gives cell offsets (+1, +2, 0)
static int linearizePosition(const V3& x, const double cw[3], const unsigned nC[3]) {
    return std::floor(x.x / cw[0]) + 1 + nC[0] * (std::floor(x.y / cw[1]) + 1 + (nC[1] * (std::floor(x.z / cw[2])) + 1));
}

void DEM::evalCellTable() {  // DEM.cpp:1326-1367
    cellTable.assign(particles.size() + (size_t)nCells[0] * nCells[1] * nCells[2], -1);
    for (size_t n = 0; n < particles.size(); ++n) {
        const int c = linearizePosition(particles[n].x0, cellWidth, nCells) + (int)particles.size();
        particles[n].tableCell = (uint32_t)c;
        if (c < 0 || (size_t)c >= cellTable.size()) {
            // the reference resets an element here and then writes out of bounds
            error = HYB_E_STATE;
            errmsg = "particle outside the neighbour-cell grid";
            return;
        }
        cellTable[n] = cellTable[c];
        cellTable[c] = (int)n;
    }
}

void DEM::evalNeighborTable() {  // DEM.cpp:1369-1494
    static const int shiftCell[14][3] = {{0, 0, 0},  {1, 0, 0},  {1, 1, 0},  {0, 1, 0},  {-1, 1, 0},
                                         {0, 0, 1},  {1, 0, 1},  {1, 1, 1},  {0, 1, 1},  {-1, 1, 1},
                                         {-1, 0, 1}, {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}};
    const double nebrRange2 = nebrRange * nebrRange;
    if (pbcs.size() != 0) {
        ghosts.clear();
        particles.resize(stdParticles);
        for (auto& e : elmts) e.components.resize(e.size);
        pbcShift();
        updateParticlesCorrected();
        createGhosts();
        updateParticlesCorrected();
        newNeighborList = true;
    }
    evalCellTable();
    if (error) return;
    neighborTable.clear();
    const uint32_t particleSize = (uint32_t)particles.size();
    unsigned i0[3];
    for (i0[0] = 0; i0[0] < nCells[0]; ++i0[0])
        for (i0[1] = 0; i0[1] < nCells[1]; ++i0[1])
            for (i0[2] = 0; i0[2] < nCells[2]; ++i0[2])
                for (int s = 0; s < 14; ++s) {
                    unsigned i1[3];
                    for (int k = 0; k < 3; ++k) {
                        i1[k] = i0[k] + shiftCell[s][k];  // unsigned: -1 wraps (DEM.cpp:1436-1448)
                        if (i1[k] == nCells[k]) i1[k] = 0;
                    }
                    const uint32_t cell0 = (i0[2] * nCells[1] + i0[1]) * nCells[0] + i0[0] + particleSize;
                    const uint32_t cell1 = (i1[2] * nCells[1] + i1[1]) * nCells[0] + i1[0] + particleSize;
                    int n0 = cellTable[cell0];
                    while (n0 > -1) {
                        if (cell1 >= cellTable.size()) {
                            error = HYB_E_STATE;
                            errmsg = "neighbour stencil left the cell table";
                            return;
                        }
                        int n1 = cellTable[cell1];
                        while (n1 > -1) {
                            if (cell0 != cell1 || n1 < n0) {
                                if (particles[n0].clusterIndex != particles[n1].clusterIndex) {
                                    const V3 r10 = particles[n1].x0 - particles[n0].x0;
                                    if (r10.norm2() < nebrRange2) {
                                        neighborTable.push_back((uint32_t)n0);
                                        neighborTable.push_back((uint32_t)n1);
                                    }
                                }
                            }
                            n1 = cellTable[n1];
                        }
                        n0 = cellTable[n0];
                    }
                }
    if (walls.size() != 0) evalNearWallTable();
    if (objects.size() != 0) evalNearObjectTable();
}

void DEM::evalNearObjectTable() {  // DEM.cpp:1515-1532: every object within nebrRange
    nearObjectTable.clear();
    const double nebrRange2 = nebrRange * nebrRange;
    for (uint32_t n = 0; n < stdParticles; ++n)
        for (size_t o = 0; o < objects.size(); ++o) {
            const V3 x0ij = particles[n].x0 - objects[o].x0;
            if (x0ij.norm2() < nebrRange2) {
                nearObjectTable.push_back(n);
                nearObjectTable.push_back((uint32_t)o);
            }
        }
}

void DEM::periodicObjects() {  // DEM.cpp:990-1062: copies of objects near a periodic plane
    struct G { uint32_t origin; V3 v; };
    std::vector<G> side;
    for (const Pbc& b : pbcs)
        for (uint32_t o = 0; o < stdObjects; ++o) {
            const double leftDist = b.pl1.dist(objects[o].x0);
            const double rightDist = b.pl2.dist(objects[o].x0);
            if (leftDist < 2.0 * nebrRange && leftDist > 0.0) side.push_back(G{objects[o].index, b.v});
            else if (rightDist < 2.0 * nebrRange && rightDist > 0.0) side.push_back(G{objects[o].index, -1.0 * b.v});
        }
    std::vector<G> corner;
    for (size_t a = 0; a < side.size(); ++a)
        for (size_t c = a + 1; c < side.size(); ++c)
            if (side[a].origin == side[c].origin) corner.push_back(G{side[a].origin, side[a].v + side[c].v});
    side.insert(side.end(), corner.begin(), corner.end());
    for (size_t g = 0; g < side.size(); ++g) {
        Object o = objects[side[g].origin];  // the reference indexes the list with the object's index
        o.x0 = o.x0 + side[g].v;
        o.originalIndex = o.index;
        o.index = stdObjects + (uint32_t)g;
        objects.push_back(o);
    }
}

void DEM::evalNearWallTable() {  // DEM.cpp:1496-1513
    nearWallTable.clear();
    for (uint32_t n = 0; n < stdParticles; ++n) {
        for (size_t w = 0; w < walls.size(); ++w) {
            if (walls[w].dist(particles[n].x0) < nebrRange) {
                nearWallTable.push_back(n);
                nearWallTable.push_back((uint32_t)w);
                if (!allWalls) break;
            }
        }
    }
}

void DEM::pbcShift() {  // DEM.cpp:1554-1584
    for (auto& b : pbcs) {
        const V3 v = b.v;
        for (auto& e : elmts) {
            const double leftDist = b.pl1.dist(e.x0);
            const double rightDist = b.pl2.dist(e.x0);
            if (leftDist < 0.0) { e.x0 += v; e.xp0 += v; }
            if (rightDist < 0.0) { const V3 m = -1.0 * v; e.x0 += m; e.xp0 += m; }
        }
    }
}

void DEM::createGhosts() {  // DEM.cpp:1586-1664
    ghosts.clear();
    for (auto& b : pbcs) {
        const V3 v = b.v;
        for (auto& e : elmts) {
            for (uint32_t j = 0; j < e.size; ++j) {
                const uint32_t pi = (uint32_t)e.components[j];
                const double leftDist = b.pl1.dist(particles[pi].x0);
                const double rightDist = b.pl2.dist(particles[pi].x0);
                if (leftDist < nebrRange) {
                    Ghost g; g.ghostIndex = particles[pi].particleIndex; g.pbcVector = v; ghosts.push_back(g);
                } else if (rightDist < nebrRange) {
                    Ghost g; g.ghostIndex = particles[pi].particleIndex; g.pbcVector = -1.0 * v; ghosts.push_back(g);
                }
            }
        }
    }
    std::vector<Ghost> corner;
    for (size_t g1 = 0; g1 < ghosts.size(); ++g1)
        for (size_t g2 = g1 + 1; g2 < ghosts.size(); ++g2)
            if (ghosts[g1].ghostIndex == ghosts[g2].ghostIndex) {
                Ghost g; g.ghostIndex = ghosts[g1].ghostIndex; g.pbcVector = ghosts[g1].pbcVector + ghosts[g2].pbcVector;
                corner.push_back(g);
            }
    ghosts.insert(ghosts.end(), corner.begin(), corner.end());
    particles.resize(stdParticles);
    for (auto& e : elmts) e.components.resize(e.size);
    uint32_t ind = stdParticles;
    for (auto& g : ghosts) {
        Particle dp = particles[g.ghostIndex];
        dp.particleIndex = ind;
        elmts[dp.clusterIndex].components.push_back((int)dp.particleIndex);
        particles.push_back(dp);
        ind++;
    }
}

// ---------------------------------------------------------------- integration
static void elmtPredict(Element& e, const double c1[], const double c2[]) {  // elmt.cpp:139-177
    e.xp0 = e.x0 + e.x1 * c2[0] + e.x2 * c2[1] + e.x3 * c2[2] + e.x4 * c2[3] + e.x5 * c2[4];
    e.xp1 = e.x1 + e.x2 * c2[0] + e.x3 * c2[1] + e.x4 * c2[2] + e.x5 * c2[3];
    e.xp2 = e.x2 + e.x3 * c2[0] + e.x4 * c2[1] + e.x5 * c2[2];
    e.xp3 = e.x3 + e.x4 * c2[0] + e.x5 * c2[1];
    e.xp4 = e.x4 + e.x5 * c2[0];
    e.xp5 = e.x5;
    e.qp0 = e.q0 + e.q1 * c2[0] + e.q2 * c2[1] + e.q3 * c2[2] + e.q4 * c2[3] + e.q5 * c2[4];
    e.qp1 = e.q1 + e.q2 * c2[0] + e.q3 * c2[1] + e.q4 * c2[2] + e.q5 * c2[3];
    e.qp2 = e.q2 + e.q3 * c2[0] + e.q4 * c2[1] + e.q5 * c2[2];
    e.qp3 = e.q3 + e.q4 * c2[0] + e.q5 * c2[1];
    e.qp4 = e.q4 + e.q5 * c2[0];
    e.qp5 = e.q5;
    e.qp0.normalize();
    e.wp0 = e.w0 + e.w1 * c1[0] + e.w2 * c1[1] + e.w3 * c1[2] + e.w4 * c1[3] + e.w5 * c1[4];
    e.wp1 = e.w1 + e.w2 * c1[0] + e.w3 * c1[1] + e.w4 * c1[2] + e.w5 * c1[3];
    e.wp2 = e.w2 + e.w3 * c1[0] + e.w4 * c1[1] + e.w5 * c1[2];
    e.wp3 = e.w3 + e.w4 * c1[0] + e.w5 * c1[1];
    e.wp4 = e.w4 + e.w5 * c1[0];
    e.wp5 = e.w5;
    e.wpGlobal = e.wp0;
    e.wpLocal = project(e.wpGlobal, e.qp0.adjoint());
}

static void elmtCorrect(Element& e, const double co1[], const double co2[]) {  // elmt.cpp:179-254
    const V3 x2Corr = e.x2 - e.xp2;
    e.x0 = e.xp0 + x2Corr * co2[0];
    e.x1 = e.xp1 + x2Corr * co2[1];
    e.x3 = e.xp3 + x2Corr * co2[3];
    e.x4 = e.xp4 + x2Corr * co2[4];
    e.x5 = e.xp5 + x2Corr * co2[5];
    e.xp0 = e.x0; e.xp1 = e.x1; e.xp2 = e.x2; e.xp3 = e.x3; e.xp4 = e.x4; e.xp5 = e.x5;
    const V3 w1Corr = e.w1 - e.wp1;
    e.w0 = e.wp0 + w1Corr * co1[0];
    e.w2 = e.wp2 + w1Corr * co1[2];
    e.w3 = e.wp3 + w1Corr * co1[3];
    e.w4 = e.wp4 + w1Corr * co1[4];
    e.w5 = e.wp5 + w1Corr * co1[5];
    e.wp0 = e.w0; e.wp1 = e.w1; e.wp2 = e.w2; e.wp3 = e.w3; e.wp4 = e.w4; e.wp5 = e.w5;
    const Q4 q2Corr = e.q2 - e.qp2;
    e.q0 = e.qp0 + q2Corr * co2[0];
    e.q1 = e.qp1 + q2Corr * co2[1];
    e.q3 = e.qp3 + q2Corr * co2[3];
    e.q4 = e.qp4 + q2Corr * co2[4];
    e.q5 = e.qp5 + q2Corr * co2[5];
    e.q0.normalize();
    e.qp0 = e.q0; e.qp1 = e.q1; e.qp2 = e.q2; e.qp3 = e.q3; e.qp4 = e.q4; e.qp5 = e.q5;
    e.wGlobal = e.w0;
    e.wLocal = project(e.wGlobal, e.q0.adjoint());
}

void DEM::predictor() { for (auto& e : elmts) elmtPredict(e, c1, c2); }           // DEM.cpp:1067-1080
void DEM::corrector() { for (auto& e : elmts) elmtCorrect(e, coeff1ord, coeff2ord); }  // DEM.cpp:1082-1112

void DEM::updateParticlesPredicted() {  // DEM.cpp:1184-1204
    for (uint32_t p = 0; p < stdParticles; ++p) particleUpdatePredicted(particles[p], elmts[particles[p].clusterIndex], prototypes);
    for (size_t g = 0; g < ghosts.size(); ++g)
        ghostUpdate(particles[stdParticles + g], particles[ghosts[g].ghostIndex], ghosts[g].pbcVector);
}

void DEM::updateParticlesCorrected() {  // DEM.cpp:1206-1227
    for (uint32_t p = 0; p < stdParticles; ++p) particleUpdateCorrected(particles[p], elmts[particles[p].clusterIndex], prototypes);
    for (size_t g = 0; g < ghosts.size(); ++g)
        ghostUpdate(particles[stdParticles + g], particles[ghosts[g].ghostIndex], ghosts[g].pbcVector);
}

// ---------------------------------------------------------------- contacts
double DEM::normalContact(double overlap, double vrel, double effRad, double effMass) const {  // DEM.cpp:2138-2174
    double fn = 0.0;
    if (mat.contactModel == HYB_CONTACT_HERTZIAN) {
        const double sqrtEffRad = std::sqrt(effRad);
        const double sqrtOverlap = std::sqrt(overlap);
        const double kn = mat.knConst * sqrtEffRad * sqrtOverlap;
        const double gamman = 2.0 * mat.dampCoeff * std::sqrt(kn * effMass);
        const double dumpfn = -gamman * vrel;
        const double elasticfn = kn * overlap;
        fn = std::max(elasticfn + dumpfn, 0.0);
    } else {
        const double elasticfn = mat.linearStiff * overlap;
        const double gamman = 2.0 * mat.dampCoeff * std::sqrt(mat.linearStiff * effMass);
        const double dumpfn = -gamman * vrel;
        fn = std::max(elasticfn + dumpfn, 0.0);
    }
    return fn;
}

double DEM::tangentialContact(double vrelt, double fn, double effRad, double effMass, double friction) const {
    // DEM.cpp:2176-2222: viscous tangential force capped by Coulomb friction (no spring, no history)
    double fs = 0.0;
    if (mat.contactModel == HYB_CONTACT_HERTZIAN) {
        const double power = 1.0 / 3.0;
        const double sqrtEffRad = std::sqrt(effRad);
        const double ks = mat.ksConst * sqrtEffRad * std::pow(std::abs(fn), power);
        const double fsMax = friction * fn;
        const double gammas = 2.0 * mat.viscTang * std::sqrt(effMass * ks);
        const double dumpfs = gammas * vrelt;
        fs = std::min(dumpfs, fsMax);
    } else {
        const double ks = mat.linearStiff;
        const double fsMax = friction * fn;
        const double gammas = 2.0 * mat.viscTang * std::sqrt(effMass * ks);
        const double dumpfs = gammas * vrelt;
        fs = std::min(dumpfs, fsMax);
    }
    return fs;
}

void DEM::particleParticleCollision(const Particle& pI, const Particle& pJ, const V3& vd) {  // DEM.cpp:1801-1900
    Element& eI = elmts[pI.clusterIndex];
    Element& eJ = elmts[pJ.clusterIndex];
    const double radI = pI.r, radJ = pJ.r;
    const double distance = vd.norm();
    const double overlap = radI + radJ - distance;
    const V3 relVel = pJ.x1 - pI.x1;
    const V3 en = vd / distance;
    const double normRelVel = relVel.dot(en);
    const V3 normalRelVel = en * normRelVel;
    const double effMass = eI.m * eJ.m / (eI.m + eJ.m);
    const double effRad = radI * radJ / (radI + radJ);
    const double fn = normalContact(overlap, normRelVel, effRad, effMass);
    const V3 normalForce = en * fn;
    const V3 vecRadI = radI * en;
    const V3 vecRadJ = -radJ * en;
    const V3 centerDistI = vecRadI + pI.radiusVec;
    const V3 centerDistJ = vecRadJ + pJ.radiusVec;
    if (pI.particleIndex < stdParticles) {
        eI.FParticle = eI.FParticle - normalForce;
        if (eI.size > 1) eI.MParticle = eI.MParticle - centerDistI.cross(normalForce);
    }
    if (pJ.particleIndex < stdParticles) {
        eJ.FParticle = eJ.FParticle + normalForce;
        if (eJ.size > 1) eJ.MParticle = eJ.MParticle + centerDistJ.cross(normalForce);
    }
    const V3 wI = eI.wpGlobal, wJ = eJ.wpGlobal;
    const V3 relVelContact = relVel - wI.cross(vecRadI) + wJ.cross(vecRadJ);
    const V3 tangRelVelContact = relVelContact - normalRelVel;
    const double nt = tangRelVelContact.norm();
    if (nt != 0.0) {
        const double ft = tangentialContact(nt, fn, effRad, effMass, mat.frictionCoefPart);
        const V3 et = tangRelVelContact / nt;
        const V3 tangForce = ft * et;
        if (pI.particleIndex < stdParticles) {
            eI.MParticle = eI.MParticle + centerDistI.cross(tangForce);
            eI.FParticle = eI.FParticle + tangForce;
        }
        if (pJ.particleIndex < stdParticles) {
            eJ.MParticle = eJ.MParticle - centerDistJ.cross(tangForce);
            eJ.FParticle = eJ.FParticle - tangForce;
        }
    }
}

void DEM::wallParticleCollision(Wall& w, const Particle& pJ, double overlap) {  // DEM.cpp:1902-1982
    Element& eJ = elmts[pJ.clusterIndex];
    const double radJ = pJ.r;
    const V3 en = w.n;
    const V3 cpv = w.getSpeed(pJ.x0);
    const V3 relVel = pJ.x1 - cpv;
    const double normRelVel = relVel.dot(en);
    const V3 normalRelVel = en * normRelVel;
    const double fn = normalContact(2.0 * overlap, normRelVel, radJ, eJ.m);
    const V3 normalForce = en * fn;
    const V3 vecRadJ = -radJ * en;
    V3 centerDistJ = vecRadJ;
    if (eJ.size > 1) centerDistJ = centerDistJ + (pJ.x0 - eJ.xp0);
    eJ.FWall = eJ.FWall + normalForce;
    w.FParticle = w.FParticle - normalForce;
    if (eJ.size > 1) eJ.MWall = eJ.MWall + centerDistJ.cross(normalForce);
    const V3 wJ = eJ.wpGlobal;
    const V3 relVelContact = relVel + wJ.cross(vecRadJ);
    const V3 tangRelVelContact = relVelContact - normalRelVel;
    const double nt = tangRelVelContact.norm();
    if (nt != 0.0) {
        const double ft = tangentialContact(nt, fn, radJ, eJ.m, mat.frictionCoefWall);
        const V3 et = tangRelVelContact / tangRelVelContact.norm();
        const V3 tangForce = ft * et;
        eJ.MWall = eJ.MWall - centerDistJ.cross(tangForce);
        eJ.FWall = eJ.FWall - tangForce;
        w.FParticle = w.FParticle + tangForce;
    }
}

// DEM.cpp:2060-2136: the object is static (its x1 is the contact-point velocity)
void DEM::objectParticleCollision(Object& o, const Particle& pJ, const V3& vectorDistance) {
    Element& eJ = elmts[pJ.clusterIndex];
    const double radJ = pJ.r;
    const double distance = vectorDistance.norm();
    const double overlap = radJ + o.r - distance;
    const V3 en = (1.0 / distance) * vectorDistance;
    const V3 contactPointVelocity = o.x1;
    const V3 relVel = pJ.x1 - contactPointVelocity;
    const double normRelVel = relVel.dot(en);
    const V3 normalRelVel = en * normRelVel;
    const double fn = normalContact(2.0 * overlap, normRelVel, radJ, eJ.m);
    const V3 normalForce = en * fn;
    const V3 vecRadJ = -radJ * en;
    const V3 centerDistJ = vecRadJ + (pJ.x0 - eJ.xp0);
    eJ.FWall = eJ.FWall + normalForce;
    o.FParticle = o.FParticle - normalForce;
    if (eJ.size > 1) eJ.MWall = eJ.MWall + centerDistJ.cross(normalForce);
    const V3 wJ = eJ.wpGlobal;
    const V3 relVelContact = relVel + wJ.cross(vecRadJ);
    const V3 tangRelVelContact = relVelContact - normalRelVel;
    const double nt = tangRelVelContact.norm();
    if (nt != 0.0) {
        const double ft = tangentialContact(nt, fn, radJ, eJ.m, mat.frictionCoefPart);
        const V3 et = tangRelVelContact / tangRelVelContact.norm();
        const V3 tangForce = ft * et;
        eJ.MWall = eJ.MWall - centerDistJ.cross(tangForce);
        eJ.FWall = eJ.FWall - tangForce;
        o.FParticle = o.FParticle + tangForce;
    }
}

void DEM::evaluateForces() {  // DEM.cpp:1114-1182
    for (auto& e : elmts) { e.FParticle.reset(); e.FWall.reset(); e.MParticle.reset(); e.MWall.reset(); }
    for (auto& w : walls) w.FParticle.reset();
    for (auto& o : objects) o.FParticle.reset();
    // particleParticleContacts, DEM.cpp:1668-1690
    for (size_t k = 0; k + 1 < neighborTable.size(); k += 2) {
        const Particle& pi = particles[neighborTable[k]];
        const Particle& pj = particles[neighborTable[k + 1]];
        const double sigij = pi.r + pj.r;
        const double sigij2 = sigij * sigij;
        const V3 vd = pj.x0 - pi.x0;
        if (vd.norm2() < sigij2) particleParticleCollision(pi, pj, vd);
    }
    // wallParticleContacts, DEM.cpp:1692-1717
    for (size_t k = 0; k + 1 < nearWallTable.size(); k += 2) {
        const Particle& pj = particles[nearWallTable[k]];
        Wall& w = walls[nearWallTable[k + 1]];
        const double distance = w.dist(pj.x0);
        const double overlap = pj.r - distance;
        if (overlap > 0.0) wallParticleCollision(w, pj, overlap);
    }
    // objectParticleContacts, DEM.cpp:1747-1770
    for (size_t k = 0; k + 1 < nearObjectTable.size(); k += 2) {
        const Particle& pj = particles[nearObjectTable[k]];
        Object& o = objects[nearObjectTable[k + 1]];
        const V3 vectorDistance = pj.x0 - o.x0;
        const double distance = vectorDistance.norm();
        const double overlap = pj.r + o.r - distance;
        if (overlap > 0.0) objectParticleCollision(o, pj, vectorDistance);
    }
    for (auto& e : elmts) {
        const V3 FVisc = -6.0 * M_PI * numVisc * e.radius * e.xp1;
        const V3 MVisc = -8.0 * M_PI * numVisc * e.radius * e.radius * e.radius * e.wpGlobal;
        e.x2 = (FVisc + e.FHydro + e.FParticle + e.FWall) / e.m + demF;
        const V3 moment = MVisc + e.MHydro + e.MParticle + e.MWall;
        const V3 momentBf = project(moment, e.qp0.adjoint());
        const V3 waBf = newtonAcc(momentBf, e.I, e.wpLocal);
        e.w1 = project(waBf, e.qp0);
        if (e.size > 1) {
            const Q4 waQuat = quatAcc(waBf, e.qp1);
            e.q2 = 0.5 * e.qp0.multiply(waQuat);
        }
    }
}

void DEM::step() {  // DEM.cpp:331-376
    if (!first_step_done) { neighListTrigger = 0.25 * nebrRange; first_step_done = true; }
    for (uint32_t it = 0; it < multiStep; ++it) {
        demTimeStep++;
        demTime += deltat;
        evalMaxDisp();
        if (maxDisp > neighListTrigger) {
            maxDisp = 0.0;
            evalNeighborTable();
            if (error) return;
        }
        predictor();
        updateParticlesPredicted();
        evaluateForces();
        corrector();
        updateParticlesCorrected();
        updateEnergy();
    }
}

// DEM.cpp:2226-2276 with energy::updateTotal (utils.cpp:324-327): kinetic sums in
// element order, potential energy above the box corner the gravity points away from.
void DEM::updateEnergy() {
    double tKin = 0.0, rKin = 0.0;
    for (const Element& e : elmts) {
        tKin += 0.5 * e.m * e.x1.norm2();
        const V3 w2(e.wLocal.x * e.wLocal.x, e.wLocal.y * e.wLocal.y, e.wLocal.z * e.wLocal.z);
        rKin += e.I.dot(w2);
    }
    double g = 0.0;
    const double gn = demF.norm();
    if (gn != 0.0) {
        const V3 n = (demF * -1.0) / gn;
        V3 p(0.0, 0.0, 0.0);
        if (n.dot(V3(1.0, 0.0, 0.0)) < 0) p += V3(demSize[0], 0.0, 0.0);
        if (n.dot(V3(0.0, 1.0, 0.0)) < 0) p += V3(0.0, demSize[1], 0.0);
        if (n.dot(V3(0.0, 0.0, 1.0)) < 0) p += V3(0.0, 0.0, demSize[2]);
        for (const Element& e : elmts) {
            const double h = n.dot(e.x0 - p);
            g += e.m * h * gn;
        }
    }
    const double elastic = 0.0;  // not computed by the reference
    energy[0] = rKin + tKin;
    energy[1] = rKin;
    energy[2] = tKin;
    energy[3] = g;
    energy[4] = elastic;
    energy[5] = rKin + tKin + g + elastic;
}

// ---------------------------------------------------------------- cycle
void Sim::cycle() {  // hybird.cpp:35-66 (IO excluded)
    ++currentTimeStep;
    ++lb.time;
    if (demSolve) dem.step();
    if (dem.error) return;
    if (lbmSolve && dem.demTime > dem.demInitialRepeat) {
        if (lb.freeSurface) lb.freeSurfaceStep();
        if (demSolve) lb.couplingStep(dem);
        lb.step(dem);
    } else if (lbmSolve && dem.demTime <= dem.demInitialRepeat) {
        if (demSolve) lb.couplingStep(dem);
    }
}

}  // namespace orc
