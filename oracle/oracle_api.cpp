// ORACLE — test infrastructure, not the product.
//
// C API of liboracle.so (loaded with ctypes by tests/ and by bench.py's CPU
// baseline leg). Field names and layouts match the golden fixtures written by
// tests/golden/make_golden.py from the compiled reference.
//
// Pinning of the restatement (tests/test_oracle_golden.py and
// tests/test_oracle_invariants.py):
//  - lattice init, types, neighbour rewiring, hydrostatic state: bit-exact vs the
//    reference fixtures (lb16, chan, couple2, dense, fs, bingham)
//  - reconstruction / BGK / Guo / streaming / bounce-back: bit-exact vs fixtures
//    after 1..10 cycles; plus closed forms (equilibrium moments, mass conservation)
//  - Bingham viscosity: bit-exact vs the bingham fixture
//  - coupling (classification, penalty force, element sums): bit-exact vs couple2, dense
//  - DEM pair lists, near-wall lists, ghosts: bit-exact vs dem60, dense, couple2;
//    pair set also against an O(P^2) brute force with the reference's cell quirks
//  - Hertz/tangential contacts, Gear integration: fixtures (dem60, dense);
//    Gear exactness for a free-fall parabola
//  - free surface mass tracking and interface mutation: bit-exact vs fs
//  - TRT: reduces to BGK at tau = 1 (closed form); otherwise "parity unpinned" by the
//    reference, checked against the Poiseuille profile.
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"

using namespace orc;

namespace {

template <class T>
long put(const std::vector<T>& v, void* out, long cap) {
    const long need = (long)(v.size() * sizeof(T));
    if (out && cap >= need && need) std::memcpy(out, v.data(), (size_t)need);
    return need;
}

void p3(std::vector<double>& a, const V3& v) { a.push_back(v.x); a.push_back(v.y); a.push_back(v.z); }
void p4(std::vector<double>& a, const Q4& q) { a.push_back(q.q0); a.push_back(q.q1); a.push_back(q.q2); a.push_back(q.q3); }

}  // namespace

extern "C" {

void* orc_create(const hyb_config* cfg, int n_elements, const double* records, int n_objects, const double* objects,
                 char* err, int errlen) {
    Sim* s = new Sim();
    s->cfg = *cfg;
    s->demSolve = cfg->demSolve;
    s->lbmSolve = cfg->lbSolve;
    s->lb.get(*cfg);
    s->dem.get(*cfg, n_elements, records);
    for (int o = 0; o < n_objects; ++o) {  // DEM.cpp:136-173 (positions are not scaled or translated)
        const double* r = objects + (size_t)o * HYB_OBJECT_RECORD;
        Object ob;
        ob.index = (uint32_t)r[0];
        ob.r = r[2];
        ob.x0 = V3(r[3], r[4], r[5]);
        ob.x1 = V3(r[6], r[7], r[8]);
        ob.originalIndex = ob.index;
        s->dem.objects.push_back(ob);
    }
    s->lb.latticeDefinition();
    s->dem.init(s->lb.boundary, s->lb.lbSize, s->lb.unit, s->lb.lbF);
    if (!s->dem.error && s->lbmSolve) s->lb.init(s->dem);
    s->lb.time = 0;
    s->dem.demTime = 0.0;
    s->dem.demTimeStep = 0;
    const int e = s->dem.error ? s->dem.error : s->lb.error;
    if (e) {
        if (err && errlen > 0) {
            std::snprintf(err, errlen, "%s", s->dem.error ? s->dem.errmsg.c_str() : s->lb.errmsg.c_str());
        }
        delete s;
        return nullptr;
    }
    return s;
}

void orc_destroy(void* h) { delete static_cast<Sim*>(h); }

int orc_cycle(void* h, int n) {
    Sim* s = static_cast<Sim*>(h);
    for (int k = 0; k < n; ++k) {
        s->cycle();
        if (s->dem.error) return s->dem.error;
        if (s->lb.error) return s->lb.error;
    }
    return 0;
}

// LB steps alone (LB.cpp:221-233): no DEM sub-steps and no coupling step, so
// the particles stay where they are (the frozen-bed model of the adjoint).
int orc_lb_step(void* h, int n) {
    Sim* s = static_cast<Sim*>(h);
    for (int k = 0; k < n; ++k) {
        s->lb.step(s->dem);
        if (s->lb.error) return s->lb.error;
    }
    return 0;
}

const char* orc_error(void* h) {
    Sim* s = static_cast<Sim*>(h);
    return s->dem.error ? s->dem.errmsg.c_str() : s->lb.errmsg.c_str();
}

long orc_get(void* h, const char* cname, void* out, long cap) {
    Sim* s = static_cast<Sim*>(h);
    const std::string name(cname);
    LB& lb = s->lb;
    DEM& dem = s->dem;
    const size_t N = lb.t.size();
    if (name == "types") return put(lb.t, out, cap);
    if (name == "pflag") return put(lb.p, out, cap);
    if (name == "solidIndex") return put(lb.solidIndex, out, cap);
    if (name == "f" || name == "n" || name == "u" || name == "hydroForce" || name == "mass" || name == "visc") {
        const int w = (name == "f") ? 19 : (name == "u" || name == "hydroForce") ? 3 : 1;
        std::vector<double> v(N * w, NAN);
        for (size_t it = 0; it < N; ++it) {
            if (!lb.alloc[it]) continue;
            const Node& nd = lb.nodes[it];
            if (name == "f") for (int j = 0; j < 19; ++j) v[it * 19 + j] = nd.f[j];
            else if (name == "n") v[it] = nd.n;
            else if (name == "u") { v[it * 3] = nd.u.x; v[it * 3 + 1] = nd.u.y; v[it * 3 + 2] = nd.u.z; }
            else if (name == "hydroForce") { v[it * 3] = nd.hydroForce.x; v[it * 3 + 1] = nd.hydroForce.y; v[it * 3 + 2] = nd.hydroForce.z; }
            else if (name == "mass") v[it] = nd.mass;
            else v[it] = nd.visc;
        }
        return put(v, out, cap);
    }
    if (name == "activeNodes") return put(lb.activeNodes, out, cap);
    if (name == "fluidNodes") return put(lb.fluidNodes, out, cap);
    if (name == "interfaceNodes") return put(lb.interfaceNodes, out, cap);
    if (name == "particleNodes") return put(lb.particleNodes, out, cap);
    if (name == "neighbors") return put(lb.nb, out, cap);
    if (name == "totalMass") return put(std::vector<double>{lb.totalMass}, out, cap);
    if (name == "fluidMass") {  // IO::totFluidMass (IO.cpp:987-999): active nodes outside particles, index order
        double m = 0.0;
        for (size_t it = 0; it < N; ++it)
            if ((lb.t[it] == 0 || lb.t[it] == 3) && !lb.p[it]) m += lb.nodes[it].mass;
        return put(std::vector<double>{m}, out, cap);
    }
    if (name == "lbF") { std::vector<double> v; p3(v, lb.lbF); return put(v, out, cap); }
    if (name == "lbSize") return put(std::vector<uint32_t>(lb.lbSize.begin(), lb.lbSize.end()), out, cap);
    if (name.rfind("e_", 0) == 0) {
        std::vector<double> v;
        std::vector<uint32_t> u;
        for (const Element& e : dem.elmts) {
            if (name == "e_x") { p3(v, e.x0); p3(v, e.x1); p3(v, e.x2); p3(v, e.x3); p3(v, e.x4); p3(v, e.x5); }
            else if (name == "e_xp") { p3(v, e.xp0); p3(v, e.xp1); p3(v, e.xp2); p3(v, e.xp3); p3(v, e.xp4); p3(v, e.xp5); }
            else if (name == "e_q") { p4(v, e.q0); p4(v, e.q1); p4(v, e.q2); p4(v, e.q3); p4(v, e.q4); p4(v, e.q5); }
            else if (name == "e_qp") { p4(v, e.qp0); p4(v, e.qp1); p4(v, e.qp2); p4(v, e.qp3); p4(v, e.qp4); p4(v, e.qp5); }
            else if (name == "e_w") { p3(v, e.w0); p3(v, e.w1); p3(v, e.w2); p3(v, e.w3); p3(v, e.w4); p3(v, e.w5); }
            else if (name == "e_wp") { p3(v, e.wp0); p3(v, e.wp1); p3(v, e.wp2); p3(v, e.wp3); p3(v, e.wp4); p3(v, e.wp5); }
            else if (name == "e_misc") {
                v.push_back(e.m); v.push_back(e.radius); p3(v, e.I); p3(v, e.wGlobal); p3(v, e.wLocal); p3(v, e.wpGlobal);
                p3(v, e.wpLocal); v.push_back(e.fluidVolume); v.push_back((double)e.size);
            } else if (name == "e_forces") {
                p3(v, e.FHydro); p3(v, e.MHydro); p3(v, e.FParticle); p3(v, e.MParticle); p3(v, e.FWall); p3(v, e.MWall); p3(v, e.FGrav);
            } else if (name == "e_comp") { for (int c : e.components) u.push_back((uint32_t)c); }
        }
        if (name == "e_comp") return put(u, out, cap);
        if (name == "e_comp_off") {
            uint32_t o = 0;
            for (const Element& e : dem.elmts) { u.push_back(o); o += (uint32_t)e.components.size(); }
            u.push_back(o);
            return put(u, out, cap);
        }
        return put(v, out, cap);
    }
    if (name == "p_data") {
        std::vector<double> v;
        for (const Particle& p : dem.particles) { p3(v, p.x0); p3(v, p.x1); p3(v, p.radiusVec); v.push_back(p.r); }
        return put(v, out, cap);
    }
    if (name == "p_idx") {
        std::vector<uint32_t> v;
        for (const Particle& p : dem.particles) { v.push_back(p.particleIndex); v.push_back(p.clusterIndex); v.push_back(p.protoIndex); v.push_back(p.tableCell); }
        return put(v, out, cap);
    }
    if (name == "neighborTable") return put(dem.neighborTable, out, cap);
    if (name == "nearWallTable") return put(dem.nearWallTable, out, cap);
    if (name == "cellTable") return put(dem.cellTable, out, cap);
    if (name == "ghostIndex") { std::vector<uint32_t> v; for (auto& g : dem.ghosts) v.push_back(g.ghostIndex); return put(v, out, cap); }
    if (name == "ghostVec") { std::vector<double> v; for (auto& g : dem.ghosts) p3(v, g.pbcVector); return put(v, out, cap); }
    if (name == "dem_scalars") {
        std::vector<double> v = {dem.maxDisp, dem.nebrRange, dem.cellWidth[0], dem.cellWidth[1], dem.cellWidth[2],
                                 (double)dem.nCells[0], (double)dem.nCells[1], (double)dem.nCells[2], dem.demTime,
                                 (double)dem.demTimeStep, dem.deltat, (double)dem.multiStep, (double)dem.stdParticles,
                                 dem.mat.dampCoeff, dem.mat.knConst, dem.mat.ksConst, dem.newNeighborList ? 1.0 : 0.0};
        return put(v, out, cap);
    }
    if (name == "energy") return put(std::vector<double>(dem.energy, dem.energy + 6), out, cap);
    if (name == "nearObjectTable") return put(dem.nearObjectTable, out, cap);
    if (name == "objects") {
        std::vector<double> v;
        for (auto& o : dem.objects) { v.push_back(o.index); v.push_back(o.r); p3(v, o.x0); p3(v, o.x1); p3(v, o.FParticle); }
        return put(v, out, cap);
    }
    if (name == "walls") {
        std::vector<double> v;
        for (auto& w : dem.walls) { p3(v, w.p); p3(v, w.n); p3(v, w.FParticle); p3(v, w.FHydro); }
        return put(v, out, cap);
    }
    return -1;
}

}  // extern "C"
