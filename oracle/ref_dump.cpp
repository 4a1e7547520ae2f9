// Checker harness: drives the compiled reference (hybird) through the same call
// sequence as its main() (hybird.cpp:245-342) and dumps the full FP64 state at
// chosen cycles. It is test infrastructure only: it is built from the reference
// sources under /root/reference by oracle/Makefile into oracle/_ref/ and is used
// to write the golden fixtures under tests/golden/ and to pin the oracle.
//
// Usage: hybird_ref_dump <cfg> <workdir> <outprefix> <cycle,cycle,...>
// Writes <outprefix>_<cycle>.bin for every listed cycle (0 = after init).
// With outprefix "-" nothing is dumped: the harness runs up to the largest
// listed cycle and prints the steady-state seconds per cycle, (T(n) - T(1))/(n-1),
// as one JSON line (CPU-baseline timing).
//
// Format of a .bin file: repeated records
//   u32 name_len | name bytes | u8 dtype ('d' f64, 'I' u32, 'B' u8, 'i' i32)
//   | u32 ndim | u64 shape[ndim] | raw little-endian data

#include <algorithm>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>
#include <sys/stat.h>

#define private public
#include "IO.h"
#include "DEM.h"
#include "LB.h"
#undef private

void goCycle(IO& io, DEM& dem, LB& lb);
void parseCommandLine(IO& io, GetPot& command_line);
void parseConfigFile(IO& io, DEM& dem, LB& lb, GetPot& lbmCfgFile, GetPot& command_line);

namespace {

struct Writer {
    FILE* fp;
    explicit Writer(const std::string& path) { fp = std::fopen(path.c_str(), "wb"); }
    ~Writer() { if (fp) std::fclose(fp); }
    void rec(const char* name, char dt, const void* data, std::vector<uint64_t> shape, size_t elsize) {
        uint32_t nl = (uint32_t)std::strlen(name);
        std::fwrite(&nl, 4, 1, fp);
        std::fwrite(name, 1, nl, fp);
        std::fwrite(&dt, 1, 1, fp);
        uint32_t nd = (uint32_t)shape.size();
        std::fwrite(&nd, 4, 1, fp);
        uint64_t count = 1;
        for (uint64_t s : shape) { std::fwrite(&s, 8, 1, fp); count *= s; }
        if (count) std::fwrite(data, elsize, count, fp);
    }
    void d(const char* name, const std::vector<double>& v, std::vector<uint64_t> shape) { rec(name, 'd', v.data(), shape, 8); }
    void u(const char* name, const std::vector<uint32_t>& v) { rec(name, 'I', v.data(), {v.size()}, 4); }
    void b(const char* name, const std::vector<uint8_t>& v) { rec(name, 'B', v.data(), {v.size()}, 1); }
    void i(const char* name, const std::vector<int32_t>& v) { rec(name, 'i', v.data(), {v.size()}, 4); }
    void scalar(const char* name, double x) { std::vector<double> v(1, x); d(name, v, {1}); }
};

void push3(std::vector<double>& a, const tVect& v) { a.push_back(v.x); a.push_back(v.y); a.push_back(v.z); }
void push4(std::vector<double>& a, const tQuat& q) { a.push_back(q.q0); a.push_back(q.q1); a.push_back(q.q2); a.push_back(q.q3); }

void dump(const std::string& path, IO& io, DEM& dem, LB& lb) {
    Writer w(path);
    const uint64_t N = lb.types.size();
    // lattice
    {
        std::vector<uint8_t> t(N), p(N);
        std::vector<uint32_t> si(N);
        std::vector<double> f(N * 19, NAN), n(N, NAN), u(N * 3, NAN), hf(N * 3, NAN), mass(N, NAN), visc(N, NAN);
        for (uint64_t it = 0; it < N; ++it) {
            t[it] = (uint8_t)lb.types[it].getType();
            p[it] = lb.types[it].isInsideParticle() ? 1 : 0;
            si[it] = lb.types[it].getSolidIndex();
            const node* nd = lb.nodes.size() ? lb.nodes[it] : nullptr;
            if (nd) {
                for (int j = 0; j < 19; ++j) f[it * 19 + j] = nd->f[j];
                n[it] = nd->n;
                u[it * 3 + 0] = nd->u.x; u[it * 3 + 1] = nd->u.y; u[it * 3 + 2] = nd->u.z;
                hf[it * 3 + 0] = nd->hydroForce.x; hf[it * 3 + 1] = nd->hydroForce.y; hf[it * 3 + 2] = nd->hydroForce.z;
                mass[it] = nd->mass;
                visc[it] = nd->visc;
            }
        }
        w.b("types", t);
        w.b("pflag", p);
        w.u("solidIndex", si);
        w.d("f", f, {N, 19});
        w.d("n", n, {N});
        w.d("u", u, {N, 3});
        w.d("hydroForce", hf, {N, 3});
        w.d("mass", mass, {N});
        w.d("visc", visc, {N});
        std::vector<uint32_t> act(lb.activeNodes.begin(), lb.activeNodes.end());
        w.u("activeNodes", act);
        std::vector<uint32_t> pn(lb.particleNodes.begin(), lb.particleNodes.end());
        w.u("particleNodes", pn);
        std::vector<uint32_t> in(lb.interfaceNodes.begin(), lb.interfaceNodes.end());
        w.u("interfaceNodes", in);
        std::vector<uint32_t> fl(lb.fluidNodes.begin(), lb.fluidNodes.end());
        w.u("fluidNodes", fl);
        std::vector<uint32_t> nb;
        if (lb.neighbors.size()) {
            nb.resize(N * 19);
            for (uint64_t it = 0; it < N; ++it)
                for (int j = 0; j < 19; ++j) nb[it * 19 + j] = lb.neighbors[it].d[j];
        }
        w.u("neighbors", nb);
        std::vector<uint32_t> sz(lb.lbSize.begin(), lb.lbSize.end());
        w.u("lbSize", sz);
        std::vector<double> lbF; push3(lbF, lb.lbF);
        w.d("lbF", lbF, {3});
        w.scalar("totalMass", lb.totalMass);
        w.scalar("lbTime", (double)lb.time);
    }
    // elements
    {
        const uint64_t E = dem.elmts.size();
        std::vector<double> x(E * 3 * 6), xp(E * 3 * 6), q(E * 4 * 6), qp(E * 4 * 6), wv(E * 3 * 6), wp(E * 3 * 6);
        std::vector<double> misc;
        std::vector<double> forces;
        std::vector<uint32_t> comps_off, comps;
        for (uint64_t e = 0; e < E; ++e) {
            const elmt& el = dem.elmts[e];
            const tVect xs[6] = {el.x0, el.x1, el.x2, el.x3, el.x4, el.x5};
            const tVect xps[6] = {el.xp0, el.xp1, el.xp2, el.xp3, el.xp4, el.xp5};
            const tQuat qs[6] = {el.q0, el.q1, el.q2, el.q3, el.q4, el.q5};
            const tQuat qps[6] = {el.qp0, el.qp1, el.qp2, el.qp3, el.qp4, el.qp5};
            const tVect ws[6] = {el.w0, el.w1, el.w2, el.w3, el.w4, el.w5};
            const tVect wps[6] = {el.wp0, el.wp1, el.wp2, el.wp3, el.wp4, el.wp5};
            for (int k = 0; k < 6; ++k) {
                x[(e * 6 + k) * 3 + 0] = xs[k].x; x[(e * 6 + k) * 3 + 1] = xs[k].y; x[(e * 6 + k) * 3 + 2] = xs[k].z;
                xp[(e * 6 + k) * 3 + 0] = xps[k].x; xp[(e * 6 + k) * 3 + 1] = xps[k].y; xp[(e * 6 + k) * 3 + 2] = xps[k].z;
                wv[(e * 6 + k) * 3 + 0] = ws[k].x; wv[(e * 6 + k) * 3 + 1] = ws[k].y; wv[(e * 6 + k) * 3 + 2] = ws[k].z;
                wp[(e * 6 + k) * 3 + 0] = wps[k].x; wp[(e * 6 + k) * 3 + 1] = wps[k].y; wp[(e * 6 + k) * 3 + 2] = wps[k].z;
                q[(e * 6 + k) * 4 + 0] = qs[k].q0; q[(e * 6 + k) * 4 + 1] = qs[k].q1; q[(e * 6 + k) * 4 + 2] = qs[k].q2; q[(e * 6 + k) * 4 + 3] = qs[k].q3;
                qp[(e * 6 + k) * 4 + 0] = qps[k].q0; qp[(e * 6 + k) * 4 + 1] = qps[k].q1; qp[(e * 6 + k) * 4 + 2] = qps[k].q2; qp[(e * 6 + k) * 4 + 3] = qps[k].q3;
            }
            // misc: m, radius, I(3), wGlobal(3), wLocal(3), wpGlobal(3), wpLocal(3), fluidVolume, size
            misc.push_back(el.m); misc.push_back(el.radius);
            push3(misc, el.I); push3(misc, el.wGlobal); push3(misc, el.wLocal); push3(misc, el.wpGlobal); push3(misc, el.wpLocal);
            misc.push_back(el.fluidVolume); misc.push_back((double)el.size);
            // forces: FHydro, MHydro, FParticle, MParticle, FWall, MWall, FGrav
            push3(forces, el.FHydro); push3(forces, el.MHydro); push3(forces, el.FParticle); push3(forces, el.MParticle);
            push3(forces, el.FWall); push3(forces, el.MWall); push3(forces, el.FGrav);
            comps_off.push_back((uint32_t)comps.size());
            for (size_t c = 0; c < el.components.size(); ++c) comps.push_back((uint32_t)el.components[c]);
        }
        comps_off.push_back((uint32_t)comps.size());
        w.d("e_x", x, {E, 6, 3});
        w.d("e_xp", xp, {E, 6, 3});
        w.d("e_q", q, {E, 6, 4});
        w.d("e_qp", qp, {E, 6, 4});
        w.d("e_w", wv, {E, 6, 3});
        w.d("e_wp", wp, {E, 6, 3});
        w.d("e_misc", misc, {E, 19});
        w.d("e_forces", forces, {E, 7, 3});
        w.u("e_comp_off", comps_off);
        w.u("e_comp", comps);
    }
    // particles
    {
        const uint64_t P = dem.particles.size();
        std::vector<double> pd;
        std::vector<uint32_t> pi;
        for (uint64_t p = 0; p < P; ++p) {
            const particle& pa = dem.particles[p];
            push3(pd, pa.x0); push3(pd, pa.x1); push3(pd, pa.radiusVec); pd.push_back(pa.r);
            pi.push_back(pa.particleIndex); pi.push_back(pa.clusterIndex); pi.push_back(pa.protoIndex); pi.push_back(pa.tableCell);
        }
        w.d("p_data", pd, {P, 10});
        w.rec("p_idx", 'I', pi.data(), {P, 4}, 4);
        std::vector<uint32_t> nt(dem.neighborTable.begin(), dem.neighborTable.end());
        w.u("neighborTable", nt);
        std::vector<uint32_t> nw(dem.nearWallTable.begin(), dem.nearWallTable.end());
        w.u("nearWallTable", nw);
        std::vector<int32_t> ct(dem.cellTable.begin(), dem.cellTable.end());
        w.i("cellTable", ct);
        std::vector<double> gh;
        std::vector<uint32_t> ghi;
        for (size_t g = 0; g < dem.ghosts.size(); ++g) { ghi.push_back(dem.ghosts[g].ghostIndex); push3(gh, dem.ghosts[g].pbcVector); }
        w.u("ghostIndex", ghi);
        w.d("ghostVec", gh, {dem.ghosts.size(), 3});
        std::vector<double> dm = {dem.maxDisp, dem.nebrRange, dem.cellWidth[0], dem.cellWidth[1], dem.cellWidth[2],
                                  (double)dem.nCells[0], (double)dem.nCells[1], (double)dem.nCells[2], dem.demTime,
                                  (double)dem.demTimeStep, dem.deltat, (double)dem.multiStep, (double)dem.stdParticles,
                                  dem.sphereMat.dampCoeff, dem.sphereMat.knConst, dem.sphereMat.ksConst,
                                  dem.newNeighborList ? 1.0 : 0.0};
        w.d("dem_scalars", dm, {dm.size()});
        const energy& en = dem.particleEnergy;
        std::vector<double> ev = {en.kin, en.rotKin, en.trKin, en.grav, en.elastic, en.total};
        w.d("energy", ev, {ev.size()});
        std::vector<double> wl;
        for (size_t k = 0; k < dem.walls.size(); ++k) {
            push3(wl, dem.walls[k].p); push3(wl, dem.walls[k].n); push3(wl, dem.walls[k].FParticle); push3(wl, dem.walls[k].FHydro);
        }
        w.d("walls", wl, {dem.walls.size(), 12});
        std::vector<uint32_t> no(dem.nearObjectTable.begin(), dem.nearObjectTable.end());
        w.u("nearObjectTable", no);
        std::vector<double> ob;
        for (size_t k = 0; k < dem.objects.size(); ++k) {
            const object& o = dem.objects[k];
            ob.push_back(o.index); ob.push_back(o.r); push3(ob, o.x0); push3(ob, o.x1); push3(ob, o.FParticle);
        }
        w.d("objects", ob, {dem.objects.size(), 11});
    }
    (void)io;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s cfg workdir outprefix cycles\n", argv[0]);
        return 1;
    }
    const std::string cfg = argv[1], workdir = argv[2], prefix = argv[3];
    std::set<unsigned> cycles;
    {
        std::stringstream ss(argv[4]);
        std::string tok;
        unsigned maxc = 0;
        while (std::getline(ss, tok, ',')) { unsigned c = (unsigned)std::stoul(tok); cycles.insert(c); maxc = std::max(maxc, c); }
    }
    const unsigned last = *cycles.rbegin();

    IO io;
    DEM dem;
    LB lb;
    time_t t = time(0);
    io.now = localtime(&t);

    std::vector<std::string> args = {"hybird", "-c", cfg, "-d", workdir};
    std::vector<char*> av;
    for (auto& s : args) av.push_back(const_cast<char*>(s.c_str()));
    GetPot command_line((int)av.size(), av.data());
    parseCommandLine(io, command_line);
    GetPot lbmCfgFile(io.lbmCfgName);
    parseConfigFile(io, dem, lb, lbmCfgFile, command_line);

    io.currentTimeStep = 0;
    io.initialize();
    lb.latticeDefinition();
    dem.discreteElementInit(lb.boundary, lb.lbSize, lb.unit, lb.lbF);
    if (io.lbmSolve) lb.latticeBolzmannInit(dem.cylinders, dem.walls, dem.particles, dem.objects);
    lb.time = 0;
    dem.demTime = 0.0;
    dem.demTimeStep = 0;

    if (prefix == "-") {
        auto now = []() {
            timespec ts;
            clock_gettime(CLOCK_MONOTONIC, &ts);
            return ts.tv_sec + 1e-9 * ts.tv_nsec;
        };
        goCycle(io, dem, lb);
        const double t1 = now();
        for (unsigned c = 2; c <= last; ++c) goCycle(io, dem, lb);
        const double t2 = now();
        std::printf("{\"cycles\": %u, \"seconds_per_cycle\": %.9g, \"active_nodes\": %zu, \"particles\": %zu}\n",
                    last - 1, (t2 - t1) / double(last > 1 ? last - 1 : 1), lb.activeNodes.size(), dem.elmts.size());
        return 0;
    }
    if (cycles.count(0)) dump(prefix + "_0.bin", io, dem, lb);
    for (unsigned c = 1; c <= last; ++c) {
        goCycle(io, dem, lb);
        if (cycles.count(c)) dump(prefix + "_" + std::to_string(c) + ".bin", io, dem, lb);
    }
    return 0;
}
