"""HIP path (through the C ABI) against the oracle on the golden cases.

The oracle is bit-exact with the compiled reference on these cases
(test_oracle_golden.py), so these tests compare the GPU with the reference.
Bars (FP64 path): integer data (node types, particle flags, solidIndex, pair
and near-wall lists, ghosts) bit-exact; populations, density, velocity and
node forces within 1e-12 relative L2 over active nodes; element states within
1e-11 relative L2 (the GPU sums per-particle hydrodynamic forces in a
different, compensated order than the reference's node loop, and device pow
may differ by one ulp in the tangential stiffness).
"""
import numpy as np
import pytest

from helpers import fixture_config, load_case, objects_of, rel_l2
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

LB_TOL = 1e-12
ELEM_TOL = 1e-11


def _runtime():
    from hybird_amd import runtime
    return runtime


def _compare_lattice(gpu, orc, tag, var_visc=False, by_element=False, force_tol=None):
    t_o = orc.get("types")
    t_g = gpu.get("types")
    assert np.array_equal(t_o, t_g), f"{tag}: node types differ"
    act = (t_o == 0) | (t_o == 3)
    pf_o = orc.get("pflag").astype(bool) & act
    pf_g = gpu.get("pflag").astype(bool) & act
    assert np.array_equal(pf_o, pf_g), f"{tag}: particle flags differ ({np.sum(pf_o != pf_g)} nodes)"
    if pf_o.any():
        si_o = orc.get("solidIndex")[pf_o]
        si_g = gpu.get("solidIndex")[pf_o]
        if by_element:
            # a node inside several spheres of one element may name any of them (the GPU re-tests every
            # node each step, the reference keeps the sphere that first covered it): compare the element
            si_o = orc.get("p_idx")[si_o, 1]
            si_g = gpu.get("p_idx")[si_g, 1]
        assert np.array_equal(si_o, si_g), f"{tag}: solidIndex differs"
    for name in ("f", "n", "u", "hydroForce"):
        a = orc.get(name)[act]
        b = gpu.get(name)[act]
        err = rel_l2(b, a) if np.linalg.norm(a) > 0 else float(np.max(np.abs(b - a)))
        tol = force_tol if (name == "hydroForce" and force_tol is not None) else LB_TOL
        assert err <= tol, f"{tag}: {name} rel L2 {err:.3e}"
    if var_visc:
        # Bingham / Smagorinsky viscosity of the last collision (node.cpp:103-145)
        a, b = orc.get("visc")[act], gpu.get("visc")[act]
        assert rel_l2(b, a) <= LB_TOL, f"{tag}: visc rel L2 {rel_l2(b, a):.3e}"


def _compare_dem(gpu, orc, tag):
    for name in ("neighborTable", "nearWallTable", "ghostIndex"):
        a, b = orc.get(name), gpu.get(name)
        assert np.array_equal(a, b), f"{tag}: {name} differs ({a.size} vs {b.size})"
    gv_o, gv_g = orc.get("ghostVec"), gpu.get("ghostVec")
    assert np.array_equal(gv_o, gv_g), f"{tag}: ghost vectors differ"
    for name in ("e_x", "e_q", "e_w"):
        a, b = orc.get(name), gpu.get(name)
        assert rel_l2(b, a) <= ELEM_TOL, f"{tag}: {name} rel L2 {rel_l2(b, a):.3e}"
    # particleEnergy (kin, rotKin, trKin, grav, elastic, total): sums in element order
    eo, eg = orc.get("energy"), gpu.get("energy")
    assert np.all(np.abs(eg - eo) <= ELEM_TOL * np.abs(eo) + 1e-300), f"{tag}: energy {eg} vs {eo}"
    fo, fg = orc.get("e_forces"), gpu.get("e_forces")
    for k, name in enumerate(["FHydro", "MHydro", "FParticle", "MParticle", "FWall", "MWall", "FGrav"]):
        a, b = fo[:, k], fg[:, k]
        scale = max(np.linalg.norm(a), 1e-300)
        err = np.linalg.norm(b - a) / scale if np.linalg.norm(a) > 0 else np.max(np.abs(b))
        assert err <= ELEM_TOL or np.max(np.abs(b - a)) < 1e-18, f"{tag}: {name} rel L2 {err:.3e}"
    # walls[].FParticle (DEM.cpp:1936-1979) and objects[].FParticle (DEM.cpp:2100-2134)
    wo, wg = orc.get("walls"), gpu.get("walls")
    assert wo.shape == wg.shape and np.array_equal(wo[:, :6], wg[:, :6]), f"{tag}: walls differ"
    _close_vec(wg[:, 6:9], wo[:, 6:9], f"{tag}: walls FParticle")
    if gpu.get("objects").size:
        _close_vec(gpu.get("objectForces"), orc.get("objects")[:, 8:11], f"{tag}: objects FParticle")


def _close_vec(b, a, what, tol=None):
    tol = ELEM_TOL if tol is None else tol
    scale = np.linalg.norm(a)
    err = np.linalg.norm(b - a) / scale if scale > 0 else np.max(np.abs(b), initial=0.0)
    assert err <= tol or np.max(np.abs(b - a), initial=0.0) < 1e-18, f"{what}: rel L2 {err:.3e}"


# Granular dynamics amplify any one-ulp difference (device pow in the tangential
# stiffness, DEM.cpp:2187) exponentially: on dem60 the GPU stays within 1e-17 of
# the reference for 120 cycles (1200 sub-steps) and the gap then grows about 10x
# every 20 cycles. Parity is therefore asserted over that short horizon.
HORIZON = {"dem60": 100}


@pytest.mark.parametrize("name", ["lb16", "chan", "bingham", "turb", "turb_bingham", "couple2", "dense", "dem60",
                                  "walls", "objects", "mercury"])
def test_gpu_fp64_matches_oracle(name):
    rt = _runtime()
    cfg, rec, z = load_case(name)
    conf = fixture_config(cfg)
    objs = objects_of(z)
    orc = OracleSim(fixture_config(cfg), rec, objs)
    print(f"[gpu parity] case={name}", flush=True)
    gpu = rt.Simulation(conf, rec, objs)
    prev = 0
    lattice = cfg.get("lbSolve", 1) != 0
    for c in (int(c) for c in z["cycles"] if int(c) <= HORIZON.get(name, 10**9)):
        orc.cycle(c - prev)
        gpu.cycle(c - prev)
        prev = c
        tag = f"{name} cycle {c}"
        if lattice:
            _compare_lattice(gpu, orc, tag, var_visc=bool(cfg.get("nonNewtonianSolve") or cfg.get("turbulenceSolve")))
            # walls[].FHydro of the moving no-slip walls (LB.cpp:1321-1341)
            _close_vec(gpu.get("wallFHydro"), orc.get("walls")[:, 9:12], f"{tag}: walls FHydro", LB_TOL)
        if rec.shape[0]:
            _compare_dem(gpu, orc, tag)
        if objs is not None:
            assert np.array_equal(gpu.get("nearObjectTable"), orc.get("nearObjectTable")), f"{tag}: nearObjectTable"
            assert np.array_equal(gpu.get("objects"), orc.get("objects")[:, :8]), f"{tag}: objects"
    gpu.close()


def test_gpu_clumps_fp64_matches_oracle():
    """Multi-sphere elements (two-, three- and four-sphere clumps and a single
    sphere) in the fluid: contacts with the floor, between clumps through the
    periodic plane, and between a clump and a sphere. Bars as above; node
    ownership is compared per element."""
    rt = _runtime()
    cfg, rec, z = load_case("clumps")
    assert sorted(set(rec[:, 1].astype(int))) == [1, 2, 3, 4]
    print("[gpu parity] case=clumps", flush=True)
    orc = OracleSim(fixture_config(cfg), rec)
    gpu = rt.Simulation(fixture_config(cfg), rec)
    prev = 0
    for c in (int(c) for c in z["cycles"]):
        orc.cycle(c - prev)
        gpu.cycle(c - prev)
        prev = c
        tag = f"clumps cycle {c}"
        _compare_lattice(gpu, orc, tag, by_element=True)
        _compare_dem(gpu, orc, tag)
        # particles: positions, velocities and radius vectors after the corrector, element and prototype
        assert np.array_equal(gpu.get("p_idx")[:, :3], orc.get("p_idx")[:, :3]), f"{tag}: particle indices"
        pd_o, pd_g = orc.get("p_data"), gpu.get("p_data")
        assert rel_l2(pd_g, pd_o) <= ELEM_TOL, f"{tag}: particle data rel L2 {rel_l2(pd_g, pd_o):.3e}"
        assert np.array_equal(gpu.get("e_misc")[:, 18], orc.get("e_misc")[:, 18])
        assert rel_l2(gpu.get("e_misc")[:, :5], orc.get("e_misc")[:, :5]) <= 1e-15, f"{tag}: mass, inertia"
    gpu.close()


def _compare_mass(gpu, orc, tag, act):
    a, b = orc.get("mass")[act], gpu.get("mass")[act]
    assert rel_l2(b, a) <= LB_TOL, f"{tag}: mass rel L2 {rel_l2(b, a):.3e}"
    fo, fg = float(orc.get("fluidMass")[0]), float(gpu.get("fluidMass")[0])
    assert abs(fg - fo) <= LB_TOL * abs(fo), f"{tag}: fluid mass {fg!r} vs {fo!r}"


@pytest.mark.parametrize("name", ["fs", "walls_fs"])
def test_gpu_free_surface_fp64_matches_oracle(name):
    """SPLASH free surface: mass exchange, filled/emptied nodes, new interface
    nodes and isolated-node removal (fs: periodic/no-slip box, 60 cycles;
    walls_fs: slip, moving slip and moving no-slip faces, 40 cycles)."""
    rt = _runtime()
    cfg, rec, z = load_case(name)
    print(f"[gpu parity] case={name}", flush=True)
    orc = OracleSim(fixture_config(cfg), rec)
    gpu = rt.Simulation(fixture_config(cfg), rec)
    gpu.set_capture(True)
    prev = 0
    for c in (int(c) for c in z["cycles"]):
        orc.cycle(c - prev)
        gpu.cycle(c - prev)
        prev = c
        tag = f"{name} cycle {c}"
        _compare_lattice(gpu, orc, tag)
        t = orc.get("types")
        _compare_mass(gpu, orc, tag, (t == 0) | (t == 3))
    gpu.close()


def _liquid_box_case():
    """A small dam break: a Bingham liquid block with three spheres collapsing
    under gravity into a gas-filled box (walls in x and z, periodic in y)."""
    from hybird_amd.cfg import base_config
    from hybird_amd._abi import element_records
    cfg = base_config(problemName="LIQUID_BOX", freeSurfaceSolve=1, lbSizeX=30, lbSizeY=12, lbSizeZ=22,
                      boundary0=7, boundary1=7, boundary2=4, boundary3=4, boundary4=7, boundary5=7,
                      nonNewtonianSolve=1, plasticVisc=0.05 / 3.0, yieldStress=1.0e-5, initVisc=0.05,
                      lbFX=2.0e-4, lbFZ=-1.0e-3, density=2.0, youngMod=2.0e3)
    rec = element_records(np.array([[5.0, 6.0, 5.0], [9.5, 5.5, 8.0], [4.5, 6.5, 12.5]]), np.array([2.5, 2.2, 2.0]))
    return cfg, rec, (0.0, 13.0, 0.0, 11.0, 0.0, 16.0)


def test_gpu_free_surface_particles_fp64_matches_oracle():
    rt = _runtime()
    cfg, rec, box = _liquid_box_case()
    print("[gpu parity] case=liquid box with particles", flush=True)
    orc = OracleSim(fixture_config(cfg, liquid_box=box), rec)
    gpu = rt.Simulation(fixture_config(cfg, liquid_box=box), rec)
    gpu.set_capture(True)
    prev = 0
    changed = False
    t0 = orc.get("types")
    for c in (1, 5, 20, 40):
        orc.cycle(c - prev)
        gpu.cycle(c - prev)
        prev = c
        tag = f"liquid box cycle {c}"
        _compare_lattice(gpu, orc, tag)
        t = orc.get("types")
        changed |= not np.array_equal(t, t0)
        _compare_mass(gpu, orc, tag, (t == 0) | (t == 3))
        _compare_dem(gpu, orc, tag)
    assert changed, "the interface never moved: the case does not exercise the reclassification"
    gpu.close()


def test_gpu_rejects_gas_without_free_surface():
    rt = _runtime()
    cfg, rec, _ = load_case("fs")
    cfg = dict(cfg, freeSurfaceSolve=0)
    with pytest.raises(rt.HybError):
        rt.Simulation(fixture_config(cfg), rec)
