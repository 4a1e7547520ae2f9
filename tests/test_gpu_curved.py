"""Cylinders with curved walls: the reference's NET and DRUM problems on the GPU.

NET (DEM.cpp:916-927, LB.cpp:552-602, 662-690) puts a static cylinder along x
through (y, z) = (4, 4) with radius 4; every node outside it is a curved wall
streamed by the Mei-Luo-Shyy rule (LB.cpp:1278-1319), the liquid fills the part
below z = 4 and z = x under a free surface, particles touch the cylinder
(DEM.cpp:1719-1744, 1984-2058). The oracle does not restate cylinders, so the
GPU is compared with the reference's own state dumps at the fixture cycles
(tests/golden/net.npz, written by the reference compiled from its sources):
node types exactly, populations, density, velocity, masses and element state
to the same tolerances as the oracle parity tests.

DRUM (DEM.cpp:783-887, LB.cpp:206-208, 734-747, 946-993, 1806-1822) turns the
cylinder (R = 1.243 around (0.95, 1.26), along y) at drumSpeed, so its curved wall
nodes carry the drum's surface speed (the BBi term of the Mei-Luo-Shyy rule) and the
particle contacts see it; the pool is the circle of the fluidMass formula and every
free-surface pass ends with the global mass fix towards fluidMass (tests/golden/drum.npz).
"""
import numpy as np
import pytest

from helpers import fixture_config, load_case, rel_l2

pytestmark = pytest.mark.gpu

LB_TOL = 1e-12
ELEM_TOL = 1e-11


def _runtime():
    from hybird_amd import runtime
    return runtime


def _ref(z, c, key):
    return z[f"c{c}/{key}"]


def _gear_err(a, b, dt, first):
    """Largest error over the Gear slots of an element array [E][6][3], each slot k
    scaled by dt^(k - first) onto the scale of the slot the forces set (first = 2
    for positions, 1 for rotations), relative to that slot's norm. The unscaled
    higher derivatives carry 1/dt^(k - first) times any rounding difference of the
    force (a 1-ulp torque difference reaches 1e-2 in w5 at dt = 1e-3)."""
    ref = max(np.linalg.norm(a[:, first]), np.linalg.norm(a[:, 0]) if first == 1 else 0.0, 1e-300)
    err = 0.0
    for k in range(a.shape[1]):
        s = dt ** max(k - first, 0)
        scale = max(np.linalg.norm(a[:, k] * s), ref) if k >= first else max(np.linalg.norm(a[:, k]), 1e-300)
        err = max(err, np.linalg.norm((b[:, k] - a[:, k]) * s) / scale)
    return err


@pytest.mark.parametrize("case", ["net", "drum"])
@pytest.mark.parametrize("prec", ["FP64", "FP32"])
def test_cylinder_problem_matches_reference(case, prec):
    rt = _runtime()
    cfg, rec, z = load_case(case)
    gpu = rt.Simulation(fixture_config(cfg, precision=prec), rec)
    gpu.set_capture(True)
    print(f"[gpu curved] {case} {prec}", flush=True)
    prev = 0
    for c in (int(c) for c in z["cycles"]):
        gpu.cycle(c - prev)
        prev = c
        tag = f"{case} {prec} cycle {c}"
        t_ref = _ref(z, c, "types").ravel()
        assert np.array_equal(gpu.get("types").ravel(), t_ref), f"{tag}: types differ"
        act = (t_ref == 0) | (t_ref == 3)
        tol = LB_TOL if prec == "FP64" else 2e-4
        for key in ("n", "u"):
            a = _ref(z, c, key).reshape(gpu.get(key).shape)[act]
            b = gpu.get(key)[act]
            assert rel_l2(b, a) <= tol, f"{tag}: {key} rel L2 {rel_l2(b, a):.3e}"
        if f"c{c}/f" in z.files and prec == "FP64":
            a = _ref(z, c, "f").reshape(-1, 19)[act]
            b = gpu.get("f")[act]
            assert rel_l2(b, a) <= LB_TOL, f"{tag}: f rel L2 {rel_l2(b, a):.3e}"
        itf = t_ref == 3
        if itf.any():
            a = _ref(z, c, "mass").ravel()[itf]
            b = gpu.get("mass")[itf]
            assert rel_l2(b, a) <= tol, f"{tag}: interface mass rel L2 {rel_l2(b, a):.3e}"
        etol = ELEM_TOL if prec == "FP64" else 1e-3
        dt = cfg["unitTime"] / cfg["multiStep"]
        for key, first in (("e_x", 2), ("e_w", 1)):
            a, b = _ref(z, c, key), gpu.get(key)
            err = _gear_err(a, b, dt, first)
            assert err <= etol, f"{tag}: {key} Gear-scaled rel error {err:.3e}"
        # FWall and MWall: plane-wall and cylinder contacts (DEM.cpp:1984-2058)
        a, b = _ref(z, c, "e_forces")[:, 4:6], gpu.get("e_forces")[:, 4:6]
        scale = np.linalg.norm(a)
        err = np.linalg.norm(b - a) / scale if scale > 0 else np.max(np.abs(b))
        assert err <= etol or np.max(np.abs(b - a)) < 1e-18, f"{tag}: FWall/MWall rel L2 {err:.3e}"
    gpu.close()
