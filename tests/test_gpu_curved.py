"""Cylinders with curved walls: the reference's NET problem on the GPU.

NET (DEM.cpp:916-927, LB.cpp:552-602, 662-690) puts a static cylinder along x
through (y, z) = (4, 4) with radius 4; every node outside it is a curved wall
streamed by the Mei-Luo-Shyy rule (LB.cpp:1278-1319), the liquid fills the part
below z = 4 and z = x under a free surface, particles touch the cylinder
(DEM.cpp:1719-1744, 1984-2058). The oracle does not restate cylinders, so the
GPU is compared with the reference's own state dumps at the fixture cycles
(tests/golden/net.npz, written by the reference compiled from its sources):
node types exactly, populations, density, velocity, masses and element state
to the same tolerances as the oracle parity tests.
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


@pytest.mark.parametrize("prec", ["FP64", "FP32"])
def test_net_cylinder_matches_reference(prec):
    rt = _runtime()
    cfg, rec, z = load_case("net")
    gpu = rt.Simulation(fixture_config(cfg, precision=prec), rec)
    gpu.set_capture(True)
    print(f"[gpu curved] NET {prec}", flush=True)
    prev = 0
    for c in (int(c) for c in z["cycles"]):
        gpu.cycle(c - prev)
        prev = c
        tag = f"NET {prec} cycle {c}"
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
        for key in ("e_x", "e_w"):
            a, b = _ref(z, c, key), gpu.get(key)
            assert rel_l2(b, a) <= etol, f"{tag}: {key} rel L2 {rel_l2(b, a):.3e}"
        # FWall and MWall: plane-wall and cylinder contacts (DEM.cpp:1984-2058)
        a, b = _ref(z, c, "e_forces")[:, 4:6], gpu.get("e_forces")[:, 4:6]
        scale = np.linalg.norm(a)
        err = np.linalg.norm(b - a) / scale if scale > 0 else np.max(np.abs(b))
        assert err <= etol or np.max(np.abs(b - a)) < 1e-18, f"{tag}: FWall/MWall rel L2 {err:.3e}"
    gpu.close()
