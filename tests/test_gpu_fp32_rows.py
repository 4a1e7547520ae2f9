"""The FP32 path on 256-node x-rows (one workgroup per row), against FP64.

The fused kernels pull +-x neighbours within the row and apply the periodic
rewiring of nodes 1 and X - 2 (LB.cpp:437-472: node 1 pulls from X - 2, node
X - 2 from 1). Checked against the FP64 path (itself pinned to the oracle) on
256-wide lattices whose flow varies strongly along x: spheres moving through
the fluid, one of them across the periodic x plane, and the same with
bounce-back walls at the x ends. A wrong shift or wrap gives O(1) errors at the
affected nodes, so the bar is both the relative L2 of the velocity field and
the largest per-node difference relative to the largest speed (1e-4, the FP32
target), plus the element forces (1e-3).
"""
import numpy as np
import pytest

from hybird_amd._abi import config_from_cfg, element_records
from hybird_amd.cfg import base_config
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _case(periodic_x: bool):
    bx = 4 if periodic_x else 7
    cfg = base_config(lbSizeX=256, lbSizeY=12, lbSizeZ=14, initVisc=0.05, lbFX=2.0e-5, multiStep=5, density=2.0,
                      youngMod=2.0e3, poisson=0.3, restitution=0.88, frictionCoefPart=0.5, frictionCoefWall=0.5,
                      viscTang=0.2, boundary0=bx, boundary1=bx, boundary2=4, boundary3=4, boundary4=7, boundary5=7)
    x0 = 1.2 if periodic_x else 5.0  # across the periodic plane (period 254) or next to the x = 0 wall
    x = np.array([[x0, 5.5, 6.5], [60.0, 6.0, 5.0], [131.3, 5.0, 8.0], [200.0, 6.5, 6.0], [245.0, 6.0, 7.0]])
    r = np.array([2.6, 2.2, 3.0, 2.4, 2.5])
    v = np.array([[0.03, 0.0, 0.01], [-0.02, 0.01, 0.0], [0.0, -0.02, 0.02], [0.04, 0.0, -0.01],
                  [-0.03, 0.01, 0.0]])
    return cfg, element_records(x, r, v)


@pytest.mark.parametrize("periodic_x", [True, False])
def test_gpu_fp32_rows_match_fp64(periodic_x):
    from hybird_amd import runtime as rt
    cfg, rec = _case(periodic_x)
    print(f"[gpu fp32 rows] periodic_x={periodic_x}", flush=True)
    a = rt.Simulation(config_from_cfg(cfg, collision="TRT", precision="FP64"), rec)
    b = rt.Simulation(config_from_cfg(cfg, collision="TRT", precision="FP32"), rec)
    for sim in (a, b):
        sim.cycle(30)
    t = a.get("types")
    assert np.array_equal(t, b.get("types"))
    act = (t == 0) | (t == 3)
    ua, ub = a.get("u")[act], b.get("u")[act]
    assert np.max(np.abs(ua)) > 1e-3  # the spheres stir the fluid
    assert rel_l2(ub, ua) <= 1e-4, rel_l2(ub, ua)
    assert np.max(np.abs(ub - ua)) <= 1e-4 * np.max(np.abs(ua)), np.max(np.abs(ub - ua))
    # the columns next to the periodic plane, where the rewired pulls act
    X = 256
    xs = np.arange(t.size) % X
    edge = act & ((xs == 1) | (xs == X - 2))
    ue_a, ue_b = a.get("u")[edge], b.get("u")[edge]
    assert np.max(np.abs(ue_b - ue_a)) <= 1e-4 * np.max(np.abs(ua))
    fa, fb = a.get("e_forces")[:, 0], b.get("e_forces")[:, 0]
    assert rel_l2(fb, fa) <= 1e-3, rel_l2(fb, fa)
    pa, pb = a.get("e_x")[:, 0], b.get("e_x")[:, 0]
    assert np.max(np.abs(pb - pa)) <= 1e-6
    a.close()
    b.close()
