"""The granular-collapse workload at reduced size: TRT collision, a free surface
and a lognormal particle column that starts partly in the gas.

The full workload is 512x512x256 D3Q19 TRT tau 0.6 with the free surface at
z = 160 and a lognormal sphere column (mean radius 4 lu, sigma 0.2) at x, y <
160, z < 220, density ratio 2.6, friction 0.6, in a closed tank. Here the same
setup is scaled to a 48x48x40 tank (liquid below z = 24, column x, y < 16.5,
z < 34.5, mean radius 2) so that the oracle (FP64 CPU restatement) runs it in a
few seconds: TRT together with the free-surface pass (LB.cpp:1492-1822), and
particles falling from the gas through the interface into the liquid.

Bars: FP64 GPU against the oracle as in the other parity tests (node types,
flags and pair lists exact; populations, density, velocity, node forces and
masses within 1e-12 relative L2; element states within 1e-11); FP32 against
FP64 as in the other free-surface FP32 test (liquid mass within 1e-6, at most
1% of the node types different, velocity within 1e-3 where the types agree).
"""
import math

import numpy as np
import pytest

from hybird_amd._abi import element_records
from hybird_amd.cfg import base_config
from hybird_amd.particles import rsa_bed
from helpers import fixture_config, rel_l2
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

CYCLES = (1, 5, 15, 30)


def collapse_case(size=(48, 48, 40), surface=24, column=(16.5, 16.5, 34.5), mean_r=2.0, sigma=0.2, count=40,
                  seed=4):
    """Closed no-slip tank, liquid below z = surface, RSA column of lognormal spheres."""
    tau = 0.6
    X, Y, Z = size
    cfg = base_config(problemName="LIQUID_BOX", freeSurfaceSolve=1, lbSizeX=X, lbSizeY=Y, lbSizeZ=Z,
                      boundary0=7, boundary1=7, boundary2=7, boundary3=7, boundary4=7, boundary5=7,
                      initVisc=(tau - 0.5) / 3.0, lbFZ=-5.0e-4, density=2.6, youngMod=2.0e3,
                      frictionCoefPart=0.6, frictionCoefWall=0.6)
    mu = math.log(mean_r) - 0.5 * sigma * sigma  # lognormal with mean radius mean_r
    bed = rsa_bed(seed, (0.5, 0.5, 0.5), column, lambda r: math.exp(mu + sigma * r.normal()), count, gap=0.05)
    box = (1.0, X - 2.0, 1.0, Y - 2.0, 1.0, surface - 1.0)
    return cfg, element_records(bed.x, bed.r), box


def test_gpu_collapse_trt_free_surface_fp64_matches_oracle():
    from test_gpu_parity import _compare_dem, _compare_lattice, _compare_mass
    from hybird_amd import runtime

    cfg, rec, box = collapse_case()
    print(f"[gpu collapse] fp64 vs oracle, {rec.shape[0]} spheres", flush=True)
    conf = fixture_config(cfg, collision="TRT", liquid_box=box)
    orc = OracleSim(fixture_config(cfg, collision="TRT", liquid_box=box), rec)
    gpu = runtime.Simulation(conf, rec)
    gpu.set_capture(True)
    z0 = rec[:, 5].copy()
    in_gas = int(np.sum(z0 - rec[:, 2] > 23.5))
    assert in_gas > 0, "no sphere starts in the gas"
    prev = 0
    for c in CYCLES:
        orc.cycle(c - prev)
        gpu.cycle(c - prev)
        prev = c
        tag = f"collapse cycle {c}"
        _compare_lattice(gpu, orc, tag)
        t = orc.get("types")
        _compare_mass(gpu, orc, tag, (t == 0) | (t == 3))
        _compare_dem(gpu, orc, tag)
    # the column moved (gravity) and the interface moved
    assert np.max(np.abs(gpu.get("e_x")[:, 0, 2] - z0)) > 1e-3
    gpu.close()
    orc.close()


def test_gpu_collapse_fp32_against_fp64():
    from hybird_amd import runtime

    cfg, rec, box = collapse_case()
    print("[gpu collapse] fp32 vs fp64", flush=True)
    a = runtime.Simulation(fixture_config(cfg, collision="TRT", precision="FP64", liquid_box=box), rec)
    b = runtime.Simulation(fixture_config(cfg, collision="TRT", precision="FP32", liquid_box=box), rec)
    a.set_capture(True)
    b.set_capture(True)
    a.cycle(CYCLES[-1])
    b.cycle(CYCLES[-1])
    ma, mb = float(a.get("fluidMass")[0]), float(b.get("fluidMass")[0])
    assert abs(mb - ma) <= 1e-6 * abs(ma), (ma, mb)
    ta, tb = a.get("types"), b.get("types")
    assert int(np.sum(ta != tb)) <= 0.01 * int(np.sum((ta == 3) | (ta == 0)))
    act = ((ta == 0) | (ta == 3)) & (ta == tb)
    assert rel_l2(b.get("u")[act], a.get("u")[act]) <= 1e-3
    xa, xb = a.get("e_x")[:, 0], b.get("e_x")[:, 0]
    assert np.max(np.abs(xb - xa)) <= 1e-3  # lattice units, over 30 cycles
    a.close()
    b.close()
