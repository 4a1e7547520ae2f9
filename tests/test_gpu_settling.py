"""The settling-suspension workload at its full size against the oracle.

64 x 64 x 64 D3Q19 BGK, tau = 0.8, periodic in x and y, no-slip walls at z = 0
and z = 63, gravity -1e-5 lu in z; 100 monodisperse spheres of radius 4 lu,
density ratio 2.5, centres uniform and non-overlapping from seed 0; Hertz
normal contact (E = 1e5 lu, Poisson 0.3), restitution 0.88, Coulomb friction
0.5; 10 DEM sub-steps per LB step.

The reference's tangential law (the viscous one, DEM.cpp:2176-2222) runs
against the oracle (the FP64 CPU restatement, bit-identical to the compiled
reference on the golden fixtures) for 120 cycles: node types, particle flags,
solidIndex, pair, near-wall and ghost lists exact; populations, density,
velocity within 1e-12 and node forces within 1e-10 relative L2 (see
FORCE_TOL); element states, forces and energy within 1e-11. The spring law (the workload's "tangential spring with
Coulomb friction") has no reference counterpart; its run over 200
cycles must stay finite, keep the spheres inside the box and, with no contact
in the first cycles, agree with the viscous run there (both laws vanish
without contacts).
"""
import numpy as np
import pytest

from hybird_amd._abi import element_records
from hybird_amd.cfg import base_config
from hybird_amd.particles import rsa_bed
from helpers import fixture_config, rel_l2
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

N = 64
# The penalty force of a particle node is m/n (u - v_p): a difference of the fluid
# velocity and the particle's rigid-body velocity, which nearly cancel inside a
# settling sphere. The element states agree to 1e-11 (sums of forces in another
# order), and that difference reaches the node force enlarged by |u| / |u - v_p|;
# over 120 cycles it stays below 1e-10.
FORCE_TOL = 1e-10


def settling_case(model="VISCOUS"):
    tau = 0.8
    cfg = base_config(lbSizeX=N, lbSizeY=N, lbSizeZ=N, boundary0=4, boundary1=4, boundary2=4, boundary3=4,
                      boundary4=7, boundary5=7, initVisc=(tau - 0.5) / 3.0, lbFZ=-1.0e-5, density=2.5,
                      youngMod=1.0e5, poisson=0.3, restitution=0.88, frictionCoefPart=0.5, frictionCoefWall=0.5,
                      multiStep=10, tangentialModel=model)
    bed = rsa_bed(0, (0.5, 0.5, 0.5), (N - 1.5, N - 1.5, N - 1.5), lambda r: 4.0, 100,
                  periodic=(True, True, False), period=(N - 2.0, N - 2.0, 0.0), gap=0.05)
    assert bed.count == 100
    return cfg, element_records(bed.x, bed.r)


def test_settling_full_size_matches_oracle():
    from test_gpu_parity import _compare_dem, _compare_lattice
    from hybird_amd import runtime

    cfg, rec = settling_case()
    print("[gpu settling] 64^3, 100 spheres, 120 cycles against the oracle", flush=True)
    orc = OracleSim(fixture_config(cfg), rec)
    gpu = runtime.Simulation(fixture_config(cfg), rec)
    prev = 0
    for c in (1, 10, 50, 120):
        orc.cycle(c - prev)
        gpu.cycle(c - prev)
        prev = c
        tag = f"settling cycle {c}"
        _compare_lattice(gpu, orc, tag, force_tol=FORCE_TOL)
        _compare_dem(gpu, orc, tag)
    assert np.mean(gpu.get("e_x")[:, 0, 2]) < np.mean(rec[:, 5]), "the suspension did not settle"
    gpu.close()
    orc.close()


def test_settling_full_size_spring_model():
    from hybird_amd import runtime

    cfg, rec = settling_case("VISCOUS")
    cfg_s, _ = settling_case("SPRING")
    print("[gpu settling] spring law over 200 cycles", flush=True)
    a = runtime.Simulation(fixture_config(cfg), rec)
    b = runtime.Simulation(fixture_config(cfg_s), rec)
    a.cycle(1)
    b.cycle(1)
    assert np.array_equal(a.get("e_x"), b.get("e_x"))  # no contact yet: both laws exert nothing
    b.cycle(199)
    x = b.get("e_x")[:, 0]
    assert np.all(np.isfinite(x))
    assert np.all((x[:, 2] > 4.0) & (x[:, 2] < N - 5.0))
    a.close()
    b.close()
