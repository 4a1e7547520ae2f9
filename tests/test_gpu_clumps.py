"""Multi-sphere elements (clumps) on the GPU against the oracle, DEM alone.

The oracle's clump code (prototypes DEM.cpp:404-433, mass and inertia with the
Steiner terms elmt.cpp:17-83, particle placement elmt.cpp:261-299, the
normal-force moments of DEM.cpp:1862-1867 and 1940-1946, the quaternion
acceleration DEM.cpp:1176-1179) is bit-identical to the compiled reference on
the clumps fixture (test_oracle_golden.py). Here a denser granular case runs
the GPU path against it over many contacts: 108 elements of one to four
spheres, random orientations, velocities and spins, in a box periodic in x and
y with floor and ceiling walls, under gravity, for 60 cycles (600 sub-steps).

Bars: pair, near-wall and ghost lists exact; element states within 1e-11
relative L2 (the short horizon keeps the chaotic growth of one-ulp
differences, from device pow in the tangential stiffness, below that).
"""
import math

import numpy as np
import pytest

from hybird_amd._abi import element_records
from hybird_amd.cfg import base_config
from helpers import fixture_config, rel_l2
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu


def clump_bed(seed=3, n_xy=6, n_z=3, r=1.4, vmax=0.05):
    """A jittered 6 x 6 x 3 grid of elements (spacing 38/6, more than the 4 r that bounds any clump of
    sphere radius r) with sizes, orientations, velocities and spins from a seeded generator."""
    rng = np.random.default_rng(seed)
    sp = 38.0 / n_xy
    x = np.array([[0.5 + sp * (i + 0.5), 0.5 + sp * (j + 0.5), 0.5 + 2.0 * r + 0.3 + sp * k]
                  for k in range(n_z) for j in range(n_xy) for i in range(n_xy)])
    x += rng.uniform(-0.25, 0.25, size=x.shape)
    n = x.shape[0]
    size = rng.integers(1, 5, size=n)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    v = rng.uniform(-vmax, vmax, size=(n, 3))
    w = rng.uniform(-0.01, 0.01, size=(n, 3))
    return element_records(x, np.full(n, r), v, size=size, w=w, q=q)


def test_gpu_clumps_dem_matches_oracle():
    from test_gpu_parity import _compare_dem
    from hybird_amd import runtime

    rec = clump_bed()
    sizes = rec[:, 1].astype(int)
    assert all(np.sum(sizes == k) >= 5 for k in (1, 2, 3, 4))
    cfg = base_config(lbSolve=0, lbSizeX=40, lbSizeY=40, lbSizeZ=40, lbFZ=-1.0e-3, youngMod=1.0e4)
    print(f"[gpu clumps] {rec.shape[0]} elements, {int(sizes.sum())} spheres", flush=True)
    orc = OracleSim(fixture_config(cfg), rec)
    gpu = runtime.Simulation(fixture_config(cfg), rec)
    contacts = 0
    for c in range(1, 61):
        orc.cycle(1)
        gpu.cycle(1)
        f = orc.get("e_forces")
        contacts += int(np.sum((np.linalg.norm(f[:, 2], axis=1) > 0) & (sizes > 1)))
        if c % 5:
            continue
        tag = f"clumps dem cycle {c}"
        _compare_dem(gpu, orc, tag)
        pd_o, pd_g = orc.get("p_data"), gpu.get("p_data")
        assert rel_l2(pd_g, pd_o) <= 1e-11, f"{tag}: particle data rel L2 {rel_l2(pd_g, pd_o):.3e}"
    assert contacts > 10, "too few clump contacts to exercise the clump forces"
    # the spheres of an element stay rigidly placed: radiusVec lengths are r |prototype|
    pd, pi = gpu.get("p_data"), gpu.get("p_idx")
    P = int(sizes.sum())
    lens = np.linalg.norm(pd[:P, 6:9], axis=1)
    r = rec[pi[:P, 1], 2]
    s = sizes[pi[:P, 1]]
    expect = {1: [0.0], 2: [0.5, 0.5], 3: [1.0, math.sqrt(3) / 2, math.sqrt(3) / 2],
              4: [1.0] + [math.sqrt(8.0 / 9.0 + 1.0 / 9.0)] * 3}
    want = np.array([expect[int(s[i])][int(pi[i, 2])] for i in range(P)]) * r
    assert np.allclose(lens, want, rtol=1e-12, atol=1e-12)
    gpu.close()
    orc.close()
