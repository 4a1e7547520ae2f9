"""Contacts with every wall in range (HYB_WALL_CONTACTS_ALL) on the HIP path.

FP64 GPU against the oracle with the same switch (the oracle's switch is pinned
by closed forms in test_wall_contacts.py): a closed, particle-laden box with
spheres in its corners and along its edges, coupled to the fluid, over 40
cycles; and the DEM-only corner sphere resting at the Hertz height.
"""
import numpy as np
import pytest

from hybird_amd._abi import config_from_cfg, element_records
from hybird_amd.cfg import base_config
from oracle.oracle import OracleSim
from test_gpu_parity import _compare_dem, _compare_lattice
from test_wall_contacts import G, _corner_setup, _resting_height

pytestmark = pytest.mark.gpu


def _runtime():
    from hybird_amd import runtime
    return runtime


def _box_case():
    cfg = base_config(lbSizeX=16, lbSizeY=14, lbSizeZ=18, initVisc=0.05, lbFZ=-4.0e-4, lbFX=1.0e-4, multiStep=10,
                      density=2.5, youngMod=2.0e3, poisson=0.3, restitution=0.88, frictionCoefPart=0.5,
                      frictionCoefWall=0.5, viscTang=0.2, boundary0=7, boundary1=7, boundary2=7, boundary3=7,
                      boundary4=7, boundary5=7, wallContacts="ALL")
    # wall planes at 0.5 and N - 1.5; each sphere starts 0.05-0.1 away from the walls it is placed at
    x = np.array([[2.72, 2.72, 2.75],   # three walls (x0, y0, floor)
                  [12.27, 2.73, 2.85],  # x1, y0, floor
                  [2.8, 10.2, 8.0],     # x0, y1 edge
                  [8.0, 6.5, 2.75],     # floor
                  [8.0, 6.8, 8.5],      # free, moving down
                  [12.3, 10.3, 14.3]])  # top corner (x1, y1, ceiling)
    r = np.array([2.2, 2.3, 2.2, 2.2, 2.0, 2.1])
    v = np.zeros_like(x)
    v[0] = [-0.01, -0.01, -0.004]  # into its corner
    v[1] = [0.01, -0.01, 0.0]
    v[4] = [0.0, 0.0, -0.02]
    return cfg, element_records(x, r, v)


def test_gpu_all_walls_fp64_matches_oracle():
    rt = _runtime()
    cfg, rec = _box_case()
    gpu = rt.Simulation(config_from_cfg(cfg), rec)
    orc = OracleSim(config_from_cfg(cfg), rec)
    assert gpu.get("nearWallTable").size > 2 * 6  # several particles near more than one wall
    step = 0
    for n in (10, 10, 20):
        step += n
        print(f"[gpu all walls] cycle {step}", flush=True)
        gpu.cycle(n)
        orc.cycle(n)
        _compare_lattice(gpu, orc, f"all walls cycle {step}")
        _compare_dem(gpu, orc, f"all walls cycle {step}")
    gpu.close()


def test_gpu_all_walls_corner_sphere_rests():
    rt = _runtime()
    gpu = rt.Simulation(*_corner_setup("ALL"))
    print("[gpu all walls] corner sphere, 600 cycles", flush=True)
    gpu.cycle(600)
    z, m = _resting_height()
    x = gpu.get("e_x")[0]
    assert abs(x[0, 2] - z) <= 1e-12 * z
    floor = gpu.get("walls")[4, 6:9]
    assert abs(floor[2] + m * abs(G)) <= 1e-12 * m * abs(G)
    gpu.close()
