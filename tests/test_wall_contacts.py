"""Wall contacts with every wall in range (HYB_WALL_CONTACTS_ALL), on the oracle.

The reference lists only the first wall within nebrRange of each particle
(DEM.cpp:1496-1513, `break`), so a sphere in the corner of a side wall and the
floor gets no floor contact and falls through the floor. With every wall in
range a resting sphere sits where the Hertz wall force (DEM.cpp:1902-1982, with
the reference's doubled overlap) carries its weight; that position and the
floor's FParticle are closed forms.
"""
import math

import numpy as np
import pytest

from hybird_amd._abi import config_from_cfg, element_records
from hybird_amd.cfg import base_config
from oracle.oracle import OracleSim

G = -1.0e-4
E_MOD, NU, RHO, R = 1.0e3, 0.3, 2.0, 2.0


def _corner_setup(mode):
    # DEM only, closed box of no-slip walls (planes at 0.5 and N - 1.5), 20 sub-steps per cycle: omega dt of
    # the resting contact is about 0.07, well inside the Gear integrator's stable range
    cfg = base_config(lbSolve=0, lbSizeX=20, lbSizeY=20, lbSizeZ=20, lbFZ=G, multiStep=20, density=RHO,
                      boundary0=7, boundary1=7, boundary2=7, boundary3=7, boundary4=7, boundary5=7,
                      youngMod=E_MOD, poisson=NU, restitution=0.5, frictionCoefWall=0.5, viscTang=0.2,
                      wallContacts=mode)
    # 0.2 from the x = 0.5 wall (inside nebrRange of it and of the floor), 0.1 above the floor
    x0 = np.array([[0.5 + R + 0.2, 10.0, 0.5 + R + 0.1]])
    return config_from_cfg(cfg), element_records(x0, np.array([R]))


def _corner(mode):
    return OracleSim(*_corner_setup(mode))


def _resting_height():
    m = RHO * 4.0 / 3.0 * math.pi * R ** 3
    kn = 2.0 / 3.0 * E_MOD / (1.0 - NU * NU)  # knConst (DEM.cpp:21)
    # fn = kn sqrt(R) (2 ov)^1.5 = m |g| (normalContact with effRad = R, overlap 2 ov for a wall)
    ov = 0.5 * (m * abs(G) / (kn * math.sqrt(R))) ** (2.0 / 3.0)
    return 0.5 + R - ov, m


def test_corner_sphere_falls_through_the_floor_with_the_first_wall_only():
    sim = _corner("FIRST")
    assert sim.get("nearWallTable").tolist() == [0, 0]  # the x = 0.5 wall only
    with pytest.raises(RuntimeError, match="outside the neighbour-cell grid"):
        sim.cycle(400)


def test_corner_sphere_rests_on_the_floor_with_every_wall():
    sim = _corner("ALL")
    assert sim.get("nearWallTable").tolist() == [0, 0, 0, 4]  # x = 0.5 wall, then the floor
    sim.cycle(600)
    z, m = _resting_height()
    x = sim.get("e_x")[0]
    assert abs(x[0, 2] - z) <= 1e-12 * z
    assert abs(x[1, 2]) <= 1e-12
    walls = sim.get("walls")
    floor = walls[4, 6:9]
    assert abs(floor[2] + m * abs(G)) <= 1e-12 * m * abs(G)  # FParticle -= fn (DEM.cpp:1936-1979)
    assert np.all(walls[[0, 1, 2, 3, 5], 6:9] == 0.0)  # the side wall is in range but not touched


def test_all_walls_list_is_every_wall_in_range():
    # 24 spheres spread over a closed box, some near edges and corners
    rng = np.random.default_rng(7)
    x = rng.uniform(2.6, 13.4, size=(24, 3))
    x[:6, :2] = rng.uniform(2.4, 3.0, size=(6, 2))  # near the x and y walls
    r = np.full(24, 1.5)
    cfg = base_config(lbSolve=0, lbSizeX=16, lbSizeY=16, lbSizeZ=16, lbFZ=G, multiStep=5, density=RHO,
                      boundary0=7, boundary1=7, boundary2=7, boundary3=7, boundary4=7, boundary5=7,
                      youngMod=E_MOD, poisson=NU)
    lists = {}
    for mode in ("FIRST", "ALL"):
        sim = OracleSim(config_from_cfg(dict(cfg, wallContacts=mode)), element_records(x, r))
        lists[mode] = sim.get("nearWallTable").reshape(-1, 2)
        walls = sim.get("walls")
        nebr = sim.get("dem_scalars")[1]
        pos = sim.get("p_data")[:, :3]
    dist = (pos[:, None, :] - walls[None, :, 0:3]) * walls[None, :, 3:6]
    near = dist.sum(axis=2) < nebr  # Wall::dist = n . (x - p)
    expect = np.argwhere(near)  # particle-major, walls ascending
    assert np.array_equal(lists["ALL"], expect)
    first = [(p, np.nonzero(near[p])[0][0]) for p in range(near.shape[0]) if near[p].any()]
    assert np.array_equal(lists["FIRST"], np.array(first).reshape(-1, 2))
    assert len(expect) > len(first)  # some particles are near two walls
