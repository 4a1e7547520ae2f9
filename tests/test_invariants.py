"""Checks that pin the oracle where the reference's fixtures leave it open:
closed forms and invariants that a dropped term, a wrong sign or a transposed
index would break.

* D3Q19 equilibrium moments (tables and weights, node.cpp:26-61, lattice.h:36-101)
* TRT with magic 1/4 at tau = 1 equals BGK (both rates are 1)
* Poiseuille flow between half-way bounce-back walls: TRT with magic 3/16 and Guo
  forcing reproduces the parabola u = F/(2 nu) (z - z0)(z1 - z) exactly
* mass conservation of stream + collide with walls and periodic planes
* momentum balance of the penalty coupling: node forces and element forces cancel
* Gear predictor-corrector integrates a constant acceleration exactly (after its start-up step)
* every listed contact pair is within the neighbour range (brute force)
"""
import numpy as np
import pytest

from helpers import fixture_config, load_case, rel_l2
from hybird_amd._abi import config_from_cfg, element_records
from hybird_amd.cfg import base_config
from oracle.oracle import OracleSim

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1])
C = np.stack([CX, CY, CZ]).astype(np.float64)
EMPTY = np.zeros((0, 20))


def _active(sim):
    t = sim.get("types")
    return (t == 0) | (t == 3)


def test_equilibrium_moments():
    cfg = base_config(lbSizeX=8, lbSizeY=6, lbSizeZ=10, initVisc=0.1, lbFX=3.0e-4, lbFZ=-2.0e-4,
                      initVelocityX=0.04, initVelocityY=-0.03, initVelocityZ=0.02)
    sim = OracleSim(config_from_cfg(cfg), EMPTY)
    act = _active(sim)
    f, n, u = sim.get("f")[act], sim.get("n")[act], sim.get("u")[act]
    assert np.allclose(f.sum(1), n, rtol=1e-14, atol=0)
    assert np.allclose(f @ C.T, n[:, None] * u, rtol=1e-13, atol=1e-17)
    second = np.einsum("na,ka,la->nkl", f, C, C)
    expect = n[:, None, None] / 3.0 * np.eye(3) + n[:, None, None] * u[:, :, None] * u[:, None, :]
    assert np.allclose(second, expect, rtol=1e-13, atol=1e-16)


def test_trt_magic_quarter_equals_bgk_at_tau_one():
    cfg = base_config(lbSizeX=10, lbSizeY=8, lbSizeZ=12, initVisc=1.0 / 6.0, lbFX=2.0e-4, lbFZ=-1.0e-4,
                      initVelocityX=0.03)
    bgk = OracleSim(config_from_cfg(cfg, collision="BGK"), EMPTY)
    trt = OracleSim(config_from_cfg(cfg, collision="TRT", trt_magic=0.25), EMPTY)
    bgk.cycle(12)
    trt.cycle(12)
    act = _active(bgk)
    assert rel_l2(trt.get("f")[act], bgk.get("f")[act]) <= 1e-14
    # a different magic parameter must change the result (the test sees the odd rate)
    trt2 = OracleSim(config_from_cfg(cfg, collision="TRT", trt_magic=0.1), EMPTY)
    trt2.cycle(12)
    assert rel_l2(trt2.get("f")[act], bgk.get("f")[act]) > 1e-8


def test_poiseuille_trt_three_sixteenths():
    Z, nu, F = 18, 0.1, 1.0e-6
    cfg = base_config(lbSizeX=4, lbSizeY=4, lbSizeZ=Z, initVisc=nu, lbFX=F)
    sim = OracleSim(config_from_cfg(cfg, collision="TRT", trt_magic=3.0 / 16.0), EMPTY)
    sim.cycle(14000)
    act = _active(sim)
    z = np.arange(4 * 4 * Z) // 16
    # momentum density n u (the force is a force density; n settles slightly below 1)
    j = sim.get("n") * sim.get("u")[:, 0]
    z0, z1 = 0.5, Z - 1.5  # half-way between the wall layers and the first fluid nodes
    exact = F / (2.0 * nu) * (z - z0) * (z1 - z)
    assert np.max(np.abs(j[act] - exact[act])) <= 1e-10 * np.max(exact)
    # BGK with the same viscosity misses the parabola (slip error grows with tau)
    bgk = OracleSim(config_from_cfg(cfg, collision="BGK"), EMPTY)
    bgk.cycle(14000)
    jb = bgk.get("n") * bgk.get("u")[:, 0]
    assert np.max(np.abs(jb[act] - exact[act])) > 1e-4 * np.max(exact)


def test_mass_conservation_stream_collide():
    cfg = base_config(lbSizeX=10, lbSizeY=9, lbSizeZ=12, initVisc=0.05, lbFX=1.0e-4, lbFZ=-3.0e-4,
                      boundary2=7, boundary3=7, initVelocityY=0.02)
    sim = OracleSim(config_from_cfg(cfg, collision="TRT"), EMPTY)
    act = _active(sim)
    m0 = sim.get("f")[act].sum()
    sim.cycle(40)
    m1 = sim.get("f")[act].sum()
    assert abs(m1 - m0) <= 1e-13 * m0


def test_penalty_momentum_balance():
    cfg, rec, _ = load_case("couple2")
    sim = OracleSim(fixture_config(cfg), rec)
    sim.cycle(5)
    act = _active(sim)
    node = np.nansum(sim.get("hydroForce")[act], axis=0)
    elem = sim.get("e_forces")[:, 0].sum(axis=0)  # FHydro, lattice units (unit.Force = 1)
    assert np.linalg.norm(elem) > 0
    assert np.linalg.norm(node + elem) <= 1e-12 * np.linalg.norm(elem)


def test_gear_integrates_free_fall_exactly():
    g = -2.0e-3
    cfg = base_config(lbSolve=0, lbSizeX=40, lbSizeY=40, lbSizeZ=40, lbFZ=g, multiStep=5, density=2.0)
    x0, v0 = np.array([[20.0, 19.0, 30.0]]), np.array([[0.01, -0.02, 0.03]])
    sim = OracleSim(config_from_cfg(cfg), element_records(x0, np.array([2.0]), v0))
    # The reference starts the Gear history with zero acceleration, so the first
    # sub-step leaves a constant offset in position and velocity; from then on the
    # trajectory is an exact parabola: its second differences over equal
    # intervals equal g h^2, with no drift.
    xs = []
    for _ in range(4):
        sim.cycle(10)
        xs.append(sim.get("e_x")[0, 0].copy())
    d2a = xs[2] - 2 * xs[1] + xs[0]
    d2b = xs[3] - 2 * xs[2] + xs[1]
    expect = np.array([0.0, 0.0, g * 10.0 * 10.0])
    assert np.max(np.abs(d2a - expect)) <= 1e-12
    assert np.max(np.abs(d2b - expect)) <= 1e-12


def test_pairs_within_range_brute_force():
    cfg, rec, _ = load_case("dem60")
    sim = OracleSim(fixture_config(cfg), rec)
    sim.cycle(10)
    pairs = sim.get("neighborTable").reshape(-1, 2)
    pd, pi = sim.get("p_data"), sim.get("p_idx")
    sc = sim.get("dem_scalars")
    max_disp, nebr = sc[0], sc[1]
    pos = pd[:, :3]
    # pairs are within nebrRange when listed; since then each particle moved at most maxDisp
    d = np.linalg.norm(pos[pairs[:, 0]] - pos[pairs[:, 1]], axis=1)
    assert np.all(d < nebr + 2.0 * max_disp)
    assert np.all(pi[pairs[:, 0], 1] != pi[pairs[:, 1], 1])  # different clusters
    # every touching pair of real particles is listed (directly or through a ghost)
    P = rec.shape[0]
    listed = {tuple(sorted((int(pi[a, 1]), int(pi[b, 1])))) for a, b in pairs}
    r = pd[:, 9]
    for a in range(P):
        for b in range(a + 1, P):
            if np.linalg.norm(pos[a] - pos[b]) < r[a] + r[b]:
                assert (a, b) in listed
