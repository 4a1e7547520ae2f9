"""Tangential spring contact law (HYB_TANGENTIAL_SPRING), pinned by closed forms.

The reference's tangential law is history-free (DEM.cpp:2176-2222): with
viscTang = 0 it exerts no tangential force at all. The spring law keeps a
tangential spring xi per contact with the reference's stiffness
k_s = ksConst sqrt(R) F_n^(1/3) (DEM.cpp:2187), limited by Coulomb friction.
Its behaviour has closed forms that a wrong sign, a missing torque or a lost
spring would break:

* stick: a sphere resting on a wall (normal force m g at the static Hertz
  overlap) given a small tangential speed v0 oscillates about its rest point;
  the slip speed at the contact point is v0 cos(w t) with
  w^2 = k_s (1/m + r^2/I) = 3.5 k_s / m for a solid sphere;
* slip: with a large v0 the spring saturates at mu F_n and the slip speed
  falls at the constant rate 3.5 mu g (the classic sliding-then-rolling ball);
* an oblique collision of two free spheres (restitution 1, no tangential
  damping) keeps the total momentum and, up to the overlap times the
  tangential force (the torques take the lever arm r, not the distance to the
  contact point), the total angular momentum, and leaves the spheres
  spinning; the kinetic energy is not kept exactly, since the spring
  stiffness falls with the normal force while the spring is still stretched
  (the stored 0.5 k_s xi^2 is let go as k_s -> 0 at separation), but the
  loss stays small;
* the spring survives a neighbour rebuild: the same collision, started so that
  a rebuild falls inside the contact, ends in the same state (up to rounding)
  as when it does not (a reset spring would lose its stored force);
* with every wall in range as a contact partner (wallContacts = ALL, one spring
  per particle and wall): a sphere touching one wall moves exactly as with the
  first wall only, and a sphere resting in the corner of the floor and a side
  wall under equal normal loads, given a speed along the corner line, sticks
  on both walls: the two springs xi_f, xi_s obey
  m xi_f'' = -k ((1 + m r^2 / I) xi_f + xi_s) and the mirror equation, so the
  symmetric mode it starts in has w^2 = 4.5 k_s / m (3.5 with one contact).

The DEM runs alone here (lbSolve = 0).
"""
import math

import numpy as np
import pytest

from hybird_amd._abi import element_records
from hybird_amd.cfg import base_config
from helpers import fixture_config

pytestmark = pytest.mark.gpu

E = 1.0          # Young's modulus (lattice units): slow contacts, well resolved by 10 sub-steps per cycle
NU = 0.3
RHO = 2.5
R = 4.0
G = 1.0e-4


def _sim(cfg, x, r, v=None, w=None, model="SPRING"):
    from hybird_amd import runtime

    rec = element_records(x, r, v)
    if w is not None:
        rec[:, 9:12] = np.asarray(w, dtype=np.float64).reshape(-1, 3)
    cfg = dict(cfg, tangentialModel=model)
    return runtime.Simulation(fixture_config(cfg), rec)


def _floor_case(v0, model="SPRING", mu=0.5, walls="FIRST"):
    """One sphere at the static overlap on the floor (wall at z = 0.5) with tangential speed v0 along x."""
    cfg = base_config(lbSolve=0, lbSizeX=24, lbSizeY=24, lbSizeZ=20, lbFZ=-G, density=RHO, youngMod=E, poisson=NU,
                      restitution=0.88, viscTang=0.0, frictionCoefWall=mu, frictionCoefPart=mu, multiStep=10,
                      wallContacts=walls)
    m = 4.0 / 3.0 * math.pi * R ** 3 * RHO
    kn = 2.0 / 3.0 * E / (1.0 - NU * NU)
    # wall contact: F_n = kn sqrt(r) (2 delta)^(3/2) = m g
    two_delta = (m * G / (kn * math.sqrt(R))) ** (2.0 / 3.0)
    z0 = 0.5 + R - 0.5 * two_delta
    sim = _sim(cfg, [[12.0, 12.0, z0]], [R], [[v0, 0.0, 0.0]], model=model)
    ks = 2.0 * E / ((2.0 - NU) * (1.0 + NU)) * math.sqrt(R) * (m * G) ** (1.0 / 3.0)
    return sim, m, ks


def _slip(sim):
    """Slip speed along x of the contact point at the bottom of the sphere: v_x - r w_y."""
    x = sim.get("e_x")[0]
    w = sim.get("e_w")[0]
    return x[1, 0] - R * w[0, 1]


def _zero_crossings(t, v):
    out = []
    for k in range(1, len(v)):
        if v[k - 1] > 0.0 >= v[k] or v[k - 1] < 0.0 <= v[k]:
            out.append(t[k - 1] + (t[k] - t[k - 1]) * v[k - 1] / (v[k - 1] - v[k]))
    return out


def test_spring_stick_oscillation_period():
    v0 = 1.0e-4
    sim, m, ks = _floor_case(v0)
    print("[gpu spring] stick oscillation on the floor", flush=True)
    t, vs = [0.0], [v0]
    for k in range(1, 141):
        sim.cycle(1)
        t.append(float(k))
        vs.append(_slip(sim))
    sim.close()
    z = _zero_crossings(np.array(t), np.array(vs))
    assert len(z) >= 3, z
    w = math.sqrt(3.5 * ks / m)
    half = (z[2] - z[0]) / 2.0
    assert abs(half - math.pi / w) <= 5e-3 * math.pi / w, (half, math.pi / w)
    assert abs(z[0] - 0.5 * math.pi / w) <= 1e-2 * math.pi / w, (z[0], 0.5 * math.pi / w)
    # amplitude kept (no damping): the extremes stay at v0
    assert max(abs(v) for v in vs[30:]) == pytest.approx(v0, rel=2e-2)


def test_viscous_law_exerts_no_force_without_tangential_damping():
    v0 = 1.0e-4
    sim, _, _ = _floor_case(v0, model="VISCOUS")
    sim.cycle(60)
    assert _slip(sim) == pytest.approx(v0, rel=1e-9)
    sim.close()


def test_spring_slip_then_coulomb_rate():
    v0 = 1.0e-2
    mu = 0.5
    sim, m, ks = _floor_case(v0, mu=mu)
    print("[gpu spring] sliding at the Coulomb limit", flush=True)
    sim.cycle(10)
    v10 = _slip(sim)
    sim.cycle(30)
    v40 = _slip(sim)
    rate = (v40 - v10) / 30.0
    assert rate == pytest.approx(-3.5 * mu * G, rel=1e-2), rate
    # after the slip speed reaches zero the ball sticks: the undamped spring then
    # oscillates from the Coulomb limit, so the slip speed stays within
    # w * (mu F_n / k_s) instead of falling on at the sliding rate
    sim.cycle(60)
    w = math.sqrt(3.5 * ks / m)
    assert abs(_slip(sim)) <= 1.05 * w * mu * m * G / ks
    sim.close()


def _collision(gap):
    """Sphere 1 at rest, sphere 2 thrown at it along -x with impact parameter 1.5; run until they separate.

    Returns the final element states, whether a neighbour rebuild happened while they touched, and the
    initial and final momentum, angular momentum and kinetic energy.
    """
    r = 3.0
    cfg = base_config(lbSolve=0, lbSizeX=40, lbSizeY=40, lbSizeZ=40, boundary4=4, boundary5=4, lbFZ=0.0,
                      density=RHO, youngMod=10.0, poisson=NU, restitution=1.0, viscTang=0.0,
                      frictionCoefPart=10.0, frictionCoefWall=10.0, multiStep=10)
    x = [[12.0, 20.0, 20.0], [12.0 + 2.0 * r + gap, 21.5, 20.0]]
    v = [[0.0, 0.0, 0.0], [-0.02, 0.0, 0.0]]
    m = 4.0 / 3.0 * math.pi * r ** 3 * RHO
    inertia = 0.4 * m * r * r

    def invariants(s):
        ex, ew = s.get("e_x"), s.get("e_w")
        pos, vel, w = ex[:, 0], ex[:, 1], ew[:, 0]
        ke = 0.5 * m * np.sum(vel ** 2) + 0.5 * inertia * np.sum(w ** 2)
        rel, vrel = pos - pos.mean(axis=0), vel - vel.mean(axis=0)  # about the centre of mass
        return m * vel.sum(axis=0), (m * np.cross(rel, vrel) + inertia * w).sum(axis=0), ke

    sim = _sim(cfg, x, [r, r], v)
    before = invariants(sim)
    touching = was = rebuild_in_contact = False
    reb = float(sim.get("rebuilds")[0])
    for _ in range(400):
        sim.cycle(1)
        pos = sim.get("e_x")[:, 0]
        touching = np.linalg.norm(pos[1] - pos[0]) < 2.0 * r
        rb = float(sim.get("rebuilds")[0])
        rebuild_in_contact |= touching and rb > reb
        reb = rb
        was |= touching
        if was and not touching:
            break
    assert was and not touching, "the spheres did not meet and separate"
    after = invariants(sim)
    state = sim.get("e_x")[:, 1].copy(), sim.get("e_w")[:, 0].copy()
    sim.close()
    return state, rebuild_in_contact, before, after, (m, inertia)


def test_spring_oblique_collision_invariants_and_rebuild():
    print("[gpu spring] oblique collision, contact starting 0.2 apart (no rebuild in contact)", flush=True)
    (v_a, w_a), reb_a, (p0, l0, e0), (p1, l1, e1), (m, inertia) = _collision(0.2)
    assert not reb_a
    assert np.allclose(p1, p0, rtol=0, atol=1e-12 * m)
    spin = np.abs(w_a).max()
    assert spin > 1e-5, "the spring left no spin"
    # the contact torques use the lever arm r en (DEM.cpp:1856-1858), not the contact point at r - overlap / 2,
    # so each contact leaves a net moment overlap x F_t: about 1 % of the spin here; a torque of the wrong
    # sign or size would leave an error of the order of 2 inertia * spin
    assert np.abs(l1 - l0).max() <= 0.05 * inertia * spin, (l0, l1, inertia * spin)
    assert -0.02 < e1 / e0 - 1.0 < 0.0, e1 / e0 - 1.0
    print("[gpu spring] the same collision with a neighbour rebuild inside the contact", flush=True)
    (v_b, w_b), reb_b, _, _, _ = _collision(2.0)
    assert reb_b, "no neighbour rebuild happened during the contact"
    assert np.allclose(v_b, v_a, rtol=0, atol=1e-9 * np.abs(v_a).max())
    assert np.allclose(w_b, w_a, rtol=0, atol=1e-9 * np.abs(w_a).max())


def test_spring_every_wall_single_contact_equals_first_wall():
    print("[gpu spring] floor slide with every wall in range against the first wall only", flush=True)
    out = {}
    for walls in ("FIRST", "ALL"):
        sim, _, _ = _floor_case(1.0e-2, walls=walls)
        sim.cycle(60)
        out[walls] = sim.get("e_x").copy(), sim.get("e_w").copy(), sim.get("walls")[:, 6:9].copy()
        sim.close()
    for a, b in zip(out["FIRST"], out["ALL"]):
        assert np.array_equal(a, b)


def test_spring_every_wall_corner_stick_oscillation():
    """Sphere in the corner of the x = 0.5 wall and the floor, gravity (-g, 0, -g), speed v0 along y."""
    v0 = 1.0e-4
    cfg = base_config(lbSolve=0, lbSizeX=24, lbSizeY=24, lbSizeZ=20, lbFX=-G, lbFZ=-G, boundary0=7, boundary1=7,
                      density=RHO, youngMod=E, poisson=NU, restitution=0.88, viscTang=0.0, frictionCoefWall=0.5,
                      frictionCoefPart=0.5, multiStep=10, wallContacts="ALL")
    m = 4.0 / 3.0 * math.pi * R ** 3 * RHO
    kn = 2.0 / 3.0 * E / (1.0 - NU * NU)
    two_delta = (m * G / (kn * math.sqrt(R))) ** (2.0 / 3.0)  # static overlap of each contact
    c0 = 0.5 + R - 0.5 * two_delta
    sim = _sim(cfg, [[c0, 12.0, c0]], [R], [[0.0, v0, 0.0]])
    assert sim.get("nearWallTable").tolist() == [0, 0, 0, 2]  # walls x = 0.5, x = 22.5, floor, ceiling
    ks = 2.0 * E / ((2.0 - NU) * (1.0 + NU)) * math.sqrt(R) * (m * G) ** (1.0 / 3.0)
    print("[gpu spring] stick oscillation in a corner, two wall springs", flush=True)
    t, vf, vsd = [0.0], [v0], [v0]
    for k in range(1, 121):
        sim.cycle(1)
        x, w = sim.get("e_x")[0], sim.get("e_w")[0]
        t.append(float(k))
        vf.append(x[1, 1] + R * w[0, 0])   # slip along y at the floor contact (0, 0, -R)
        vsd.append(x[1, 1] - R * w[0, 2])  # at the side contact (-R, 0, 0)
    pos = sim.get("e_x")[0][0]
    sim.close()
    # the sphere stays at its static position: the oscillation is purely tangential
    assert abs(pos[0] - c0) <= 1e-9 * c0 and abs(pos[2] - c0) <= 1e-9 * c0
    # symmetric mode: both contacts slip alike
    assert np.max(np.abs(np.array(vf) - np.array(vsd))) <= 1e-6 * v0
    z = _zero_crossings(np.array(t), np.array(vf))
    assert len(z) >= 3, z
    w = math.sqrt(4.5 * ks / m)
    half = (z[2] - z[0]) / 2.0
    assert abs(half - math.pi / w) <= 5e-3 * math.pi / w, (half, math.pi / w, math.pi / math.sqrt(3.5 * ks / m))
    assert abs(z[0] - 0.5 * math.pi / w) <= 1e-2 * math.pi / w, (z[0], 0.5 * math.pi / w)
    assert max(abs(v) for v in vf[30:]) == pytest.approx(v0, rel=2e-2)
