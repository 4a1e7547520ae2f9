"""Frictional apparent viscosity (HYB_RHEOLOGY_FRICTIONAL), pinned by its definition.

Not in the reference, whose non-Newtonian law is Bingham (node.cpp:103-145).
The frictional law replaces the constant yield stress by one proportional to the
gauge pressure (n - rho0) c_s^2 (none under tension):
nu = plasticVisc + frictionCoefFluid * max(n - 1, 0) / 3 / (2 n shearRate),
clamped to the Bingham range [(0.501 - 0.5)/3, (1.8 - 0.5)/3], with shearRate
the second invariant of the strain-rate estimate from the non-equilibrium
populations, evaluated with the node's previous viscosity (node.cpp:103-123).

Checks: (1) with frictionCoefFluid = 0 it is the Newtonian fluid of viscosity
plasticVisc; (2) one collision from a sheared, gravity-loaded channel state
gives, node by node, the viscosity that an independent numpy evaluation of the
definition gives from the same pre-collision populations (FP64, 1e-12); a
share of the nodes lies strictly inside the clamp range, others have no yield
stress or sit at the upper clamp, so every branch is exercised; (3) the FP32
form follows FP64.
"""
import numpy as np
import pytest

from hybird_amd.cfg import base_config
from helpers import fixture_config, rel_l2

pytestmark = pytest.mark.gpu

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1], dtype=np.float64)
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0], dtype=np.float64)
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1], dtype=np.float64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
NU_P = 0.02
F = np.array([1.0e-4, 0.0, -2.0e-4])
MIN_V, MAX_V = (0.501 - 0.5) / 3.0, (1.8 - 0.5) / 3.0


def _channel(mu, **kw):
    return base_config(lbSizeX=12, lbSizeY=8, lbSizeZ=18, initVisc=NU_P, plasticVisc=NU_P, lbFX=F[0], lbFZ=F[2],
                       initVelocityX=0.03, rheology="FRICTIONAL", frictionCoefFluid=mu, **kw)


def _sim(cfg, precision="FP64"):
    from hybird_amd import runtime
    return runtime.Simulation(fixture_config(cfg, precision=precision), np.zeros((0, 20)))


def frictional_visc(f, nu_old, mu):
    """The definition, from pre-collision populations f [n, 19] and the previous viscosity."""
    n = f.sum(axis=1)
    u = np.stack([f @ CX, f @ CY, f @ CZ], axis=1) / n[:, None]
    u = u + 0.5 * F[None, :] / n[:, None]  # velocity shifted by half the force (node.cpp:85-88)
    cu = u[:, 0:1] * CX + u[:, 1:2] * CY + u[:, 2:3] * CZ
    usq = (u * u).sum(axis=1)[:, None]
    feq = W * n[:, None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
    d = f - feq
    tau = 0.5 + 3.0 * nu_old
    c = [CX, CY, CZ]
    g = np.empty((f.shape[0], 3, 3))
    for a in range(3):
        for b in range(3):
            g[:, a, b] = (d * (c[a] * c[b])).sum(axis=1) * 1.5 / (tau * n)
    sr = np.sqrt(0.5 * (g[:, 0, 0] ** 2 + g[:, 1, 1] ** 2 + g[:, 2, 2] ** 2
                        + 2.0 * (g[:, 0, 1] * g[:, 1, 0] + g[:, 2, 0] * g[:, 0, 2] + g[:, 1, 2] * g[:, 2, 1])))
    ty = mu * np.maximum(n - 1.0, 0.0) / 3.0
    with np.errstate(divide="ignore", invalid="ignore"):
        nu = np.where(ty > 0.0, NU_P + ty / (n * 2.0 * sr), NU_P)
    nu = np.where(np.isnan(nu), MAX_V, nu)
    return np.maximum(MIN_V, np.minimum(MAX_V, nu))


def test_frictional_without_friction_is_newtonian():
    a = _sim(_channel(0.0))
    b = _sim(dict(_channel(0.0), rheology="BINGHAM"))
    print("[gpu rheology] frictional(0) vs Newtonian", flush=True)
    a.cycle(15)
    b.cycle(15)
    t = a.get("types")
    act = (t == 0) | (t == 3)
    assert rel_l2(a.get("f")[act], b.get("f")[act]) <= 1e-15
    assert np.all(a.get("visc")[act] == NU_P)
    a.close()
    b.close()


def test_frictional_viscosity_matches_its_definition():
    mu = 1.5
    sim = _sim(_channel(mu))
    print("[gpu rheology] frictional viscosity, one collision against the definition", flush=True)
    sim.cycle(12)
    t = sim.get("types")
    act = (t == 0) | (t == 3)
    f = sim.get("f")[act]
    nu_old = sim.get("visc")[act]
    sim.cycle(1)
    nu_gpu = sim.get("visc")[act]
    nu_ref = frictional_visc(f, nu_old, mu)
    sim.close()
    assert rel_l2(nu_gpu, nu_ref) <= 1e-12, rel_l2(nu_gpu, nu_ref)
    # every branch is exercised: no yield stress (gauge pressure <= 0), the pressure term inside the clamp
    # range, and the upper clamp
    inside = (nu_ref > NU_P * 1.01) & (nu_ref < MAX_V * 0.999)
    assert inside.mean() > 0.05, inside.mean()
    assert np.mean(nu_ref == NU_P) > 0.05 and np.mean(nu_ref == MAX_V) > 0.05


def test_frictional_fp32_follows_fp64():
    a = _sim(_channel(1.5), "FP64")
    b = _sim(_channel(1.5), "FP32")
    a.cycle(20)
    b.cycle(20)
    t = a.get("types")
    act = (t == 0) | (t == 3)
    assert rel_l2(b.get("u")[act], a.get("u")[act]) <= 1e-4
    assert rel_l2(b.get("visc")[act], a.get("visc")[act]) <= 1e-3
    a.close()
    b.close()
