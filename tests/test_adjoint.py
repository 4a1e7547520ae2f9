"""Adjoint of the LB step with frozen particles.

CPU: the torch model of tests/adjoint_reference.py reproduces the oracle's LB
steps (so its autograd gradient is a gradient of the reference's step).
GPU: the hand-written adjoint (through the autograd.Function harness) matches
the torch autograd gradient (FP64, 1e-9 relative L2) and central finite
differences of the GPU forward on tau and on 32 random nodes (h = 1e-4
relative to tau and to f0, 1e-2 relative to the uniform body force for the
force field; 1e-5 relative error plus a noise floor of 1e-7 of the largest
gradient component).
"""
import numpy as np
import pytest
import torch

from adjoint_harness import TRT_MAGIC, LBMPlaneFlux, frozen_bed_case, make_sim, plane_seed
from adjoint_reference import Lattice
from helpers import rel_l2

STEPS, EVERY, PLANE = 30, 7, 17
TAU = 0.5 + 3.0 * ((0.7 - 0.5) / 3.0)  # the relaxation time the library derives from initVisc


def _lattice(types, pflag, cfg):
    size = (int(cfg["lbSizeX"]), int(cfg["lbSizeY"]), int(cfg["lbSizeZ"]))
    per = tuple(int(cfg[f"boundary{2 * k}"]) == 4 for k in range(3))
    t = np.asarray(types)
    act = (t == 0) | (t == 3)
    return Lattice(t, np.asarray(pflag).astype(bool) & act, size, per), size


def _uniform_force(cfg, lat, ff=None):
    F = torch.tensor([cfg["lbFX"], cfg.get("lbFY", 0.0), cfg["lbFZ"]], dtype=torch.float64)[:, None]
    F = F.expand(3, lat.active.size)
    return F if ff is None else F + ff


def test_torch_model_matches_oracle_lb_steps():
    from hybird_amd._abi import config_from_cfg
    from oracle.oracle import OracleSim
    cfg, rec = frozen_bed_case()
    orc = OracleSim(config_from_cfg(cfg, collision="TRT", trt_magic=TRT_MAGIC), rec)
    lat, _ = _lattice(orc.get("types"), orc.get("pflag"), cfg)
    assert lat.pmask.sum() > 100, "the case must contain particle nodes"
    g = lat.to_active(orc.get("f"))
    F = _uniform_force(cfg, lat)
    for step in range(1, 6):
        orc.lb_step(1)
        g = lat.stream(lat.collide(g, F, torch.tensor(TAU, dtype=torch.float64), TRT_MAGIC))
        ref = lat.to_active(orc.get("f"))
        assert rel_l2(g.numpy(), ref.numpy()) <= 1e-13, f"step {step}"


@pytest.mark.gpu
def test_gpu_adjoint_matches_autograd_and_fd():
    cfg, rec = frozen_bed_case()
    print("[gpu adjoint] frozen bed 24x20x22, 30 steps, checkpoints every 7", flush=True)
    sim = make_sim(cfg, rec, TAU)
    types, pflag, f0 = sim.get("types"), sim.get("pflag"), sim.get("f")
    sim.close()
    lat, size = _lattice(types, pflag, cfg)
    N = types.size
    rng = np.random.default_rng(3)
    ff32 = rng.normal(0.0, 1.0e-7, size=(N, 3)).astype(np.float32)
    act = lat.active
    f0 = np.where(np.isfinite(f0), f0, 0.0)
    # GPU adjoint through the autograd.Function harness
    ffield = torch.tensor(ff32.astype(np.float64), requires_grad=True)
    tau = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
    f0t = torch.tensor(f0, requires_grad=True)
    case = (cfg, rec, STEPS, EVERY, PLANE)
    loss = LBMPlaneFlux.apply(ffield, tau, f0t, case)
    loss.backward()
    gF, gtau, g0 = ffield.grad.numpy(), float(tau.grad), f0t.grad.numpy()
    # torch autograd of the reference model
    ffr = torch.tensor(ff32[act].T.astype(np.float64), requires_grad=True)
    taur = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
    f0r = lat.to_active(f0).clone().requires_grad_(True)
    gT = lat.run(f0r, _uniform_force(cfg, lat, ffr), taur, TRT_MAGIC, STEPS)
    Lr = (lat.plane_flux_weights(PLANE) * gT).sum()
    Lr.backward()
    lv = float(loss.detach())
    rv = float(Lr.detach()) if hasattr(Lr, "detach") else float(Lr)
    assert abs(lv - rv) <= 1e-12 * abs(rv), (lv, rv)
    assert rel_l2(gF[act].T, ffr.grad.numpy()) <= 1e-9
    assert abs(gtau - float(taur.grad)) <= 1e-9 * abs(float(taur.grad))
    assert rel_l2(g0[act].T, f0r.grad.numpy()) <= 1e-9
    assert np.all(gF[~np.isin(np.arange(N), act)] == 0.0)

    # central finite differences of the GPU forward (h = 1e-4 relative)
    def L(tau_v, ff, f0v):
        s = make_sim(cfg, rec, tau_v)
        s.set_force_field(ff)
        s.set_populations(f0v)
        s.step_lb(STEPS)
        fT = s.get("f")
        s.close()
        seed = plane_seed(types, size, PLANE)
        ok = np.isfinite(fT[:, 0])
        return float(np.sum(seed[ok] * fT[ok]))

    h = 1e-4
    tp, tm = TAU * (1 + h), TAU * (1 - h)
    # the library derives tau from nu = (tau - 1/2)/3 and back: use the values it sees
    tp_eff, tm_eff = 0.5 + 3.0 * ((tp - 0.5) / 3.0), 0.5 + 3.0 * ((tm - 0.5) / 3.0)
    fd_tau = (L(tp, ff32, f0) - L(tm, ff32, f0)) / (tp_eff - tm_eff)
    assert abs(fd_tau - gtau) <= 1e-5 * abs(gtau), (fd_tau, gtau)
    # Force-field perturbations are 100 h relative to the uniform body force: relative to the 1e-7
    # field itself the loss change (~1e-14) would sink into the FP64 rounding noise of two 30-step
    # runs (~1e-17 in L). Central differences cancel the quadratic term, and the loss is linear in
    # the force to far below 1e-5 at this size. The remaining noise floor enters as 1e-7 of the
    # largest gradient component.
    fscale = 100.0 * float(np.hypot(cfg["lbFX"], cfg["lbFZ"]))
    gFmax = float(np.max(np.abs(gF)))
    # 32 random active nodes: one force component and one initial population each
    picks = rng.choice(act, size=32, replace=False)
    for n_i, node in enumerate(picks):
        for k in ((0,) if n_i % 2 == 0 else (2,)):
            ffp, ffm = ff32.copy(), ff32.copy()
            base = float(ff32[node, k])
            d = h * fscale
            ffp[node, k] = np.float32(base + d)
            ffm[node, k] = np.float32(base - d)
            fd = (L(TAU, ffp, f0) - L(TAU, ffm, f0)) / (float(ffp[node, k]) - float(ffm[node, k]))
            assert abs(fd - gF[node, k]) <= 1e-5 * abs(gF[node, k]) + 1e-7 * gFmax, (node, k, fd, gF[node, k])
        j = int(rng.integers(0, 19))
        fp, fm = f0.copy(), f0.copy()
        d = h * abs(f0[node, j])
        fp[node, j] += d
        fm[node, j] -= d
        fd = (L(TAU, ff32, fp) - L(TAU, ff32, fm)) / (fp[node, j] - fm[node, j])
        assert abs(fd - g0[node, j]) <= 1e-5 * abs(g0[node, j]) + 1e-7 * np.max(np.abs(g0)), (node, j, fd, g0[node, j])
