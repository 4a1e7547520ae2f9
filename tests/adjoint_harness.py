"""Python check harness of the GPU adjoint: a torch.autograd.Function over the
C ABI (hyb_adj_forward / hyb_adj_backward), and the small frozen-bed case the
tests use. Test infrastructure only; the library itself has no torch types.

Loss: mean x-momentum sum_i f_i c_ix over the active nodes of the plane
x = plane_x, on the post-stream populations after `steps` LB steps.
"""
from __future__ import annotations

import numpy as np
import torch

from adjoint_reference import CX

TRT_MAGIC = 0.25


def frozen_bed_case(size=(24, 20, 22), tau=0.7):
    """TRT, periodic x/y, walls in z, body force along x, four spheres at rest."""
    from hybird_amd.cfg import base_config
    from hybird_amd._abi import element_records
    X, Y, Z = size
    cfg = base_config(lbSizeX=X, lbSizeY=Y, lbSizeZ=Z, initVisc=(tau - 0.5) / 3.0, lbFX=2.0e-5, lbFZ=-0.5e-5,
                      density=2.5, youngMod=1.0e5, poisson=0.3, restitution=0.88)
    xs = np.array([[6.0, 5.5, 6.0], [15.2, 12.0, 9.5], [9.0, 14.5, 15.0], [19.5, 4.0, 14.0]])
    rs = np.array([3.0, 3.5, 2.5, 3.0])
    return cfg, element_records(xs, rs)


def plane_seed(types, size, plane_x):
    """dL/df [N, 19] of the plane-flux loss."""
    X = size[0]
    t = np.asarray(types).reshape(-1)
    act = (t == 0) | (t == 3)
    sel = act & ((np.arange(t.size) % X) == plane_x)
    seed = np.zeros((t.size, 19))
    seed[sel] = CX[None, :] / max(int(sel.sum()), 1)
    return seed


def make_sim(cfg, rec, tau):
    from hybird_amd import runtime
    from hybird_amd._abi import config_from_cfg
    c = dict(cfg, initVisc=(tau - 0.5) / 3.0)
    return runtime.Simulation(config_from_cfg(c, collision="TRT", trt_magic=TRT_MAGIC), rec)


class LBMPlaneFlux(torch.autograd.Function):
    """loss = LBMPlaneFlux.apply(force_field [N,3] f64, tau (0-d f64), f0 [N,19] f64, case).

    forward: builds the GPU simulation at this tau, sets the force field (as
    float32) and the initial populations, runs hyb_adj_forward; backward: seeds
    hyb_adj_backward with dL/df_T and returns the three gradients.
    """

    @staticmethod
    def forward(ctx, force_field, tau, f0, case):
        cfg, rec, steps, every, plane_x = case
        sim = make_sim(cfg, rec, float(tau))
        sim.set_force_field(force_field.detach().numpy().astype(np.float32))
        sim.set_populations(f0.detach().numpy())
        sim.adj_forward(steps, every)
        fT = sim.get("f")
        types = sim.get("types")
        seed = plane_seed(types, (int(cfg["lbSizeX"]), int(cfg["lbSizeY"]), int(cfg["lbSizeZ"])), plane_x)
        act = np.isfinite(fT[:, 0])
        loss = float(np.sum(seed[act] * fT[act]))
        ctx.sim, ctx.seed = sim, seed
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        gF, gtau, g0 = ctx.sim.adj_backward(ctx.seed * float(grad_out))
        ctx.sim.close()
        return (torch.from_numpy(gF), torch.tensor(gtau, dtype=torch.float64), torch.from_numpy(g0), None)
