"""FP64 CPU reference for the adjoint: the lattice-Boltzmann step with frozen
particles written in PyTorch (CPU, float64), differentiated by torch.autograd.

Test infrastructure only. The forward model is the reference's LB step
(LB.cpp:221-233: reconstruction, penalty force of particle nodes LB.cpp:1851-1919,
Guo shift and forcing node.cpp:85-184, BGK node.cpp:158-165 or TRT, pull
streaming with half-way bounce-back LB.cpp:1144-1488 and the periodic rewiring
LB.cpp:437-472) plus a per-node body-force field; it is checked against the
oracle's LB step, and its autograd gradient is the independent reference for
the hand-written GPU adjoint.

State convention: g_t are the populations that enter collision t (g_1 = f0,
the initial populations); s_t = C(g_t); g_{t+1} = P(s_t). The loss sees g_{T+1},
the post-stream populations after T steps (the "f" view).
"""
from __future__ import annotations

import numpy as np
import torch

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1])
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])
W = np.array([12.0 / 36.0] + [2.0 / 36.0] * 6 + [1.0 / 36.0] * 12)


class Lattice:
    """Active nodes, the streaming gather and the particle mask of a lattice.

    types: node type codes [N] (0 fluid, 3 interface, 7 static wall, ...);
    pflag: particle-node flags [N]; size (X, Y, Z); periodic flags per axis.
    """

    def __init__(self, types, pflag, size, periodic):
        X, Y, Z = size
        self.size = size
        t = np.asarray(types).reshape(-1)
        self.active = np.flatnonzero((t == 0) | (t == 3))
        n = t.size
        pos = -np.ones(n, dtype=np.int64)
        pos[self.active] = np.arange(self.active.size)
        self.pos = pos
        x = self.active % X
        y = (self.active // X) % Y
        z = self.active // (X * Y)

        def wrap(v, d, L, per):
            w = v + d
            if per:
                w = np.where((d < 0) & (v == 1), L - 2, w)
                w = np.where((d > 0) & (v == L - 2), 1, w)
            return w

        na = self.active.size
        src = np.zeros((19, na), dtype=np.int64)
        src[0] = np.arange(na)
        for i in range(1, 19):
            j = OPP[i]  # population i arrives from the neighbour in direction j
            nx = wrap(x, CX[j], X, periodic[0])
            ny = wrap(y, CY[j], Y, periodic[1])
            nz = wrap(z, CZ[j], Z, periodic[2])
            nb = nx + X * (ny + Y * nz)
            wall = t[nb] == 7
            if np.any((pos[nb] < 0) & ~wall):
                raise ValueError("an active node links to a node that is neither active nor a static wall")
            src[i] = np.where(wall, j * na + np.arange(na), i * na + np.maximum(pos[nb], 0))
        self.src = torch.from_numpy(src.reshape(-1))
        self.pmask = torch.from_numpy(np.asarray(pflag).reshape(-1)[self.active].astype(np.float64))
        self.c = torch.from_numpy(np.stack([CX, CY, CZ]).astype(np.float64))  # [3, 19]
        self.w = torch.from_numpy(W)[:, None]
        self.opp = torch.from_numpy(OPP)

    def to_active(self, f_nodes):
        """[N, 19] host view -> [19, Nact] tensor."""
        return torch.from_numpy(np.ascontiguousarray(np.asarray(f_nodes)[self.active].T))

    def stream(self, s):
        return s.reshape(-1)[self.src].reshape(19, -1)

    def collide(self, g, F, tau, trt_magic, trt=True):
        """C(g): moments, penalty force (frozen particles at rest, node mass 1),
        Guo shift, equilibrium, TRT (or BGK) relaxation with Guo forcing."""
        c, w = self.c, self.w
        n = g.sum(0)
        u = (c @ g) / n
        hF = -(self.pmask / n) * u
        T = F + hF
        us = u + T / (2.0 * n)
        a = c.T @ us  # [19, Nact]
        b = c.T @ T
        q = (us * us).sum(0)
        r = (us * T).sum(0)
        feq = w * n * (1.0 + 3.0 * a + 4.5 * a * a - 1.5 * q)
        omP = 1.0 / tau
        omM = 1.0 / (0.5 + trt_magic / (tau - 0.5)) if trt else omP
        go, eo = g[self.opp], feq[self.opp]
        sp = w * (9.0 * a * b - 3.0 * r)
        sm = w * 3.0 * b
        return (g - omP * (0.5 * (g + go) - 0.5 * (feq + eo)) - omM * (0.5 * (g - go) - 0.5 * (feq - eo))
                + (1.0 - 0.5 * omP) * sp + (1.0 - 0.5 * omM) * sm)

    def run(self, f0, F, tau, trt_magic, steps, trt=True):
        """g_{T+1} after `steps` collisions starting from g_1 = f0 ([19, Nact])."""
        g = f0
        for _ in range(steps):
            g = self.stream(self.collide(g, F, tau, trt_magic, trt))
        return g

    def plane_flux_weights(self, plane_x):
        """Loss weights: mean x-momentum (sum_i f_i c_ix) over the active nodes of plane x = plane_x."""
        X = self.size[0]
        sel = (self.active % X) == plane_x
        wts = np.zeros((19, self.active.size))
        wts[:, sel] = CX[:, None] / max(int(sel.sum()), 1)
        return torch.from_numpy(wts)
