#!/usr/bin/env python3
"""Throughput of the coupled LBM-DEM step on one MI355X.

Workload (dense suspension): 256^3 D3Q19 lattice, TRT collision (magic 1/4),
tau = 0.6, periodic in x and y, bounce-back walls in z, body force 1e-6 (lattice
units) along x; single spheres with radii uniform in [3, 6] at solid fraction
0.2 (7735 spheres) placed by seeded random sequential addition (seed 1);
Hertz normal + Coulomb-capped viscous tangential contacts (E = 1e5, nu = 0.3,
e = 0.88, mu = 0.5), 10 DEM sub-steps per LB step. A "step" is one coupled
cycle: DEM sub-steps, particle-node classification, the fused LB kernel and the
per-particle hydrodynamic reduction.

The headline value is the MLUPS of the FP32 deviation-stored path; the plain
FP64 path runs the same workload beside it, and the velocity-field error of
FP32 against FP64 is reported with them; the FP32 path is timed once more with
the tangential contact spring (history per contact) in place of the viscous law.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from hybird_amd import runtime  # noqa: E402
from hybird_amd._abi import config_from_cfg, element_records  # noqa: E402
from hybird_amd.cfg import base_config  # noqa: E402
from hybird_amd.particles import rsa_bed  # noqa: E402

METRIC = "MLUPS + particle-steps/s, FP32-shift vs FP64 plain at <=1e-4 velocity L2 error"
HBM_PEAK_GBS = 8000.0
BYTES_PER_LUP = {"fp32": 2 * 19 * 4, "fp64": 2 * 19 * 8}
# HBM bytes per launch of k_fused32 on the 256^3 workload: FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc,
# scaled by the same counters on a plain copy (profiles/hbm_traffic_fused_dense_suspension_256.md)
TRAFFIC_FUSED32_256 = 1.529e9 + 1.212e9


def workload(size: int, seed: int = 1, phi: float = 0.2):
    tau = 0.6
    cfg = base_config(lbSizeX=size, lbSizeY=size, lbSizeZ=size, initVisc=(tau - 0.5) / 3.0, lbFX=1.0e-6,
                      boundary0=4, boundary1=4, boundary2=4, boundary3=4, boundary4=7, boundary5=7,
                      density=2.5, youngMod=1.0e5, poisson=0.3, restitution=0.88, frictionCoefPart=0.5,
                      frictionCoefWall=0.5, viscTang=0.2, multiStep=10)
    span = size - 2.0
    mean_vol = 4.0 / 3.0 * math.pi * ((6.0 ** 4 - 3.0 ** 4) / (4.0 * 3.0))
    count = int(round(phi * span ** 3 / mean_vol))
    bed = rsa_bed(seed, (0.5, 0.5, 0.5), (0.5 + span, 0.5 + span, 0.5 + span), lambda r: r.uniform(3.0, 6.0), count,
                  periodic=(True, True, False), period=(span, span, 0.0))
    return cfg, bed


def make_sim(cfg, bed, precision):
    conf = config_from_cfg(cfg, collision="TRT", precision=precision)
    return runtime.Simulation(conf, element_records(bed.x, bed.r))


def timed(sim, steps, warmup):
    """Seconds for `steps` cycles, mean fused-kernel ms, neighbour rebuilds inside the timed cycles."""
    sim.cycle(warmup)
    sim.sync()
    r0 = float(sim.get("rebuilds")[0])
    sim.set_profiling(True)
    t0 = time.perf_counter()
    sim.cycle(steps)
    sim.sync()
    t1 = time.perf_counter()
    ms, launches = sim.kernel_time()
    sim.set_profiling(False)
    return t1 - t0, ms / max(launches, 1), int(float(sim.get("rebuilds")[0]) - r0)


def host_cpu():
    """Model name, physical cores and logical CPUs of this host (from /proc/cpuinfo)."""
    model, cores, logical = "unknown", set(), 0
    try:
        with open("/proc/cpuinfo") as fh:
            phys = core = None
            for line in fh:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    return {"model": model, "physical_cores": len(cores) or None, "logical_cpus": logical or os.cpu_count()}


def cpu_baseline(size_sub: int, cycles: int):
    """The FP64 CPU restatement of the reference's cycle (oracle/, one thread) on a bounded
    sub-box of the same workload: same radii, solid fraction, TRT collision and contacts."""
    from oracle.oracle import OracleSim

    cfg, bed = workload(size_sub, seed=1)
    conf = config_from_cfg(cfg, collision="TRT", precision="FP64")
    o = OracleSim(conf, element_records(bed.x, bed.r))
    o.cycle(1)
    t0 = time.perf_counter()
    o.cycle(cycles - 1)
    dt = (time.perf_counter() - t0) / (cycles - 1)
    t = o.get("types")
    act = int(np.sum((t == 0) | (t == 3)))
    o.close()
    sample = (f"{size_sub}^3 sub-box of the same workload ({bed.count} spheres, same radii and solid fraction, TRT "
              f"tau 0.6, 10 DEM sub-steps), FP64, {cycles} cycles, steady-state (T(n)-T(1))/(n-1), one thread")
    return {"value": round(act / dt / 1e6, 3), "unit": "MLUPS", "cores": 1, "kind": "port", "sample": sample,
            "particle_steps_per_s": round(bed.count * 10 / dt, 1), "host": host_cpu()}


def adjoint_cost(size: int = 128, steps: int = 1000, every: int = 50, count: int = 2000, seed: int = 3):
    """Cost of the hand-written adjoint against the forward on a frozen bed: 128^3 D3Q19 TRT tau 0.7,
    2000 spheres of radius 4 (seed 3) at rest, body force 1e-6 x plus a per-node Gaussian field
    (sigma 1e-7, float32), loss = mean x-momentum through the plane x = 120 after 1000 steps, state
    checkpoints every 50 steps. Gradients w.r.t. the force field, tau and the initial populations."""
    tau = 0.7
    cfg = base_config(lbSizeX=size, lbSizeY=size, lbSizeZ=size, initVisc=(tau - 0.5) / 3.0, lbFX=1.0e-6,
                      boundary0=4, boundary1=4, boundary2=4, boundary3=4, boundary4=7, boundary5=7, density=2.5)
    span = size - 2.0
    bed = rsa_bed(seed, (0.5, 0.5, 0.5), (0.5 + span, 0.5 + span, 0.5 + span), lambda r: 4.0, count,
                  periodic=(True, True, False), period=(span, span, 0.0))
    conf = config_from_cfg(cfg, collision="TRT", precision="FP64")
    rec = element_records(bed.x, bed.r)
    N = size ** 3
    ff = np.random.default_rng(seed).normal(0.0, 1.0e-7, size=(N, 3)).astype(np.float32)
    sim = runtime.Simulation(conf, rec)
    sim.set_force_field(ff)
    sim.sync()
    t0 = time.perf_counter()
    sim.adj_forward(steps, every)
    sim.sync()
    t_fwd = time.perf_counter() - t0
    fT = sim.get("f")
    types = sim.get("types")
    act = (types == 0) | (types == 3)
    sel = act & ((np.arange(N) % size) == 120)
    cxv = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1], dtype=np.float64)
    sd = np.zeros((N, 19))
    sd[sel] = cxv[None, :] / sel.sum()
    loss = float(np.sum(sd[sel] * fT[sel]))
    t0 = time.perf_counter()
    gF, gtau, g0 = sim.adj_backward(sd)
    t_bwd = time.perf_counter() - t0
    # device time of the two C-side loops (HIP events on the library stream): the forward's LB steps
    # and checkpoint copies, the reverse pass's segment replays and adjoint steps
    d_fwd, d_bwd = sim.adj_device_time()
    sim.close()
    return {"workload": f"frozen bed {size}^3 TRT tau {tau}, {bed.count} spheres r=4 (seed {seed}), "
                        f"loss = mean x-momentum at x=120 after {steps} steps, checkpoints every {every}",
            "steps": steps, "loss": loss, "forward_ms_per_step": round(d_fwd / steps, 4),
            "backward_ms_per_step": round(d_bwd / steps, 4),
            "backward_over_forward": round(d_bwd / d_fwd, 3),
            "wall_forward_s": round(t_fwd, 3), "wall_backward_s": round(t_bwd, 3),
            "grad_tau": gtau, "grad_force_norm": float(np.linalg.norm(gF)), "grad_f0_norm": float(np.linalg.norm(g0))}


def copy_ceiling(nbytes: int = 1 << 31):
    """GB/s of a plain non-temporal device copy (in the library) with 4-, 8- and 16-byte lanes: what a
    streaming kernel can reach on this HBM at each access width."""
    return {f"{w}B_per_lane": round(runtime.copy_bandwidth(nbytes, w, 20), 1) for w in (4, 8, 16)}


def _rel(a, b):
    n = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))


def horizon(sim32, sim64, done: int, steps: int, every: int):
    """FP32 against FP64 on the same workload, continued to `steps` cycles in total:
    velocity-field and drag relative L2 and the largest particle-position difference
    every `every` cycles."""
    t = sim64.get("types")
    act = (t == 0) | (t == 3)
    rows = []
    while done < steps:
        n = min(every - done % every, steps - done)
        sim32.sync()
        t0 = time.perf_counter()
        sim32.cycle(n)
        sim32.sync()
        ms32 = (time.perf_counter() - t0) / n * 1e3  # neighbour rebuilds of these cycles included
        sim64.cycle(n)
        done += n
        u32, u64 = sim32.get("u")[act], sim64.get("u")[act]
        f32, f64 = sim32.get("e_forces")[:, 0], sim64.get("e_forces")[:, 0]
        x32, x64 = sim32.get("e_x")[:, 0], sim64.get("e_x")[:, 0]
        rows.append({"step": done, "velocity_rel_l2": _rel(u32, u64), "drag_rel_l2": _rel(f32, f64),
                     "position_max_abs": float(np.max(np.abs(x32 - x64))),
                     "neighbour_rebuilds_fp32": int(float(sim32.get("rebuilds")[0])),
                     "fp32_ms_per_step": round(ms32, 4)})
        print(f"[bench] horizon step {done}: velocity {rows[-1]['velocity_rel_l2']:.3e}", file=sys.stderr,
              flush=True)
    return rows


# Young's modulus (lattice units) of the two free-surface workloads, which state no contact stiffness: at 10
# DEM sub-steps per LB step (dt = 0.1) a Hertz wall contact of a radius-3 sphere of density 2.5 has omega dt
# of about 1.2 at E = 1e5 (the Gear integration then pumps energy into resting contacts and particles sink
# through the floor within ~2000 steps of the dam break) and about 0.4 at E = 1e4.
YOUNG_FS = 1.0e4
# Both free-surface workloads are closed tanks whose particles reach the corners of the floor and the side
# walls: they take the contact of every wall in range (wallContacts = ALL). With the reference's first wall
# only (DEM.cpp:1496-1513) a particle in such a corner has no floor contact and sinks through the floor (the
# dam break then stops between steps 5000 and 5250, a particle leaving the neighbour-cell grid).


def dam_break(steps: int = 10000, every: int = 2000, phi: float = 0.25, seed: int = 2):
    """Free-surface dam break with Bingham rheology and particles: 384x128x192 D3Q19 closed no-slip tank, liquid
    block x < 128, z < 120, plastic viscosity tau 0.55 (nu = 0.05/3), yield stress 1e-5, gravity -2e-5 (lattice
    units); spheres r ~ U[3,5] filling the block at solid fraction `phi` by RSA (seed 2), E = YOUNG_FS; BGK; FP32 populations
    (FP64 masses) against the plain FP64 path: liquid-mass drift and front position every `every` steps."""
    X, Y, Z = 384, 128, 192
    visc = 0.05 / 3.0
    cfg = base_config(problemName="LIQUID_BOX", freeSurfaceSolve=1, lbSizeX=X, lbSizeY=Y, lbSizeZ=Z,
                      boundary0=7, boundary1=7, boundary2=7, boundary3=7, boundary4=7, boundary5=7,
                      nonNewtonianSolve=1, plasticVisc=visc, yieldStress=1.0e-5, initVisc=visc, lbFZ=-2.0e-5,
                      density=2.5, youngMod=YOUNG_FS, multiStep=10, wallContacts="ALL")
    box = (1.0, 127.0, 1.0, Y - 2.0, 1.0, 119.0)
    vol = 127.0 * (Y - 2.0) * 119.0
    count = int(phi * vol / (4.0 / 3.0 * math.pi * (5.0 ** 4 - 3.0 ** 4) / 8.0))
    bed = rsa_bed(seed, (0.5, 0.5, 0.5), (127.5, Y - 1.5, 119.5), lambda r: r.uniform(3.0, 5.0), count,
                  max_attempts=400000)
    rec = element_records(bed.x, bed.r)
    sims = {k: runtime.Simulation(config_from_cfg(cfg, precision=k, liquid_box=box), rec) for k in ("FP64", "FP32")}

    def front(sim):
        t = sim.get("types").reshape(Z, Y, X)
        liq = (t[1] == 0) | (t[1] == 3)  # the bottom liquid layer
        cols = np.nonzero(liq.any(axis=0))[0]
        return int(cols.max()) if cols.size else -1

    m0 = {k: float(s.get("fluidMass")[0]) for k, s in sims.items()}
    rows = []
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        n = min(every, steps - done)
        ms = {}
        for k, s in sims.items():
            s.sync()
            tc = time.perf_counter()
            s.cycle(n)
            s.sync()
            ms[k] = (time.perf_counter() - tc) / n * 1e3
        done += n
        m = {k: float(s.get("fluidMass")[0]) for k, s in sims.items()}
        rows.append({"step": done, "mass_drift_fp64": m["FP64"] / m0["FP64"] - 1.0,
                     "mass_drift_fp32": m["FP32"] / m0["FP32"] - 1.0,
                     "mass_rel_fp32_vs_fp64": m["FP32"] / m["FP64"] - 1.0,
                     "front_x_fp64": front(sims["FP64"]), "front_x_fp32": front(sims["FP32"]),
                     "neighbour_rebuilds_fp32": int(float(sims["FP32"].get("rebuilds")[0])),
                     "fp32_ms_per_step": round(ms["FP32"], 4), "fp64_ms_per_step": round(ms["FP64"], 4)})
        print(f"[bench] dam break step {done}: front {rows[-1]['front_x_fp64']} / {rows[-1]['front_x_fp32']}",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    for s in sims.values():
        s.close()
    return {"workload": f"dam break {X}x{Y}x{Z}, liquid block x<128 z<120, Bingham tau 0.55 yield 1e-5, gravity "
                        f"-2e-5, {bed.count} spheres r~U[3,5] (phi {phi}, seed {seed}), closed no-slip tank, BGK, "
                        f"contacts with every wall in range",
            "steps": steps, "seconds_both_precisions": round(wall, 2), "rows": rows}


def collapse(steps: int, warmup: int, err_steps: int = 2000, phi: float = 0.2, seed: int = 4):
    """Granular collapse into fluid: 512x512x256 D3Q19 TRT tau 0.6 closed no-slip tank, free surface at z = 160,
    lognormal spheres (mean radius 4, sigma 0.2) packed by RSA at solid fraction `phi` in the column x, y < 160,
    z < 220 (seed 4), density ratio 2.6, Hertz (E = YOUNG_FS) + Coulomb (mu 0.6), gravity -1e-5. Timed FP32 and
    FP64 cycles, then
    FP32 against FP64 after `err_steps` cycles."""
    X, Y, Z = 512, 512, 256
    tau = 0.6
    cfg = base_config(problemName="LIQUID_BOX", freeSurfaceSolve=1, lbSizeX=X, lbSizeY=Y, lbSizeZ=Z,
                      boundary0=7, boundary1=7, boundary2=7, boundary3=7, boundary4=7, boundary5=7,
                      initVisc=(tau - 0.5) / 3.0, lbFZ=-1.0e-5, density=2.6, youngMod=YOUNG_FS, poisson=0.3,
                      restitution=0.88, frictionCoefPart=0.6, frictionCoefWall=0.6, viscTang=0.2, multiStep=10,
                      wallContacts="ALL")
    box = (1.0, X - 2.0, 1.0, Y - 2.0, 1.0, 159.0)
    sig = 0.2
    mu = math.log(4.0) - 0.5 * sig * sig
    ev = 4.0 / 3.0 * math.pi * math.exp(3.0 * mu + 4.5 * sig * sig)
    count = int(phi * 160.0 * 160.0 * 219.5 / ev)
    bed = rsa_bed(seed, (0.5, 0.5, 0.5), (160.5, 160.5, 220.0), lambda r: math.exp(mu + sig * r.normal()), count,
                  max_attempts=400000)
    rec = element_records(bed.x, bed.r)
    out = {"workload": f"granular collapse {X}x{Y}x{Z} TRT tau 0.6, free surface z=160, {bed.count} lognormal spheres "
                       f"(mean r 4, sigma 0.2; phi {phi} in x,y<160 z<220, seed {seed}), rho 2.6, mu 0.6, gravity -1e-5, "
                       f"closed no-slip tank, contacts with every wall in range"}
    sims = {}
    for prec in ("FP32", "FP64"):
        sim = runtime.Simulation(config_from_cfg(cfg, collision="TRT", precision=prec, liquid_box=box), rec)
        n_act = sim.active_nodes()
        sec, kms, reb = timed(sim, steps, warmup)
        key = prec.lower()
        out[f"{key}_mlups"] = round(n_act * steps / sec / 1e6, 1)
        out[f"{key}_ms_per_step"] = round(sec / steps * 1e3, 4)
        out[f"{key}_lb_kernel_ms"] = round(kms, 4)
        if prec == "FP32":
            out["active_nodes_initial"] = n_act
            out["particle_steps_per_s"] = round(bed.count * 10 * steps / sec, 1)
            out["dem_rebuilds_in_timed_steps"] = reb
            gbs = BYTES_PER_LUP["fp32"] * n_act / (kms * 1e-3) / 1e9
            out["lb_kernel_hbm_frac"] = round(gbs / HBM_PEAK_GBS, 4)
        sims[prec] = sim
        print(f"[bench] collapse {prec} timed: {sec:.3f}s", file=sys.stderr, flush=True)
    done = warmup + steps
    if err_steps > done:
        for s in sims.values():
            s.cycle(err_steps - done - 1)
            s.set_capture(True)
            s.cycle(1)
        a, b = sims["FP64"], sims["FP32"]
        ta, tb = a.get("types"), b.get("types")
        act = ((ta == 0) | (ta == 3)) & (ta == tb)
        ma, mb = float(a.get("fluidMass")[0]), float(b.get("fluidMass")[0])
        out["errors_at_step"] = err_steps
        out["velocity_rel_l2_fp32_vs_fp64"] = _rel(b.get("u")[act], a.get("u")[act])
        out["node_types_differing"] = int(np.sum(ta != tb))
        out["mass_rel_fp32_vs_fp64"] = mb / ma - 1.0
        out["position_max_abs_fp32_vs_fp64"] = float(np.max(np.abs(b.get("e_x")[:, 0] - a.get("e_x")[:, 0])))
        out["neighbour_rebuilds_fp32"] = int(float(b.get("rebuilds")[0]))
    for s in sims.values():
        s.close()
    return out


def copy_probe(nbytes: int = 1 << 31, reps: int = 5) -> int:
    """A plain device-to-device copy of nbytes, reps times: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "reps": reps}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-size", type=int, default=64)
    ap.add_argument("--cpu-cycles", type=int, default=60)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost measurement")
    ap.add_argument("--horizon", type=int, default=5000,
                    help="cycles over which FP32 is compared with FP64 on the headline workload (0: skip)")
    ap.add_argument("--dam-break", type=int, default=10000, help="dam-break steps (0: skip)")
    ap.add_argument("--collapse", type=int, default=2000,
                    help="cycles after which the granular collapse compares FP32 with FP64 (0: skip the leg)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of known size (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU slab design is "
              f"described but not built)", file=sys.stderr)
        return 2
    t_setup = time.perf_counter()
    cfg, bed = workload(args.size)
    print(f"[bench] workload {args.size}^3, {bed.count} spheres (setup {time.perf_counter() - t_setup:.1f}s)",
          file=sys.stderr, flush=True)
    sim32 = make_sim(cfg, bed, "FP32")
    n_act = sim32.active_nodes()
    sec32, kms32, reb32 = timed(sim32, args.steps, args.warmup)
    print(f"[bench] fp32 done: {sec32:.3f}s", file=sys.stderr, flush=True)
    sim64 = make_sim(cfg, bed, "FP64")
    sec64, kms64, _ = timed(sim64, args.steps, args.warmup)
    print(f"[bench] fp64 done: {sec64:.3f}s", file=sys.stderr, flush=True)
    # the same workload with the tangential contact spring (history per pair and wall contact) in place of
    # the reference's viscous, history-free tangential law
    sim_sp = make_sim(dict(cfg, tangentialModel="SPRING"), bed, "FP32")
    sec_sp, _, reb_sp = timed(sim_sp, args.steps, args.warmup)
    sim_sp.close()
    print(f"[bench] fp32 spring law done: {sec_sp:.3f}s", file=sys.stderr, flush=True)
    # accuracy of the FP32 form against FP64 after the same number of cycles
    t_o = sim64.get("types")
    act = (t_o == 0) | (t_o == 3)
    u32, u64 = sim32.get("u")[act], sim64.get("u")[act]
    vel_err = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    f32, f64 = sim32.get("e_forces")[:, 0], sim64.get("e_forces")[:, 0]
    drag_err = float(np.linalg.norm(f32 - f64) / max(np.linalg.norm(f64), 1e-300))
    extra = {}
    if args.horizon > args.warmup + args.steps:
        try:
            extra["precision_horizon"] = {"steps": args.horizon, "rows": horizon(sim32, sim64, args.warmup + args.steps,
                                                                                  args.horizon, 1000)}
        except Exception as e:  # the headline line is printed in any case
            extra["precision_horizon"] = {"error": str(e)}
    sim32.close()
    sim64.close()
    if args.collapse:
        try:
            extra["granular_collapse"] = collapse(20, 3, args.collapse)
        except Exception as e:
            extra["granular_collapse"] = {"error": str(e)}
    if args.dam_break:
        try:
            extra["dam_break"] = dam_break(args.dam_break)
        except Exception as e:
            extra["dam_break"] = {"error": str(e)}
    K = args.steps
    mlups32 = n_act * K / sec32 / 1e6
    mlups64 = n_act * K / sec64 / 1e6
    psteps = bed.count * 10 * K / sec32
    achieved = BYTES_PER_LUP["fp32"] * n_act / (kms32 * 1e-3) / 1e9
    # the algorithmic LB bytes of one step over the whole coupled step (DEM, coupling and sums included)
    whole_step_gbs = BYTES_PER_LUP["fp32"] * n_act / (sec32 / K) / 1e9
    line = {
        "metric": METRIC,
        "value": round(mlups32, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(sec32 / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "dense suspension, D3Q19 TRT (magic 1/4) tau 0.6, periodic x/y, walls z, "
                               "body force 1e-6 x, spheres r~U[3,6] at solid fraction 0.2 (seed 1), "
                               "Hertz + Coulomb contacts, 10 DEM sub-steps per LB step",
                   "lattice": [args.size] * 3, "active_nodes": n_act, "particles": bed.count},
        "particle_steps_per_s": round(psteps, 1),
        "dem_rebuilds_in_timed_steps": reb32,
        "fp64_mlups": round(mlups64, 1),
        "fp64_ms_per_step": round(sec64 / K * 1e3, 4),
        "fp32_over_fp64": round(mlups32 / mlups64, 3),
        "spring_law": {"mlups": round(n_act * K / sec_sp / 1e6, 1), "ms_per_step": round(sec_sp / K * 1e3, 4),
                       "particle_steps_per_s": round(bed.count * 10 * K / sec_sp, 1),
                       "dem_rebuilds_in_timed_steps": reb_sp},
        "whole_step_hbm_frac": round(whole_step_gbs / HBM_PEAK_GBS, 4),
        "velocity_rel_l2_fp32_vs_fp64": vel_err,
        "drag_rel_l2_fp32_vs_fp64": drag_err,
        "lb_kernel_ms": {"fp32": round(kms32, 4), "fp64": round(kms64, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC_FUSED32_256 if args.size == 256 else None,
                     "kernel": "k_fused32 (TRT, pull)", "bytes_per_lup": BYTES_PER_LUP["fp32"],
                     "copy_ceiling": copy_ceiling()},
    }
    line.update(extra)
    if not args.no_adjoint:
        print("[bench] adjoint ...", file=sys.stderr, flush=True)
        line["adjoint"] = adjoint_cost()
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(args.cpu_size, args.cpu_cycles)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
