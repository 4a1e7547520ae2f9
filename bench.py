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
FP32 against FP64 is reported with them. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from hybird_amd import runtime  # noqa: E402
from hybird_amd._abi import config_from_cfg, element_records  # noqa: E402
from hybird_amd.cfg import base_config, write_cfg  # noqa: E402
from hybird_amd.particles import rsa_bed, write_empty_object_file, write_particle_file  # noqa: E402

METRIC = "MLUPS + particle-steps/s, FP32-shift vs FP64 plain at <=1e-4 velocity L2 error"
HBM_PEAK_GBS = 8000.0
BYTES_PER_LUP = {"fp32": 2 * 19 * 4, "fp64": 2 * 19 * 8}
# HBM bytes per launch of k_fused32 on the 256^3 workload: FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc,
# scaled by the same counters on a plain copy (profiles/hbm_traffic_fused_dense_suspension_256.md)
TRAFFIC_FUSED32_256 = 1.542e9 + 1.231e9


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


def cpu_baseline(size_sub: int, cycles: int):
    """The reference itself (oracle/_ref, built from its sources) on a bounded sub-box."""
    harness = os.path.join(ROOT, "oracle", "_ref", "hybird_ref_dump")
    threads = min(16, os.cpu_count() or 1)
    cfg, bed = workload(size_sub, seed=1)
    with tempfile.TemporaryDirectory() as tmp:
        pf, of = os.path.join(tmp, "p.dat"), os.path.join(tmp, "o.dat")
        write_particle_file(pf, bed)
        write_empty_object_file(of)
        cfg = dict(cfg, particleFile=pf, objectFile=of, maximumTimeSteps=cycles)
        cp = os.path.join(tmp, "c.cfg")
        write_cfg(cp, cfg)
        work = os.path.join(tmp, "w")
        os.makedirs(work)
        sample = (f"{size_sub}^3 sub-box of the same workload ({bed.count} spheres, same radii and solid fraction), "
                  f"BGK tau=0.6 (the reference has no TRT), FP64, {cycles} cycles, steady-state (T(n)-T(1))/(n-1)")
        if os.path.exists(harness):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            out = subprocess.run([harness, cp, work, "-", str(cycles)], env=env, capture_output=True, text=True,
                                 timeout=600, check=True).stdout.strip().splitlines()[-1]
            r = json.loads(out)
            mlups = r["active_nodes"] / r["seconds_per_cycle"] / 1e6
            return {"value": round(mlups, 3), "unit": "MLUPS", "cores": threads, "kind": "reference",
                    "sample": sample}
        # restatement fallback: one thread
        sys.path.insert(0, ROOT)
        from oracle.oracle import OracleSim

        conf = config_from_cfg(cfg)
        o = OracleSim(conf, element_records(bed.x, bed.r))
        o.cycle(1)
        t0 = time.perf_counter()
        o.cycle(cycles - 1)
        dt = (time.perf_counter() - t0) / (cycles - 1)
        act = int(np.sum((o.get("types") == 0) | (o.get("types") == 3)))
        return {"value": round(act / dt / 1e6, 3), "unit": "MLUPS", "cores": 1, "kind": "port", "sample": sample}


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
    # plain forward of the same steps
    plain = runtime.Simulation(conf, rec)
    plain.set_force_field(ff)
    plain.sync()
    t0 = time.perf_counter()
    plain.step_lb(steps)
    plain.sync()
    t_plain = time.perf_counter() - t0
    plain.close()
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
    sim.close()
    return {"workload": f"frozen bed {size}^3 TRT tau {tau}, {bed.count} spheres r=4 (seed {seed}), "
                        f"loss = mean x-momentum at x=120 after {steps} steps, checkpoints every {every}",
            "steps": steps, "loss": loss, "forward_ms_per_step": round(t_plain / steps * 1e3, 4),
            "forward_with_checkpoints_ms_per_step": round(t_fwd / steps * 1e3, 4),
            "backward_ms_per_step": round(t_bwd / steps * 1e3, 4),
            "backward_over_forward": round(t_bwd / t_plain, 3),
            "grad_tau": gtau, "grad_force_norm": float(np.linalg.norm(gF)), "grad_f0_norm": float(np.linalg.norm(g0))}


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
    ap.add_argument("--cpu-cycles", type=int, default=150)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost measurement")
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
    # accuracy of the FP32 form against FP64 after the same number of cycles
    t_o = sim64.get("types")
    act = (t_o == 0) | (t_o == 3)
    u32, u64 = sim32.get("u")[act], sim64.get("u")[act]
    vel_err = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    f32, f64 = sim32.get("e_forces")[:, 0], sim64.get("e_forces")[:, 0]
    drag_err = float(np.linalg.norm(f32 - f64) / max(np.linalg.norm(f64), 1e-300))
    sim32.close()
    sim64.close()
    K = args.steps
    mlups32 = n_act * K / sec32 / 1e6
    mlups64 = n_act * K / sec64 / 1e6
    psteps = bed.count * 10 * K / sec32
    achieved = BYTES_PER_LUP["fp32"] * n_act / (kms32 * 1e-3) / 1e9
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
        "velocity_rel_l2_fp32_vs_fp64": vel_err,
        "drag_rel_l2_fp32_vs_fp64": drag_err,
        "lb_kernel_ms": {"fp32": round(kms32, 4), "fp64": round(kms64, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC_FUSED32_256 if args.size == 256 else None,
                     "kernel": "k_fused32 (TRT, pull)", "bytes_per_lup": BYTES_PER_LUP["fp32"]},
    }
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
