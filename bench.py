"""Benchmark of the gfx950 DEM step (one GPU).

    python bench.py [--steps K] [--warmup W] [--workload headline|small|large]

Times K steps of the headline configuration (250,000 tetrahedral 4-sphere clumps, FP64,
sphere-cell list rebuilt every step, spatial reorder every 50 steps) with inputs resident
in HBM, and prints one JSON line: clump-steps/s, the roofline of the dominant kernel
(k_force, from HIP events over the timed region) and the CPU baseline (the oracle on one
host core, bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "clump-steps/s + HBM GB/s at 250k clumps; batched vs plain; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per clump and launch of k_force: carried state read (rr, R, v, w, a,
# alpha = 24 doubles, 192 B), v, w, a, alpha written (96 B), radius (8 B) and the sorted
# sphere index of its 4 spheres (16 B).
FORCE_BYTES_PER_CLUMP = 312.0
# Algorithmic bytes per clump-step of the whole step (480 B).
STEP_BYTES_PER_CLUMP = 480.0


def cpu_baseline(inst, budget_s: float) -> dict:
    """The oracle (single-threaded C restatement of the reference host path) on the same
    workload: initial force pass, then as many steps as fit in about budget_s seconds."""
    from oracle import oracle
    o = oracle.OracleSim(inst.sys, inst.geo, inst.fun, inst.ntags, inst.phase, inst.particles)
    o.initial()
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    k = max(1, int(budget_s / max(one, 1e-6)) - 1)
    t0 = time.perf_counter()
    o.step(k)
    dt = time.perf_counter() - t0
    n = len(inst.particles)
    return {"value": n * k / dt, "unit": "clump-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle, {n} clumps, steps 2..{k + 1} of the same workload "
                      f"({dt:.1f} s on one host core)"}


def batch_compare(steps: int, warmup: int = 20) -> dict:
    """64 independent flumes (N ~ log-uniform[5e3, 6e4], per-instance seeds, mu, eps and
    slope) stepped together in one context vs the same instances as sequential plain runs
    (one context each, synchronised between instances)."""
    from avalancheflume3d_amd import dem, workloads
    insts = workloads.batch(64)
    total = sum(len(i.particles) for i in insts)
    ctx = dem.Context(insts)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    t_batch = time.perf_counter() - t0
    ctx.close()
    t_seq = 0.0
    for inst in insts:
        c = dem.Context([inst])
        c.regrid()
        c.initial_forces()
        c.step(warmup)
        c.sync()
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        t_seq += time.perf_counter() - t0
        c.close()
    return {"instances": len(insts), "clumps": total, "steps": steps,
            "batched_clump_steps_per_s": total * steps / t_batch,
            "sequential_clump_steps_per_s": total * steps / t_seq,
            "speedup": t_seq / t_batch}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="headline", choices=["headline", "small", "large"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding of "
              "this path is design-only (see DESIGN.md)", file=sys.stderr)
        return 2

    import torch

    from avalancheflume3d_amd import dem, workloads

    make = {"headline": workloads.headline, "small": workloads.small, "large": workloads.large}
    inst = make[args.workload]()
    n = len(inst.particles)
    opt = dem.default_options()
    ctx = dem.Context([inst], opt)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(args.warmup)
    ctx.sync()
    torch.cuda.synchronize()
    # timed region: K steps with HIP events between consecutive kernels on the library's
    # stream (per-kernel durations for the roofline)
    t0 = time.perf_counter()
    kern, total_ms = ctx.step_profiled(args.steps)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    diag = ctx.diag()
    # the same number of steps through the captured-graph path, for comparison
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.sync()
    torch.cuda.synchronize()
    wall_graph = time.perf_counter() - t0
    ctx.close()

    value = n * args.steps / wall
    force_ms = kern["force"]
    achieved = FORCE_BYTES_PER_CLUMP * n / (force_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get(args.workload, {}).get("force_bytes_per_launch")
    out = {
        "metric": METRIC, "value": value, "unit": "clump-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} clumps ({4 * n} spheres), FP64, sphere "
                               "cells rebuilt every step, spatial reorder every "
                               f"{opt.reorder_interval} steps", "clumps": n,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "kernel_ms": kern,
        "step_hbm_gbs_algorithmic": STEP_BYTES_PER_CLUMP * n / (wall / args.steps) / 1e9,
        "events_total_ms": total_ms,
        "graph_path_clump_steps_per_s": n * args.steps / wall_graph,
        "diag": diag,
    }
    if args.batch_steps > 0:
        out["batched"] = batch_compare(args.batch_steps)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(inst, args.cpu_seconds)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
