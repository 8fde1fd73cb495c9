"""Benchmark of the gfx950 DEM step (one GPU).

    python bench.py [--steps K] [--warmup W] [--workload headline|small|large]

Times K steps of the headline configuration as stated (250,000 tetrahedral 4-sphere clumps,
FP64, the cell list rebuilt every step, spatial reorder every 50 steps) with inputs resident
in HBM, and prints one JSON line: clump-steps/s, the roofline of the path's dominant kernel
(from HIP events over the timed region), the step-level roofline (480 B per clump-step), and
the CPU baseline (the oracle pinned to one host core, a 200-step window on a slab of the
same workload).  The cell list is binned by clump (the library default; --binning sphere
bins the four spheres of every clump instead, with the same results).  The same workload
binned by sphere, on candidate lists (identical results, fewer rebuilds), the batched
ensemble and the adjoint window are reported under their own keys.
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
# Algorithmic bytes per clump and launch of k_force on a candidate-list step of a lean chunk
# (which also predicts the next step): reads the clump's stage record (rr, radius, v, w, R,
# a: 176 B), its density and reference index (12 B), the angular acceleration (24 B), mass
# and inertia (32 B), the instance index and list length (8 B) and the displacement bound
# (8 B); writes the next stage record (176 B), the angular acceleration (24 B) and the
# bound (8 B).  The list entries (6 B each) and the neighbours' stage records (L2 hits) are
# not counted.
FORCE_BYTES_PER_CLUMP = 468.0
# Mass and inertia are not read where the steps recompute them from radius and density
# (avf_diag.closed_instances says so): 32 B less there.
MASS_INERTIA_BYTES = 32.0
# Algorithmic bytes per clump-step of the whole step (480 B per clump-step).
STEP_BYTES_PER_CLUMP = 480.0
# Candidate-list skin (cm) of the headline runs: 10 % of the sphere radius (a wider skin lists
# more candidates; a narrower one ends chunks and rebuilds the lists sooner).
DEFAULT_SKIN = 0.02
# The timed region brackets runs of TIMED_EVERY consecutive k_force launches (candidate-list
# steps, back to back) with one pair of HIP events, one run in every 2 * TIMED_EVERY steps: a
# regular sample over the whole region, with the events' own cost shared by the run.
TIMED_EVERY = 8
# Spatial reorder interval (steps) of the candidate-list comparison runs.
DEFAULT_REORDER = 400
# Per-step binning by sphere (--binning sphere): algorithmic bytes per clump and launch of
# k_force (a chunk's steps but its last): the clump's stage record (176 B), its density and
# reference index (12 B), the angular acceleration (24 B), mass and inertia (32 B), the
# instance index (4 B), the four per-sphere contact counts (16 B) read; the next stage
# record (176 B), the angular acceleration (24 B) and the four sphere-cell keys (16 B)
# written.  The contact entries (4 B each) and the neighbours' records (gathered, cache
# hits) are not counted; the keys are counted by their own kernel.
CELLS_FORCE_BYTES = 480.0
# Per-step binning by sphere: algorithmic bytes per clump of the contact sweep (k_sweep):
# each binned sphere's sorted record
# (x, y, z, r: 32 B) and id (4 B) read once, two cell starts per sphere (the grid holds about
# two cells per sphere: 8 B), its contact count (4 B) written; contact entries not counted.
SWEEP_BYTES_PER_CLUMP = 4 * (32.0 + 4.0 + 8.0 + 4.0)
# Per-step binning by clump (the default): k_force reads the stage record (176 B), the
# density and reference index (12 B), the angular acceleration (24 B), mass and inertia
# (32 B), the instance index and list length (8 B), and writes the next stage record
# (176 B), the angular acceleration (24 B), the next clump-cell key and the clump's rank in
# its cell (8 B) and the sweep's float record (centre and radius as floats: 16 B): 476 B; list entries, the neighbours' records and the cell
# counters (atomics in L2) not counted.  k_csweep reads each binned clump's centre and radius
# in cell order (16 B), its storage index and key (8 B) and two cell starts (8 B): 32 B; the
# candidates' centres (read again by neighbouring clumps, from cache), the stage records of
# clumps with candidates that pass the bounding test, and the list entries not counted.
CLUMP_FORCE_BYTES = 476.0
CSWEEP_BYTES_PER_CLUMP = 32.0
KERNEL_NAMES = {"sphere": {"force": "k_force", "detect": "k_sweep"},
                "clump": {"force": "k_force", "detect": "k_csweep"}}


def cpu_batch_baseline(budget_s: float, cores: int) -> dict:
    """The oracle on independent instances of the batched ensemble, one per host thread (one
    core each), every thread stepping its instance for about budget_s seconds after its
    initial force pass: the all-cores CPU figure of the batch."""
    from concurrent.futures import ThreadPoolExecutor
    from avalancheflume3d_amd import workloads
    from oracle import oracle
    insts = workloads.batch(64)[:cores]

    def run(inst):
        o = oracle.OracleSim(inst.sys, inst.geo, inst.fun, inst.ntags, inst.phase, inst.particles)
        o.initial()
        t0, k = time.perf_counter(), 0
        while True:
            o.step(1)
            k += 1
            if time.perf_counter() - t0 >= budget_s:
                break
        return len(inst.particles) * k, time.perf_counter() - t0

    with ThreadPoolExecutor(max_workers=len(insts)) as pool:  # ctypes releases the GIL
        res = list(pool.map(run, insts))
    clumps = sum(len(i.particles) for i in insts)
    return {"value": sum(w / t for w, t in res), "unit": "clump-steps/s", "cores": len(insts),
            "kind": "port",
            "sample": f"oracle, the first {len(insts)} of the 64 batch instances ({clumps} clumps), "
                      f"one per host thread, about {budget_s:.0f} s of steps each"}


def batch_compare(steps: int, warmup: int = 20, skin: float = DEFAULT_SKIN,
                  sequential: bool = True, reorder: int = DEFAULT_REORDER) -> dict:
    """64 independent flumes (N ~ log-uniform[5e3, 6e4], per-instance seeds, mu, eps and
    slope) stepped together in one context vs the same instances as sequential plain runs
    (one context each, synchronised between instances)."""
    from avalancheflume3d_amd import dem, workloads
    insts = workloads.batch(64)
    total = sum(len(i.particles) for i in insts)
    opt = dem.default_options()
    opt.verlet_skin = skin
    opt.reorder_interval = reorder
    ctx = dem.Context(insts, opt)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    t_batch = time.perf_counter() - t0
    ctx.close()
    if not sequential:
        return {"instances": len(insts), "clumps": total, "steps": steps, "skin_cm": skin,
                "batched_clump_steps_per_s": total * steps / t_batch}
    t_seq = 0.0
    for inst in insts:
        c = dem.Context([inst], opt)
        c.regrid()
        c.initial_forces()
        c.step(warmup)
        c.sync()
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        t_seq += time.perf_counter() - t0
        c.close()
    return {"instances": len(insts), "clumps": total, "steps": steps, "skin_cm": skin,
            "batched_clump_steps_per_s": total * steps / t_batch,
            "sequential_clump_steps_per_s": total * steps / t_seq,
            "speedup": t_seq / t_batch}


def adjoint_bench(steps: int, checkpoint: int = 100, count: int = 16, n: int = 20_000,
                  pre: int = 8000, skin: float = 0.1) -> dict:
    """Differentiable window: count slope releases of n clumps (35 deg, different seeds)
    after `pre` plain steps (the layer settles onto the floor and starts to slide), then
    a steps-long window with checkpoints every `checkpoint` steps.  Loss = mean runout +
    L2 misfit of a 64-bin deposit profile against the profile the forward window gives at
    mu = 0.45, eps = 0.55; gradients wrt (mu, eps, kappa) for both contact laws and the
    per-clump start velocities (dv ~ N(0, 1 cm/s)).  Cost = reverse pass (segment
    recompute + adjoint steps) / forward window, both from HIP events inside the library.
    Checks: the dual-number tangent of the same window along each parameter (same scale),
    and central differences of the forward window on a 2,000-clump, 200-step window."""
    import numpy as np

    from avalancheflume3d_amd import abi, dem, workloads
    insts = workloads.adjoint_batch(count, n)
    ctx = dem.Context(insts)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(pre)
    ctx.sync()
    ps = abi.ProfileSpec()
    ps.x0, ps.x1, ps.width, ps.sigma, ps.nbins = 0.0, 40.0, 10.0, 0.0, 64
    s0 = insts[0].sys
    base = np.array([[s0.mu_pp, s0.eps_pp, s0.kappa_pp, s0.mu_pw, s0.eps_pw, s0.kappa_pw]] * count)
    tgt_theta = base.copy()
    tgt_theta[:, [0, 3]] = 0.45
    tgt_theta[:, [1, 4]] = 0.55
    kw = dict(checkpoint=checkpoint, skin=skin, w_runout=1.0, w_profile=1.0, profile=ps)
    target = ctx.window_run(steps, theta=tgt_theta, grad=False, **kw)["profile"]
    dv = np.random.default_rng(5).normal(scale=1.0, size=(count * n, 3))
    out = ctx.window_run(steps, theta=base, dv=dv, target=target, **kw)
    d = out["diag"]
    # dual-number tangent along each parameter (all instances at once)
    tangent = {}
    for k in (1, 2, 3, 4, 5):
        e = np.zeros_like(base)
        e[:, k] = base[:, k] * 1e-2
        o2 = ctx.window_run(steps, theta=base, dv=dv, target=target, grad=False, tangent=True,
                            tan_theta=e, **kw)
        want = out["grad_theta"][:, k] * e[:, k]
        got = o2["dloss"]
        rel = np.abs(got - want) / np.maximum(np.maximum(np.abs(got), np.abs(want)), 1e-300)
        tangent[abi.THETA[k]] = {"adjoint_inst0": float(want[0]), "dual_inst0": float(got[0]),
                                 "max_rel": float(rel.max())}
    # the plain forward over the same number of steps
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    plain_ms = 1e3 * (time.perf_counter() - t0)
    ctx.close()
    return {"instances": count, "clumps": count * n, "window": steps,
            "checkpoint": checkpoint, "skin_cm": skin, "forward_ms": d["forward_ms"],
            "backward_ms": d["backward_ms"], "cost_ratio": d["backward_ms"] / d["forward_ms"],
            "plain_forward_ms": plain_ms,
            "cost_ratio_vs_plain": d["backward_ms"] / plain_ms,
            "list_valid": d["list_valid"], "missed_pairs": d["missed"],
            "max_disp_cm": d["max_disp"], "max_list": d["max_list"],
            "loss_mean": float(np.mean(out["loss"])),
            "grad_theta_inst0": out["grad_theta"][0].tolist(),
            "grad_dv_rms": float(np.sqrt(np.mean(out["grad_dv"] ** 2))),
            "tangent": tangent, "fd": adjoint_fd_check()}


def adjoint_fd_check(steps: int = 300) -> dict:
    """Central differences of the GPU forward window against the adjoint, per parameter, on
    two windows whose loss resolves every parameter that acts: two packets of clumps
    colliding head-on in free space (eps_pp, kappa_pp) and a packet hitting the floor at an
    angle (mu_pw, eps_pw, kappa_pw and its own pp contacts).  Step sizes 1e-6, 1e-5, 1e-4
    relative; mu_pp is inactive under the host friction rule (gradient and differences 0)."""
    import numpy as np

    from avalancheflume3d_amd import abi, dem, workloads
    insts = [workloads.colliding_packets(1), workloads.impacting_packet(2)]
    ctx = dem.Context(insts)
    ctx.regrid()
    ctx.initial_forces()
    ps = abi.ProfileSpec()
    ps.x0, ps.x1, ps.width, ps.sigma, ps.nbins = -4.0, 4.0, 2.0, 0.0, 32
    th = np.array([[i.sys.mu_pp, i.sys.eps_pp, i.sys.kappa_pp, i.sys.mu_pw, i.sys.eps_pw,
                    i.sys.kappa_pw] for i in insts])
    kw = dict(checkpoint=50, skin=0.1, w_runout=1.0, w_profile=10.0, profile=ps,
              target=np.zeros((2, ps.nbins)))
    out = ctx.window_run(steps, theta=th, **kw)
    res = {"window": steps, "clumps": [len(i.particles) for i in insts],
           "list_valid": out["diag"]["list_valid"]}
    worst = 0.0
    for i, name in enumerate(("collision", "impact")):
        for k in range(6):
            g = float(out["grad_theta"][i, k])
            row = {"adjoint": g}
            best = None
            for hr in (1e-6, 1e-5, 1e-4):
                tp, tm = th.copy(), th.copy()
                h = hr * th[i, k]
                tp[i, k] += h
                tm[i, k] -= h
                lp = ctx.window_run(steps, theta=tp, grad=False, **kw)["loss"][i]
                lm = ctx.window_run(steps, theta=tm, grad=False, **kw)["loss"][i]
                f = (lp - lm) / (2 * h)
                row[f"fd_h{hr:.0e}"] = f
                err = abs(f - g)
                best = err if best is None else min(best, err)
            if g != 0.0:
                row["best_rel"] = best / abs(g)
                worst = max(worst, row["best_rel"])
            res[f"{name}:{abi.THETA[k]}"] = row
    res["worst_best_rel"] = worst
    ctx.close()
    return res


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_window(inst, n_sample: int, steps: int) -> dict:
    """The oracle on one host core (the calling thread pinned to the first core it may run
    on): the first n_sample clumps of the workload (its lowest layers), the initial force
    pass, then a timed window of `steps` steps."""
    from oracle import oracle
    pp = inst.particles[:n_sample].copy()
    sys_ = type(inst.sys).from_buffer_copy(inst.sys)
    sys_.npart = len(pp)
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        o = oracle.OracleSim(sys_, inst.geo, inst.fun, inst.ntags, inst.phase, pp)
        o.initial()
        t0 = time.perf_counter()
        o.step(steps)
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    return {"value": len(pp) * steps / dt, "unit": "clump-steps/s", "cores": 1,
            "kind": "port", "cpu": cpu_model(),
            "sample": f"oracle (gcc -O2 C restatement of the reference host path) pinned to 1 "
                      f"core: the first {len(pp)} of the workload's {len(inst.particles)} "
                      f"clumps, a {steps}-step window after the initial force pass "
                      f"({dt:.1f} s)"}


def kernel_bytes(name: str, n: int, diag: dict, binning: str = "clump") -> float:
    """Algorithmic bytes per launch of the kernels a roofline is quoted for."""
    mi = MASS_INERTIA_BYTES if diag["closed_instances"] else 0.0
    if name == "force":
        return ((CLUMP_FORCE_BYTES if binning == "clump" else CELLS_FORCE_BYTES) - mi) * n
    if name == "detect":
        return (CSWEEP_BYTES_PER_CLUMP if binning == "clump" else SWEEP_BYTES_PER_CLUMP) * n
    raise ValueError(name)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="headline", choices=["headline", "small", "large"])
    ap.add_argument("--cpu-clumps", type=int, default=25000,
                    help="clumps of the workload in the CPU baseline's sample")
    ap.add_argument("--cpu-steps", type=int, default=200, help="steps of the CPU window")
    ap.add_argument("--binning", default="clump", choices=["clump", "sphere"],
                    help="per-step binning: one key per clump in clump cells (default) or four "
                         "per clump in sphere cells (same results)")
    ap.add_argument("--compare-steps", type=int, default=2000,
                    help="steps of the same workload binned the other way (0: skip)")
    ap.add_argument("--reorder", type=int, default=50,
                    help="spatial reorder interval in steps (the headline configuration: 50)")
    ap.add_argument("--list-steps", type=int, default=20000,
                    help="steps of the candidate-list comparison run (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-only", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched run of the ensemble (no sequential comparison)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device-to-device copy kernel of this many bytes 5 times "
                         "(calibrates the HBM byte counters against a known traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding of "
              "this path is a design only, not implemented", file=sys.stderr)
        return 2

    import torch

    if args.copy_probe:
        # a device-to-device copy of known size: calibrates the byte counters, and its rate is
        # the practical read + write bandwidth a streaming kernel of that size reaches
        x = torch.ones(args.copy_probe // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for k in range(5):
            if k == 1:
                t0.record()
            y.copy_(x)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / 4
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "launches": 5, "ms_per_copy": ms,
                          "read_plus_write_gbs": 2 * args.copy_probe / (ms * 1e-3) / 1e9}),
              flush=True)
        return 0

    from avalancheflume3d_amd import dem, workloads

    if args.batch_only:
        print(json.dumps({"batched": batch_compare(args.batch_steps, sequential=False)}),
              flush=True)
        return 0

    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench(args.adjoint_steps)}), flush=True)
        return 0

    make = {"headline": workloads.headline, "small": workloads.small, "large": workloads.large}
    inst = make[args.workload]()
    n = len(inst.particles)
    # the stated configuration: the cell list rebuilt every step, reorder every 50 steps
    opt = dem.default_options()
    opt.verlet_skin = 0.0
    opt.reorder_interval = args.reorder
    opt.step_binning = 0 if args.binning == "clump" else 1
    ctx = dem.Context([inst], opt)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(args.warmup)
    # per-kernel breakdown (events between all kernels, a few us each): names the dominant
    # kernel of this path, whose launches the timed region brackets
    kern, _ = ctx.step_profiled(min(args.steps, 400))
    dominant = max(("force", "detect"), key=lambda k: kern[k])
    ctx.sync()
    torch.cuda.synchronize()
    # the state after the warmup: the graph-path run below replays the same segment
    t0 = time.perf_counter()
    snap = ctx.download()
    ctx.sync()
    download_s = time.perf_counter() - t0
    t_snap, cc_snap = ctx.time()
    # timed region: K steps launched on the library's stream, HIP events around the dominant
    # kernel on every TIMED_EVERY-th step
    t0 = time.perf_counter()
    kernel_ms, total_ms = ctx.step_timed(args.steps, dominant, TIMED_EVERY)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    diag = ctx.diag()
    if diag["neigh_overflow"] or diag["cell_below"]:
        raise RuntimeError(f"contact list overflow or clump below the grid: {diag}")
    # the same steps from the same state, replayed as captured hipGraphs (the production path)
    t0 = time.perf_counter()
    ctx.upload(0, *snap)
    ctx.sync()
    upload_s = time.perf_counter() - t0
    ctx.set_phase(0, inst.phase, t_snap, inst.sys.side.x, cc_snap)
    ctx.regrid()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.sync()
    torch.cuda.synchronize()
    wall_graph = time.perf_counter() - t0
    ctx.close()
    value = n * args.steps / wall
    kb = kernel_bytes(dominant, n, diag, args.binning)
    achieved = kb / (kernel_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get(args.workload, {}).get(f"{dominant}_bytes_per_launch")
    step_gbs = STEP_BYTES_PER_CLUMP * value / 1e9
    out = {
        "metric": METRIC, "value": value, "unit": "clump-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} clumps ({4 * n} spheres), FP64, cell "
                               f"list rebuilt every step ({args.binning} cells), spatial "
                               f"reorder every {opt.reorder_interval} steps",
                   "clumps": n, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": KERNEL_NAMES[args.binning][dominant], "kernel_ms": kernel_ms,
                     "bytes_per_launch": kb},
        # the whole step at 480 B per clump-step (algorithmic), against the HBM peak
        "roofline_step": {"achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS,
                          "bytes_per_clump_step": STEP_BYTES_PER_CLUMP},
        "kernel_ms_breakdown": {k: v for k, v in kern.items() if v > 0},
        "events_total_ms": total_ms,
        "graph_path_clump_steps_per_s": n * args.steps / wall_graph,
        # the boundary's host hand-over (reference particle records, 408 B, plus the two
        # acceleration arrays, 48 B, per clump, from and to pageable host memory) and the rate
        # of a job that uploads the state, runs the timed steps and downloads it
        "host_transfer": {"bytes": 456 * n, "upload_ms": 1e3 * upload_s,
                          "download_ms": 1e3 * download_s,
                          "pcie_inclusive_clump_steps_per_s":
                              n * args.steps / (upload_s + wall + download_s)},
        "diag": diag,
    }
    if args.compare_steps > 0:
        other = "sphere" if args.binning == "clump" else "clump"
        out[f"{other}_binning"] = binning_run(inst, other, args.compare_steps, args.warmup,
                                              args.reorder)
    if args.list_steps > 0:
        out["candidate_lists"] = list_run(inst, args.list_steps, args.warmup)
    if args.batch_steps > 0:
        out["batched"] = batch_compare(args.batch_steps)
    if args.adjoint_steps > 0:
        out["adjoint"] = adjoint_bench(args.adjoint_steps)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_window(inst, min(args.cpu_clumps, n), args.cpu_steps)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
        if "batched" in out:  # the batch against the oracle on all cores (16: one GPU's share)
            cb = cpu_batch_baseline(6.0, min(16, os.cpu_count() or 1))
            out["batched"]["cpu_all_cores"] = cb
            out["batched"]["speedup_vs_cpu_all_cores"] = (
                out["batched"]["batched_clump_steps_per_s"] / cb["value"])
    print(json.dumps(out), flush=True)
    return 0


def binning_run(inst, binning: str, steps: int, warmup: int, reorder: int) -> dict:
    """The same workload and steps with the cell list binned the other way (identical
    results): its throughput and per-kernel breakdown."""
    from avalancheflume3d_amd import dem
    n = len(inst.particles)
    opt = dem.default_options()
    opt.reorder_interval = reorder
    opt.step_binning = 0 if binning == "clump" else 1
    ctx = dem.Context([inst], opt)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(warmup)
    kern, _ = ctx.step_profiled(min(steps, 400))
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.close()
    return {"workload": f"the same {n} clumps, cell list rebuilt every step ({binning} cells), "
                        f"spatial reorder every {reorder} steps",
            "clump_steps_per_s": n * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "steps": steps, "kernel_ms_breakdown": {k: v for k, v in kern.items() if v > 0}}


def list_run(inst, steps: int, warmup: int) -> dict:
    """The same workload on candidate lists (skin 0.02 cm, reorder every 400 steps; the
    library's per-step-binning results to the bit), with its k_force roofline (events around
    runs of consecutive list-step launches)."""
    from avalancheflume3d_amd import dem
    n = len(inst.particles)
    opt = dem.default_options()
    opt.verlet_skin = DEFAULT_SKIN
    opt.reorder_interval = DEFAULT_REORDER
    ctx = dem.Context([inst], opt)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    force_ms, _ = ctx.step_timed(steps, "force", TIMED_EVERY)
    ctx.sync()
    wall = time.perf_counter() - t0
    diag = ctx.diag()
    ctx.close()
    bpc = FORCE_BYTES_PER_CLUMP - (MASS_INERTIA_BYTES if diag["closed_instances"] else 0.0)
    ach = bpc * n / (force_ms * 1e-3) / 1e9
    return {"workload": f"the same {n} clumps on candidate lists within sep + {DEFAULT_SKIN} cm, "
                        f"spatial reorder every {DEFAULT_REORDER} steps",
            "clump_steps_per_s": n * steps / wall, "steps": steps,
            "k_force_roofline": {"achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": ach / HBM_PEAK_GBS, "kernel_ms": force_ms,
                                 "bytes_per_clump": bpc},
            "list_builds": diag["list_builds"], "steps_undone": diag["steps_undone"]}


if __name__ == "__main__":
    sys.exit(main())
