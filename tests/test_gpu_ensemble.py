"""Long runs: the seed ensemble of the GPU path against the host oracle.

Trajectories are chaotic, so after many steps runs are compared through the statistics the
reference's users read off a run: the runout (mass-weighted mean slope coordinate x' of the
clumps, slope frame of rotate_vectors.cu:7-56), the front, and the deposit height profile
over x'.  Six seeds of the same small flume run 20,000 steps (0.29 s of flow: the layer
leaves the release box, runs down the slope and piles onto the deposit) on the GPU in one
batched context on candidate lists, and on the oracle (one host thread per seed).  The
ensemble means of the GPU runs must lie within the seed-ensemble spread of the oracle runs.
"""
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from avalancheflume3d_amd import dem, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu

SEEDS = (1, 2, 3, 4, 5, 6)
STEPS = 20_000
BINS = np.linspace(0.0, 10.0, 21)   # x' bins (cm) over the slope end and the deposit


def instance(seed: int) -> dem.Instance:
    # contacts 100x softer than the reference stiffness: a 10x larger step, so the flow
    # develops within a run the oracle finishes in seconds; 24 clumps per clump cell
    spec = workloads.FlumeSpec(n=120, width=2.0, length=3.0, box=(2.0, 2.0, 2.01),
                               deposit_length=6.0, spacing=0.4, jitter=0.02, seed=seed,
                               k_n=workloads.K_N / 100)
    inst = workloads.build(spec)
    inst.ntags = 24
    return inst


def statistics(pp: np.ndarray, inst: dem.Instance) -> dict:
    s, c = inst.sys.sint, inst.sys.cost
    xs = pp["rr"][:, 0] * c - pp["rr"][:, 2] * s
    m = pp["mass"]
    vol = m / pp["density"]
    hist, _ = np.histogram(xs, bins=BINS, weights=vol)
    return {"runout": float((m * xs).sum() / m.sum()), "front": float(xs.max()),
            "profile": hist / (np.diff(BINS) * inst.sys.side.y)}


def oracle_run(inst: dem.Instance) -> dict:
    o = oracle.OracleSim(inst.sys, inst.geo, inst.fun, inst.ntags, inst.phase, inst.particles)
    o.initial()
    o.step(STEPS)
    return statistics(o.get()["pp"], inst)


def test_long_run_statistics_within_seed_spread(gpu):
    insts = [instance(s) for s in SEEDS]
    with ThreadPoolExecutor(max_workers=len(insts)) as pool:   # ctypes releases the GIL
        ref = list(pool.map(oracle_run, insts))
    o = dem.default_options()
    o.verlet_skin = 0.01
    ctx = dem.Context(insts, o)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(STEPS)
    got = [statistics(ctx.download(k)[0], inst) for k, inst in enumerate(insts)]
    ctx.close()
    for key in ("runout", "front"):
        a = np.array([g[key] for g in got])
        b = np.array([r[key] for r in ref])
        spread = b.std(ddof=1)
        print(f"{key}: GPU {a.mean():.4f} +- {a.std(ddof=1):.4f}, oracle {b.mean():.4f} "
              f"+- {spread:.4f}; largest per-seed difference {np.abs(a - b).max():.3e}")
        assert spread > 0.0
        assert abs(a.mean() - b.mean()) <= spread / math.sqrt(len(SEEDS))
    pa = np.array([g["profile"] for g in got])
    pb = np.array([r["profile"] for r in ref])
    diff = np.linalg.norm(pa.mean(0) - pb.mean(0))
    spread = math.sqrt((pb.std(0, ddof=1) ** 2).sum())
    print(f"deposit profile: |mean difference| {diff:.3e}, oracle seed spread {spread:.3e}")
    assert spread > 0.0
    assert diff <= spread / math.sqrt(len(SEEDS))


def test_batched_ensemble_configuration_equals_plain_runs(gpu):
    """The batched-ensemble configuration: 64 flumes (clumps log-uniform in [5,000, 60,000],
    per-instance seeds, friction and restitution U[0.3, 0.7], slope U[28, 40] deg) stepped
    10,000 steps together in one context, against each instance run alone: per-instance
    runout, front, mean speeds and 64-bin deposit profile bit for bit, and the whole state.
    Both on candidate lists (skin 0.02 cm, reorder every 400 steps), whose results equal
    per-step binning bit for bit (tests/test_gpu_parity.py)."""
    from avalancheflume3d_amd import abi
    insts = workloads.batch(64)
    assert len(insts) == 64
    o = dem.default_options()
    o.verlet_skin, o.reorder_interval = 0.02, 400
    steps = 10_000
    ctx = dem.Context(insts, o)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(steps)
    assert ctx.diag()["neigh_overflow"] == 0
    ps = abi.ProfileSpec()
    ps.x0, ps.x1, ps.width, ps.sigma, ps.nbins = -2.0, 30.0, 4.0, 0.0, 64
    batched = [(ctx.observe(k, ps), ctx.download(k)[0]) for k in range(len(insts))]
    ctx.close()
    moved = 0
    for k, inst in enumerate(insts):
        one = dem.Context([inst], o)
        one.regrid()
        one.initial_forces()
        one.step(steps)
        (obs, prof), pp = one.observe(0, ps), one.download(0)[0]
        one.close()
        (bobs, bprof), bpp = batched[k]
        assert obs == bobs, k  # every field bit for bit (floats compared exactly)
        assert np.array_equal(prof.view(np.uint64), bprof.view(np.uint64)), k
        for fld in ("rr", "vv", "ww", "Rmatx", "Rmaty", "Rmatz", "radius"):
            assert np.array_equal(pp[fld].view(np.uint64), bpp[fld].view(np.uint64)), (k, fld)
        moved += obs["vmean"] > 0
    assert moved == len(insts)
