"""The contact sweep's rare paths, against the oracle:

* sphere pairs within the float test's margin of sep (|d - sep| below 2^-16 cell sides),
  which the sweep decides with force_part's FP64 test: clump pairs placed so that their
  closest sphere pair sits at sep (1 + delta) for delta from 1e-12 to 1e-4 either side;
* brick halos holding more spheres than the sweep stages in LDS, swept from global memory:
  lattices of clumps packed closer than their diameter.

Forces and torques are compared with the oracle (within 1e-12, bit-identical in practice):
a touching pair the sweep missed, or a listed pair that does not touch, changes them.
"""
import numpy as np
import pytest

from avalancheflume3d_amd import dem, workloads
from oracle import oracle
from tests.test_gpu_parity import F_RTOL, gpu_ctx, oracle_sim, vec_close

pytestmark = pytest.mark.gpu

DELTAS = (-1e-4, -1e-6, -1e-9, -1e-12, 1e-12, 1e-9, 1e-6, 1e-4)


def closest_pair_d2(pp, a, b):
    """Squared distance of the closest sphere pair of clumps a[k], b[k], in force_part's
    operation order ((dx^2 + dy^2) + dz^2)."""
    keys = ("rr_a", "rr_b", "rr_c", "rr_d")
    best = np.full(len(a), np.inf)
    for ka in keys:
        for kb in keys:
            d = pp[kb][b] - pp[ka][a]
            d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
            best = np.minimum(best, d2)
    return best


def margin_pairs(seed, per_delta=6):
    """Clump pairs in free space whose closest sphere pair is at sep (1 + delta)."""
    deltas = np.repeat(np.array(DELTAS), per_delta)
    m = len(deltas)
    inst = workloads.colliding_packets(seed, per=(m, 1, 1), a=0.4, v=0.0)
    pp = inst.particles
    rng = np.random.default_rng(seed)
    a, b = np.arange(0, 2 * m, 2), np.arange(1, 2 * m, 2)
    g = np.stack(np.meshgrid(np.arange(8), np.arange(4), np.arange(4), indexing="ij"), -1)
    sites = g.reshape(-1, 3)[:m] * 1.2 + np.array([0.6, 8.0, 5.0])
    pp["rr"][a] = sites
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    sep = pp["radius"][a] + pp["radius"][b]
    target2 = (sep * (1.0 + deltas)) ** 2
    lo, hi = np.full(m, 0.02), np.full(m, 1.0)
    for _ in range(100):  # bisection on the offset of clump b along u
        t = 0.5 * (lo + hi)
        pp["rr"][b] = pp["rr"][a] + u * t[:, None]
        dem.sphere_centres(pp, inst.geo)
        below = closest_pair_d2(pp, a, b) < target2
        lo = np.where(below, t, lo)
        hi = np.where(below, hi, t)
    pp["rr"][b] = pp["rr"][a] + u * hi[:, None]
    dem.sphere_centres(pp, inst.geo)
    touching = closest_pair_d2(pp, a, b) < sep * sep
    return inst, deltas, touching


def test_pairs_at_the_float_margin_of_sep(gpu):
    inst, deltas, touching = margin_pairs(31)
    # the placement reproduces the intended side of sep for every pair
    assert np.array_equal(touching, deltas < 0)
    brute = oracle.brute_contacts(inst.particles, inst.sys)
    assert len(set(zip(brute["l"] // 2, brute["m"] // 2))) == int(touching.sum())
    ctx = gpu_ctx(inst)
    ctx.initial_forces()
    f, t = ctx.forces()
    o = oracle_sim(inst)
    o.initial()
    r = o.get()
    hit = np.linalg.norm(r["ff"], axis=1) > 0
    assert hit.sum() == 2 * touching.sum()  # both clumps of every touching pair
    exact = [vec_close(f, r["ff"], F_RTOL, "force"), vec_close(t, r["tq"], F_RTOL, "torque")]
    for k in (1, 2):  # per-step binning
        ctx.step(1)
        o.step(1)
        f, t = ctx.forces()
        r = o.get()
        exact += [vec_close(f, r["ff"], F_RTOL, f"force, step {k}"),
                  vec_close(t, r["tq"], F_RTOL, f"torque, step {k}")]
    d = ctx.diag()
    print("FP64-decided pairs", d["sweep_fp64_pairs"], "bit-exact", all(exact))
    assert d["sweep_fp64_pairs"] > 0 and d["neigh_overflow"] == 0
    ctx.close()


@pytest.mark.parametrize("a", [0.26, 0.3])
def test_overfull_halos_swept_from_global_memory(gpu, a):
    inst = workloads.colliding_packets(5, per=(8, 8, 8), a=a, v=20.0)
    ctx = gpu_ctx(inst)
    ctx.initial_forces()
    o = oracle_sim(inst)
    o.initial()
    f, t = ctx.forces()
    r = o.get()
    assert (np.linalg.norm(r["ff"], axis=1) > 0).mean() > 0.9
    exact = [vec_close(f, r["ff"], F_RTOL, "force"), vec_close(t, r["tq"], F_RTOL, "torque")]
    ctx.step(3)
    o.step(3)
    f, t = ctx.forces()
    r = o.get()
    exact += [vec_close(f, r["ff"], F_RTOL, "force after 3 steps"),
              vec_close(t, r["tq"], F_RTOL, "torque after 3 steps")]
    d = ctx.diag()
    print(f"spacing {a}: bricks swept from global memory", d["sweep_global_bricks"],
          "bit-exact", all(exact))
    assert d["sweep_global_bricks"] > 0 and d["neigh_overflow"] == 0
    ctx.close()
