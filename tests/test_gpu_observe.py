"""Observables, cell checks and the autotuner on the GPU.

* avf_observe against closed forms on a synthetic deposit: the reference statistics
  (mean |v|, mean |w|, centre of mass, avalanche_hst.cu:551-574, 819-829), the slope-frame
  runout, and the Gaussian deposit profile, whose integral over the bins equals the deposited
  volume; clumps past side.x are excluded once the gate is open.
* avf_check_cells reports the reference's fatal cell conditions (cells.cu:79-102) with the
  offending clump instead of exiting.
* The autotuner's choice changes no bit of the result.
"""
import math

import numpy as np
import pytest

from avalancheflume3d_amd import abi, autotune, dem, workloads
from tests import scenarios

pytestmark = pytest.mark.gpu


def deposit_state(seed=3, n=2000):
    inst = workloads.small(seed)
    pp = inst.particles.copy()
    rng = np.random.default_rng(seed)
    s, c = inst.sys.sint, inst.sys.cost
    xs = rng.uniform(2.0, 12.0, n)          # slope coordinate x'
    zs = rng.uniform(0.2, 1.0, n)           # height above the floor
    pp["rr"][:, 0] = xs * c + zs * s        # slope frame -> lab (rotate_vectors inverse)
    pp["rr"][:, 2] = zs * c - xs * s
    pp["rr"][:, 1] = rng.uniform(0.3, 3.7, n)
    pp["vv"] = rng.normal(size=(n, 3))
    pp["ww"] = rng.normal(size=(n, 3))
    dem.sphere_centres(pp, inst.geo)
    inst.particles = pp
    return inst


def test_observables_closed_forms(gpu):
    inst = deposit_state()
    pp = inst.particles
    ctx = dem.Context([inst])
    spec = abi.ProfileSpec()
    spec.x0, spec.x1, spec.width, spec.sigma, spec.nbins = -2.0, 16.0, 4.0, 0.0, 144
    obs, prof = ctx.observe(0, spec)
    s, c = inst.sys.sint, inst.sys.cost
    x, z = pp["rr"][:, 0], pp["rr"][:, 2]
    xs = x * c - z * s
    m = pp["mass"]
    vol = m / pp["density"]
    assert obs["live"] == len(pp)
    assert math.isclose(obs["runout_mean"], (m * xs).sum() / m.sum(), rel_tol=1e-12)
    assert obs["runout_front"] == xs.max()
    assert math.isclose(obs["vmean"], np.linalg.norm(pp["vv"], axis=1).mean(), rel_tol=1e-12)
    assert math.isclose(obs["wmean"], np.linalg.norm(pp["ww"], axis=1).mean(), rel_tol=1e-12)
    assert math.isclose(obs["xcm"], (m * x).sum() / inst.sys.tot_mass, rel_tol=1e-12)
    assert math.isclose(obs["zcm"], (m * z).sum() / inst.sys.tot_mass, rel_tol=1e-12)
    bw = (spec.x1 - spec.x0) / spec.nbins
    cb = spec.x0 + (np.arange(spec.nbins) + 0.5) * bw
    u = xs[None, :] - cb[:, None]
    g = np.where(np.abs(u) < 8 * bw, np.exp(-0.5 * (u / bw) ** 2) / (bw * math.sqrt(2 * math.pi)), 0)
    want = (vol[None, :] * g).sum(1) / spec.width
    assert np.allclose(prof, want, rtol=1e-12, atol=1e-15)
    # the profile integrates to the deposited volume (bins cover the support)
    assert math.isclose(prof.sum() * bw * spec.width, vol.sum(), rel_tol=1e-6)
    # gate open: clumps past side.x leave the statistics of the live material
    cut = np.quantile(x, 0.7)
    ctx.set_phase(0, abi.PHASE_AVALANCHE, 0.5, cut)
    obs2 = ctx.observe(0)
    live = x <= cut
    assert obs2["live"] == live.sum()
    assert math.isclose(obs2["mass_live"], m[live].sum(), rel_tol=1e-12)
    assert math.isclose(obs2["runout_mean"], (m[live] * xs[live]).sum() / m[live].sum(),
                        rel_tol=1e-12)
    ctx.close()


def test_check_cells_reports_reference_fatal_conditions(gpu):
    inst = scenarios.hopper_state(500, 71)
    ctx = dem.Context([inst])
    ctx.regrid()
    assert ctx.check_cells() == (0, -1)
    cells = ctx.clump_cells()
    ctx.close()
    # ntags = 2: the first clump that finds its cell already holding ntags - 1 = 1 clump
    tight = dem.Instance(sys=inst.sys, geo=inst.geo, fun=inst.fun, particles=inst.particles,
                         ntags=2, phase=inst.phase)
    ctx = dem.Context([tight])
    ctx.regrid()
    seen, want = set(), None
    for k, cl in enumerate(cells):
        if cl < 0:
            continue
        if cl in seen:
            want = k
            break
        seen.add(cl)
    assert want is not None
    assert ctx.check_cells() == (abi.STATUS_CODES["AVF_ERR_CELL_FULL"], want)
    ctx.close()
    # a clump moved below the clump-cell grid after the regrid (cells.cu:79-88)
    ctx = dem.Context([inst])
    ctx.regrid()
    pp = inst.particles.copy()
    pp["rr"][17, 2] -= 500.0
    dem.sphere_centres(pp, inst.geo)
    ctx.upload(0, pp)
    assert ctx.check_cells() == (abi.STATUS_CODES["AVF_ERR_CELL_NEG"], 17)
    ctx.close()


def test_autotuner_choice_changes_no_bit(gpu):
    inst = scenarios.avalanche_state(12, 81)
    best, table = autotune.autotune([inst], steps=30, warmup=5, skins=(0.0, 0.01),
                                    factors=(1.0,), reorders=(50,), blocks=(64, 128))
    assert len(table) == 4 and best is not None
    outs = []
    for o in (dem.default_options(), best):
        ctx = dem.Context([inst], o)
        ctx.regrid()
        ctx.initial_forces()
        ctx.step(60)
        outs.append(ctx.download())
        ctx.close()
    for fld in ("rr", "vv", "ww", "Rmatx", "Rmaty", "Rmatz"):
        assert np.array_equal(outs[0][0][fld].view(np.uint64), outs[1][0][fld].view(np.uint64))
