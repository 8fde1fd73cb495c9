"""GPU test of avf_observe_noalloc, the entry point that allocates no device memory: through
the C ABI it returns avf_observe's observables and deposit profile bit for bit, on two calls
in a row (the second on a smaller instance, the scratch array keeping the first call's
contents), and refuses an instance whose partial sums exceed its fixed scratch."""
import numpy as np
import pytest

from avalancheflume3d_amd import abi, dem, workloads
from avalancheflume3d_amd.abi import ptr
from tests import scenarios

pytestmark = pytest.mark.gpu


def spec_of(nbins):
    s = abi.ProfileSpec()
    s.x0, s.x1, s.width, s.sigma, s.nbins = -2.0, 30.0, 4.0, 0.0, nbins
    return s


def observe(ctx, fn, inst, spec):
    o = abi.Observables()
    prof = np.zeros(spec.nbins, np.float64)
    rc = getattr(dem.lib(), fn)(ctx.h, inst, ptr(o), ptr(spec), ptr(prof))
    return rc, {k: getattr(o, k) for k, _ in abi.Observables._fields_}, prof


def same(a, b):
    return all(np.float64(a[k]).tobytes() == np.float64(b[k]).tobytes() for k in a)


def run_context(inst, steps):
    ctx = dem.Context([inst])
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(steps)
    return ctx


def test_noalloc_observe_equals_observe(gpu):
    big = workloads.build(workloads.FlumeSpec(n=6000, width=4.0, length=20.0,
                                              box=(4.0, 4.0, 0.4 * 60 + 0.01), spacing=0.4,
                                              jitter=0.02, seed=11))
    small = scenarios.avalanche_state(16, 12)
    cases = [(big, 200, 96), (small, 30, 40)]  # the second call: fewer clumps and bins
    ctxs = [run_context(inst, steps) for inst, steps, _ in cases]
    got = []
    for ctx, (inst, _, nbins) in zip(ctxs, cases):
        rc, o, prof = observe(ctx, "avf_observe_noalloc", 0, spec_of(nbins))
        assert rc == 0, dem.lib().avf_last_error(ctx.h)
        got.append((o, prof))
    for ctx, (inst, _, nbins), (o, prof) in zip(ctxs, cases, got):
        rc, want, wprof = observe(ctx, "avf_observe", 0, spec_of(nbins))
        assert rc == 0
        assert same(o, want), (o, want)
        assert prof.tobytes() == wprof.tobytes()
        assert 0 < o["live"] <= len(inst.particles) and o["mass_live"] > 0.0  # some exited
        print("observables", o)
    # again after avf_observe (the scratch holds other contents now), then without a profile
    rc, o2, prof2 = observe(ctxs[0], "avf_observe_noalloc", 0, spec_of(cases[0][2]))
    assert rc == 0 and same(o2, got[0][0]) and prof2.tobytes() == got[0][1].tobytes()
    o3 = abi.Observables()
    assert dem.lib().avf_observe_noalloc(ctxs[1].h, 0, ptr(o3), None, None) == 0
    assert np.float64(o3.runout_mean).tobytes() == np.float64(got[1][0]["runout_mean"]).tobytes()
    for ctx in ctxs:
        ctx.close()


def test_noalloc_observe_refuses_oversized_partials(gpu):
    # 300,000 clumps with 1024 bins: (1172 + 1) * 1033 doubles of partial sums > 2^20
    inst = workloads.headline(n=300_000)
    ctx = dem.Context([inst])
    rc, _, _ = observe(ctx, "avf_observe_noalloc", 0, spec_of(1024))
    assert rc == abi.STATUS_CODES["AVF_ERR_NOMEM"]
    assert b"scratch" in dem.lib().avf_last_error(ctx.h)
    rc, o, _ = observe(ctx, "avf_observe_noalloc", 0, spec_of(64))  # within the scratch
    assert rc == 0 and o["live"] == 300_000
    ctx.close()
