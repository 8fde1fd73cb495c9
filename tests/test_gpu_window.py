"""GPU tests of the differentiable window (avf_window_run).

* The window forward (frozen contact list) reproduces the plain step (avf_step) over the
  window when the list misses no contact.
* Dual-number tangent of the GPU forward vs the hand-written adjoint (dot-product test):
  dloss along a random direction equals grad . direction.
* Gradients vs central finite differences of the host path (the oracle) at 1e-5 relative
  per parameter (pp and pw), the best of h = 1e-6, 1e-5, 1e-4 relative, on windows whose
  loss resolves every parameter that acts.
* A batch of instances gives the same losses and gradients as separate runs.
"""
import math

import numpy as np
import pytest

import torch

from avalancheflume3d_amd import abi, calibrate, dem, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu

PRE_STEPS = 40000   # release -> flowing layer with contacts on the floor, walls and deposit
FD_RTOL = 1e-5


def flowing_state(seed: int, n: int = 240, length: float = 1.6, deposit: float = 6.0):
    """A short 35 deg flume ending on a flat deposit, released from a 2 x 2 cm box and run
    PRE_STEPS steps on the GPU: returns (instance, pp, aa_i, anga_i, time)."""
    spec = workloads.FlumeSpec(n=n, width=2.0, length=length, box=(2.0, 2.0, 0.4 * math.ceil(n / 25) + 0.01),
                               deposit_length=deposit, spacing=0.4, jitter=0.02, seed=seed)
    inst = workloads.build(spec)
    ctx = dem.Context([inst])
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(PRE_STEPS)
    pp, aa, anga = ctx.download()
    t, _ = ctx.time()
    ctx.close()
    return inst, pp, aa, anga, t


def context_at(inst, pp, aa, anga, t, extra=()):
    """Context holding the given state (and further (inst, pp, aa, anga, t) tuples)."""
    items = [(inst, pp, aa, anga, t)] + list(extra)
    insts = []
    for it, p, a, g, tt in items:
        i2 = dem.Instance(sys=abi.System.from_buffer_copy(it.sys), geo=it.geo, fun=it.fun,
                          particles=p, ntags=it.ntags, phase=abi.PHASE_AVALANCHE, aa_i=a,
                          anga_i=g)
        i2.sys.time = tt
        insts.append(i2)
    ctx = dem.Context(insts)
    for k, (it, p, a, g, tt) in enumerate(items):
        ctx.set_phase(k, abi.PHASE_AVALANCHE, tt, it.sys.side.x)
    ctx.regrid()  # the clump-cell grid of the start state, as the oracle builds it
    return ctx


@pytest.fixture(scope="module")
def state():
    return flowing_state(5)


def profile_spec(nbins=16):
    ps = abi.ProfileSpec()
    ps.x0, ps.x1, ps.width, ps.sigma, ps.nbins = -0.5, 5.0, 2.0, 0.0, nbins
    return ps


def host_loss(pp, sys, ps, target, w_r=1.0, w_p=10.0):
    """runout_mean + profile misfit of a final state (the device definitions)."""
    x, z = pp["rr"][:, 0], pp["rr"][:, 2]
    live = ~((sys.time > 0) & (x > sys.side.x))
    m = pp["mass"][live]
    xs = (x * sys.cost - z * sys.sint)[live]
    vol = m / pp["density"][live]
    runout = (m * xs).sum() / m.sum()
    bw = (ps.x1 - ps.x0) / ps.nbins
    sig = ps.sigma if ps.sigma > 0 else bw
    cb = ps.x0 + (np.arange(ps.nbins) + 0.5) * bw
    u = xs[None, :] - cb[:, None]
    g = np.where(np.abs(u) < 8 * sig, np.exp(-0.5 * (u / sig) ** 2) / (sig * math.sqrt(2 * math.pi)), 0.0)
    h = (vol[None, :] * g).sum(1) / ps.width
    return w_r * runout + w_p * ((h - target) ** 2).sum(), h


def gamma_of(eps, kappa):
    L = math.log(eps)
    return -2.0 * L * math.sqrt(kappa / (math.pi ** 2 + L * L))


def oracle_loss(inst, pp, aa, anga, t, steps, theta, dv, ps, target):
    sys = abi.System.from_buffer_copy(inst.sys)
    sys.time = t
    mu, eps, kap = theta[0], theta[1], theta[2]
    sys.mu_pp, sys.kappa_pp, sys.gamma_pp, sys.fr_fact_pp = mu, kap, gamma_of(eps, kap), mu / 1e12
    mu, eps, kap = theta[3], theta[4], theta[5]
    sys.mu_pw, sys.kappa_pw, sys.gamma_pw, sys.fr_fact_pw = mu, kap, gamma_of(eps, kap), mu / 1e12
    p2 = pp.copy()
    p2["vv"] += dv
    o = oracle.OracleSim(sys, inst.geo, inst.fun, inst.ntags, abi.PHASE_AVALANCHE, p2)
    o.set(p2, aa, anga)
    o.step(steps)
    r = o.get()
    return host_loss(r["pp"], r["sys"], ps, target)[0]


def base_theta(sys):
    return np.array([sys.mu_pp, sys.eps_pp, sys.kappa_pp, sys.mu_pw, sys.eps_pw, sys.kappa_pw])


def test_window_forward_matches_plain_step(gpu, state):
    inst, pp, aa, anga, t = state
    steps = 230
    ctx = context_at(inst, pp, aa, anga, t)
    out = ctx.window_run(steps, checkpoint=50, skin=0.05, grad=False, final_state=True,
                         profile=profile_spec())
    d = out["diag"]
    print("window diag", d)
    assert d["max_list"] > 0 and d["list_valid"] == 1
    ctx.step(steps)
    p2, a2, g2 = ctx.download()
    fs = out["final_state"]
    order = {"rr": slice(0, 3), "vv": slice(3, 6), "ww": slice(6, 9)}
    exact = True
    for fld, sl in order.items():
        a, b = fs[:, sl], p2[fld]
        err = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)
        assert err.max() <= 1e-9, (fld, err.max())
        exact &= bool(np.array_equal(a, b))
    R = np.stack([p2["Rmatx"], p2["Rmaty"], p2["Rmatz"]], 1).reshape(-1, 9)
    assert np.abs(fs[:, 9:18] - R).max() <= 1e-9
    assert np.abs(fs[:, 18:21] - a2).max() <= 1e-9 * max(1.0, np.abs(a2).max())
    # the window's observables equal avf_observe of the plain run
    o2, prof2 = ctx.observe(0, profile_spec())
    assert abs(out["obs"][0]["runout_mean"] - o2["runout_mean"]) <= 1e-9 * abs(o2["runout_mean"])
    assert np.allclose(out["profile"][0], prof2, rtol=1e-9, atol=1e-12)
    print("window forward bit-identical to the plain step:", exact)
    ctx.close()


def test_tangent_matches_adjoint(gpu, state):
    inst, pp, aa, anga, t = state
    ps = profile_spec()
    target = np.linspace(0.05, 0.01, ps.nbins)
    ctx = context_at(inst, pp, aa, anga, t)
    rng = np.random.default_rng(3)
    th = base_theta(inst.sys)
    tan_theta = rng.normal(size=6) * th * 0.1
    tan_dv = rng.normal(size=(len(pp), 3))
    out = ctx.window_run(300, checkpoint=64, skin=0.05, w_profile=10.0, profile=ps,
                         target=target, tan_theta=tan_theta, tan_dv=tan_dv, tangent=True)
    lhs = out["dloss"][0]
    rhs = float((out["grad_theta"][0] * tan_theta).sum() + (out["grad_dv"] * tan_dv).sum())
    print("dual tangent", lhs, "adjoint", rhs, "diag", out["diag"])
    assert abs(lhs - rhs) <= 1e-9 * max(abs(lhs), abs(rhs))
    # per-parameter: tangent along each theta alone
    for k in range(6):
        e = np.zeros(6)
        e[k] = th[k] * 1e-2
        o2 = ctx.window_run(300, checkpoint=64, skin=0.05, w_profile=10.0, profile=ps,
                            target=target, tan_theta=e, tangent=True, grad=False)
        want = out["grad_theta"][0][k] * e[k]
        got = o2["dloss"][0]
        print(abi.THETA[k], got, want)
        assert abs(got - want) <= 1e-9 * max(abs(got), abs(want), 1e-30), abi.THETA[k]
    ctx.close()


def fd_agrees(label, g, fds):
    """Central differences at step sizes 1e-6, 1e-5 and 1e-4 (relative to the parameter, or
    in cm/s): the best must agree with the adjoint to FD_RTOL relative to |g|; a component
    the loss does not depend on (g = 0, e.g. mu_pp under the host friction rule) must leave
    the loss unchanged to the last bit at every step size.  The smallest steps reach the
    double-precision noise floor of the loss, the largest its curvature."""
    print(f"{label:10s} adjoint {g:+.12e} " +
          " ".join(f"h={h:.0e}: {fd:+.12e}" for h, fd in fds), end="")
    if g == 0.0:
        print()
        return all(fd == 0.0 for _, fd in fds)
    rel = min(abs(fd - g) for _, fd in fds) / abs(g)
    print(f"  best rel {rel:.2e}")
    return rel <= FD_RTOL


def test_gradients_vs_central_fd_of_host_path(gpu):
    """Adjoint gradients of a 300-step window against central differences of the host path
    (the oracle), per parameter, pp and pw, at 1e-5 relative.  Two windows batched in one
    context, chosen so that every parameter that acts moves the loss far above its noise
    floor: two packets colliding head-on in free space (particle-particle contacts:
    eps_pp, kappa_pp) and a packet hitting the floor at an angle (impacts and sliding
    friction: mu_pw, eps_pw, kappa_pw, plus the packet's own pp contacts).  Loss: mean
    runout + 10 x the squared 32-bin deposit profile; dv ~ N(0, 1 cm/s)."""
    steps = 300
    items = []
    for inst in (workloads.colliding_packets(1), workloads.impacting_packet(2)):
        o = oracle.OracleSim(inst.sys, inst.geo, inst.fun, inst.ntags, inst.phase, inst.particles)
        o.initial()
        g = o.get()
        items.append((inst, g["pp"], g["aa_i"], g["anga_i"], inst.sys.time))
    ps = abi.ProfileSpec()
    ps.x0, ps.x1, ps.width, ps.sigma, ps.nbins = -4.0, 4.0, 2.0, 0.0, 32
    target = np.zeros((2, ps.nbins))
    ctx = context_at(*items[0], extra=[items[1]])
    th = np.stack([base_theta(it[0].sys) for it in items])
    rng = np.random.default_rng(11)
    n0 = len(items[0][1])
    dv = rng.normal(scale=1.0, size=(n0 + len(items[1][1]), 3))
    out = ctx.window_run(steps, checkpoint=50, skin=0.1, w_profile=10.0, profile=ps,
                         target=target, theta=th, dv=dv)
    ctx.close()
    assert out["diag"]["list_valid"] == 1
    ok = True
    for i, (inst, pp, aa, anga, t) in enumerate(items):
        lo = 0 if i == 0 else n0
        dvi = dv[lo:lo + len(pp)]
        L0 = oracle_loss(inst, pp, aa, anga, t, steps, th[i], dvi, ps, target[i])
        print(f"instance {i}: loss gpu {out['loss'][i]} host {L0}")
        assert abs(L0 - out["loss"][i]) <= 1e-9 * abs(L0)
        gth = out["grad_theta"][i]
        for k in range(6):
            fds = []
            for hr in (1e-6, 1e-5, 1e-4):
                h = hr * th[i, k]
                tp, tm = th[i].copy(), th[i].copy()
                tp[k] += h
                tm[k] -= h
                lp = oracle_loss(inst, pp, aa, anga, t, steps, tp, dvi, ps, target[i])
                lm = oracle_loss(inst, pp, aa, anga, t, steps, tm, dvi, ps, target[i])
                fds.append((hr, (lp - lm) / (2 * h)))
            ok &= fd_agrees(f"{i}:{abi.THETA[k]}", gth[k], fds)
        # the parameters that act in each window are resolved (not zero)
        acting = (1, 2) if i == 0 else (1, 2, 3, 4, 5)
        assert all(gth[k] != 0.0 for k in acting)
        gd = out["grad_dv"][lo:lo + len(pp)]
        for flat in np.argsort(-np.abs(gd).ravel())[:3]:
            c, ax = divmod(int(flat), 3)
            fds = []
            for h in (1e-6, 1e-5, 1e-4):  # cm/s
                dp, dm = dvi.copy(), dvi.copy()
                dp[c, ax] += h
                dm[c, ax] -= h
                fds.append((h, (oracle_loss(inst, pp, aa, anga, t, steps, th[i], dp, ps, target[i]) -
                                oracle_loss(inst, pp, aa, anga, t, steps, th[i], dm, ps, target[i])) / (2 * h)))
            ok &= fd_agrees(f"{i}:dv[{c},{ax}]", gd[c, ax], fds)
    assert ok


def test_batch_equals_separate(gpu, state):
    inst, pp, aa, anga, t = state
    other = flowing_state(6, n=180)
    ps = profile_spec()
    ctx = context_at(inst, pp, aa, anga, t, extra=[other])
    both = ctx.window_run(120, checkpoint=40, skin=0.05, w_profile=5.0, profile=ps)
    ctx.close()
    for k, st in enumerate([(inst, pp, aa, anga, t), other]):
        c1 = context_at(*st)
        one = c1.window_run(120, checkpoint=40, skin=0.05, w_profile=5.0, profile=ps)
        c1.close()
        assert both["loss"][k] == one["loss"][0]
        assert np.array_equal(both["grad_theta"][k], one["grad_theta"][0])
        lo = 0 if k == 0 else len(pp)
        assert np.array_equal(both["grad_dv"][lo:lo + len(st[1])], one["grad_dv"])


def test_window_rejects_unsupported_states(gpu, state):
    inst, pp, aa, anga, t = state
    ctx = context_at(inst, pp, aa, anga, t)
    ctx.set_phase(0, abi.PHASE_FILL, t, inst.sys.side.x)
    with pytest.raises(dem.AvfError):
        ctx.window_run(10)
    ctx.close()
    funnel = workloads.build(workloads.FlumeSpec(n=100, funnel_width=6.0, funnel_length=4.0))
    c2 = dem.Context([funnel])
    c2.regrid()
    c2.step(2)
    with pytest.raises(dem.AvfError):
        c2.window_run(10)
    c2.close()


def test_autograd_harness_matches_dual_tangent(gpu, state):
    inst, pp, aa, anga, t = state
    ps = profile_spec()
    target = np.linspace(0.05, 0.01, ps.nbins)
    ctx = context_at(inst, pp, aa, anga, t)
    prob = calibrate.WindowProblem(ctx, 150, checkpoint=50, skin=0.05, w_profile=10.0,
                                   profile=ps, target=target[None, :])
    ms = prob.sphere_mass
    kn = inst.sys.kappa_pp * ms / 2.0
    phys = torch.tensor([0.5, 0.5, kn], dtype=torch.float64, requires_grad=True)
    dv = torch.zeros(len(pp), 3, dtype=torch.float64)
    loss = calibrate.window_loss(calibrate.tied_theta(phys, ms), dv, prob).sum()
    loss.backward()
    d = np.array([0.01, 0.02, 1e-3 * kn])
    tan = np.array([d[0], d[1], 2 * d[2] / ms] * 2)
    out = ctx.window_run(150, checkpoint=50, skin=0.05, w_profile=10.0, profile=ps,
                         target=target, theta=calibrate.tied_theta(phys.detach(), ms).numpy(),
                         tan_theta=tan, tangent=True, grad=False)
    got, want = float(phys.grad.numpy() @ d), out["dloss"][0]
    print("autograd", got, "dual", want)
    assert abs(got - want) <= 1e-9 * abs(want)
    hist = calibrate.calibrate(prob, [0.5, 0.5, kn], iters=2, lr=0.01, log=None)
    assert len(hist) == 2 and np.isfinite(hist[-1]["loss"])
    ctx.close()


def test_adjoint_configuration_scale(gpu):
    """The adjoint configuration at its size: 16 slope releases of 20,000 clumps (different
    seeds, 35 deg) after 8,000 plain steps, a 2,000-step window with checkpoints every 100
    steps, loss = mean runout + L2 misfit of a 64-bin deposit profile against the profile of
    the forward window at mu 0.45, eps 0.55; gradients wrt (mu, eps, kappa) pp and pw and
    the per-clump start velocities (dv ~ N(0, 1 cm/s)).  The frozen contact list holds
    every touching pair of the checkpoint states (list_valid), and the dual-number tangent
    of the GPU forward along each parameter equals the adjoint within 1e-9 relative in
    every instance."""
    count, n, steps = 16, 20_000, 2000
    insts = workloads.adjoint_batch(count, n)
    ctx = dem.Context(insts)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(8000)
    ps = abi.ProfileSpec()
    ps.x0, ps.x1, ps.width, ps.sigma, ps.nbins = 0.0, 40.0, 10.0, 0.0, 64
    s0 = insts[0].sys
    base = np.array([[s0.mu_pp, s0.eps_pp, s0.kappa_pp, s0.mu_pw, s0.eps_pw, s0.kappa_pw]] * count)
    tgt = base.copy()
    tgt[:, [0, 3]] = 0.45
    tgt[:, [1, 4]] = 0.55
    kw = dict(checkpoint=100, skin=0.1, w_runout=1.0, w_profile=1.0, profile=ps)
    target = ctx.window_run(steps, theta=tgt, grad=False, **kw)["profile"]
    dv = np.random.default_rng(5).normal(scale=1.0, size=(count * n, 3))
    out = ctx.window_run(steps, theta=base, dv=dv, target=target, **kw)
    d = out["diag"]
    print("adjoint configuration diag", d)
    assert d["list_valid"] == 1 and d["missed"] == 0
    assert np.isfinite(out["grad_theta"]).all() and np.isfinite(out["grad_dv"]).all()
    for k in range(6):
        e = np.zeros_like(base)
        e[:, k] = base[:, k] * 1e-2
        o2 = ctx.window_run(steps, theta=base, dv=dv, target=target, grad=False, tangent=True,
                            tan_theta=e, **kw)
        want = out["grad_theta"][:, k] * e[:, k]
        got = o2["dloss"]
        scale = np.maximum(np.maximum(np.abs(got), np.abs(want)), 1e-300)
        rel = np.abs(got - want) / scale
        print(abi.THETA[k], "max rel", rel.max(), "inst0", got[0], want[0])
        assert (rel <= 1e-9).all(), abi.THETA[k]
    # along the start velocities too (one random direction)
    tdv = np.random.default_rng(6).normal(size=dv.shape)
    o3 = ctx.window_run(steps, theta=base, dv=dv, target=target, grad=False, tangent=True,
                        tan_dv=tdv, **kw)
    lo = 0
    for i in range(count):
        want = float((out["grad_dv"][lo:lo + n] * tdv[lo:lo + n]).sum())
        assert abs(o3["dloss"][i] - want) <= 1e-9 * max(abs(want), abs(o3["dloss"][i]))
        lo += n
    ctx.close()
