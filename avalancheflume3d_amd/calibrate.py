"""Calibration harness over the differentiable window.

The window (avf_window_run) maps contact-law parameters theta (per instance: mu, eps, kappa
for the particle-particle and particle-wall laws) and per-clump start-velocity
perturbations dv to a loss per instance (mean runout plus the L2 misfit of the deposit
profile against a target), with the hand-written adjoint giving dloss/dtheta and
dloss/ddv.  `window_loss` exposes it as a torch.autograd.Function, so a calibration is
ordinary torch code:

    problem = WindowProblem(ctx, steps=2000, checkpoint=100, profile=spec, target=h_obs)
    phys = torch.tensor([0.5, 0.5, 1e7], dtype=torch.float64, requires_grad=True)
    loss = window_loss(tied_theta(phys, problem.sphere_mass), dv, problem).sum()
    loss.backward()        # phys.grad = dloss / d(mu, eps, k_n)

Units are the reference's (cgs).  The normal stiffness k_n enters the law as
kappa = 2 k_n / m_s (the reference's spring km = 0.5 kappa m_eff, force_part.cu:121).
"""
from __future__ import annotations

import numpy as np
import torch

from . import abi, dem


class WindowProblem:
    """A context (its current state is the window start, left unchanged) plus the window
    settings and the loss definition."""

    def __init__(self, ctx: dem.Context, steps: int, checkpoint: int = 100, skin: float = 0.1,
                 w_runout: float = 1.0, w_profile: float = 1.0,
                 profile: abi.ProfileSpec | None = None, target=None):
        self.ctx = ctx
        self.steps, self.checkpoint, self.skin = steps, checkpoint, skin
        self.w_runout, self.w_profile = w_runout, w_profile
        self.profile = profile
        self.target = None if target is None else np.asarray(target, np.float64)
        self.ninst = len(ctx.instances)
        sizes = [ctx.n(i) for i in range(self.ninst)]
        self.nclumps = int(sum(sizes))
        # instance of every dv row (instance-major reference order)
        self.rows = torch.from_numpy(np.repeat(np.arange(self.ninst), sizes))
        inst0 = ctx.instances[0]
        self.sphere_mass = float(abi.SPH_FACT * inst0.sys.rmin ** 3 *
                                 float(inst0.particles["density"][0]))
        self.last = None

    def run(self, theta: np.ndarray | None, dv: np.ndarray | None, grad: bool) -> dict:
        out = self.ctx.window_run(self.steps, checkpoint=self.checkpoint, skin=self.skin,
                                  w_runout=self.w_runout, w_profile=self.w_profile,
                                  profile=self.profile, theta=theta, dv=dv, target=self.target,
                                  grad=grad)
        self.last = out
        return out


class _WindowLoss(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta: torch.Tensor, dv: torch.Tensor, problem: WindowProblem):
        need = bool(fctx.needs_input_grad[0] or fctx.needs_input_grad[1])
        out = problem.run(theta.detach().cpu().double().numpy(),
                          dv.detach().cpu().double().numpy(), need)
        if need:
            fctx.save_for_backward(torch.from_numpy(out["grad_theta"]),
                                   torch.from_numpy(out["grad_dv"]))
        fctx.rows = problem.rows
        return torch.from_numpy(out["loss"].copy()).to(theta.device)

    @staticmethod
    def backward(fctx, gloss: torch.Tensor):
        gth, gdv = fctx.saved_tensors
        g = gloss.detach().cpu().double()
        g_theta = (g[:, None] * gth) if fctx.needs_input_grad[0] else None
        g_dv = (g[fctx.rows][:, None] * gdv) if fctx.needs_input_grad[1] else None
        return g_theta, g_dv, None


def window_loss(theta: torch.Tensor, dv: torch.Tensor, problem: WindowProblem) -> torch.Tensor:
    """Loss per instance of the window; differentiable in theta (ninst x 6, float64) and
    dv (clumps x 3, float64, cm/s)."""
    return _WindowLoss.apply(theta, dv, problem)


def tied_theta(phys: torch.Tensor, sphere_mass: float, ninst: int = 1) -> torch.Tensor:
    """theta (ninst x 6) from shared physical parameters phys = (mu, eps, k_n), the same
    for particle-particle and particle-wall contacts: kappa = 2 k_n / m_s."""
    mu, eps, kn = phys[0], phys[1], phys[2]
    kappa = 2.0 * kn / sphere_mass
    one = torch.stack([mu, eps, kappa, mu, eps, kappa])
    return one.unsqueeze(0).expand(ninst, 6)


def calibrate(problem: WindowProblem, phys0, iters: int = 10, lr: float = 0.05,
              dv: np.ndarray | None = None, log=print) -> list[dict]:
    """Adam on log(mu), logit(eps), log(k_n): fits the shared contact-law parameters to the
    problem's target.  Returns the history (loss and parameters per iteration)."""
    p0 = torch.as_tensor(phys0, dtype=torch.float64)
    z = torch.stack([torch.log(p0[0]), torch.log(p0[1] / (1 - p0[1])), torch.log(p0[2])])
    z = z.clone().requires_grad_(True)
    opt = torch.optim.Adam([z], lr=lr)
    dvt = torch.zeros(problem.nclumps, 3, dtype=torch.float64) if dv is None else \
        torch.as_tensor(dv, dtype=torch.float64)
    hist = []
    for it in range(iters):
        opt.zero_grad()
        phys = torch.stack([torch.exp(z[0]), torch.sigmoid(z[1]), torch.exp(z[2])])
        loss = window_loss(tied_theta(phys, problem.sphere_mass, problem.ninst), dvt,
                           problem).sum()
        loss.backward()
        rec = {"iter": it, "loss": loss.item(), "mu": phys[0].item(), "eps": phys[1].item(),
               "k_n": phys[2].item()}
        hist.append(rec)
        if log:
            log(rec)
        opt.step()
    return hist
