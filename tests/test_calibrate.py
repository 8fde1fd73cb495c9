"""The torch.autograd.Function of the calibration harness, on a stand-in problem with a
known quadratic loss (no GPU): gradcheck of the wiring (per-instance scaling of the
window's gradients, dv rows per instance) and of the tied parameterisation."""
import numpy as np
import torch

from avalancheflume3d_amd import calibrate


class QuadProblem:
    """loss_i = sum_k a_ik theta_ik^2 + sum_{rows of i} b_r . dv_r^2 (instance-major rows)."""

    def __init__(self, sizes, seed=0):
        rng = np.random.default_rng(seed)
        self.ninst = len(sizes)
        self.nclumps = int(sum(sizes))
        self.rows = torch.from_numpy(np.repeat(np.arange(self.ninst), sizes))
        self.a = rng.uniform(0.5, 2.0, (self.ninst, 6))
        self.b = rng.uniform(0.5, 2.0, (self.nclumps, 3))
        self.sphere_mass = 0.0105

    def run(self, theta, dv, grad):
        rows = self.rows.numpy()
        loss = (self.a * theta ** 2).sum(1) + np.bincount(rows, (self.b * dv ** 2).sum(1),
                                                           minlength=self.ninst)
        out = {"loss": loss}
        if grad:
            out["grad_theta"] = 2 * self.a * theta
            out["grad_dv"] = 2 * self.b * dv
        return out


def test_window_loss_gradcheck():
    prob = QuadProblem([3, 5, 2])
    theta = torch.rand(3, 6, dtype=torch.float64, requires_grad=True)
    dv = torch.rand(10, 3, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t, d: calibrate.window_loss(t, d, prob), (theta, dv))


def test_tied_theta_chain_rule():
    prob = QuadProblem([4])
    phys = torch.tensor([0.5, 0.4, 1.0e7], dtype=torch.float64, requires_grad=True)
    dv = torch.zeros(4, 3, dtype=torch.float64)
    loss = calibrate.window_loss(calibrate.tied_theta(phys, prob.sphere_mass), dv, prob).sum()
    loss.backward()
    th = calibrate.tied_theta(phys.detach(), prob.sphere_mass).numpy()[0]
    g = 2 * prob.a[0] * th
    want = [g[0] + g[3], g[1] + g[4], (g[2] + g[5]) * 2.0 / prob.sphere_mass]
    assert np.allclose(phys.grad.numpy(), want, rtol=1e-12)
