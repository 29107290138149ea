"""torch.autograd binding of the batched CT-HYB sampler (no counterpart in
the reference).

``observables(theta, run)`` returns G(tau) [N][2][Ntau] and the double
occupancies <n_b,up n_b,down> [N] estimated by ``run`` (a ``ScoreRun``: the
batched chains at the model ``builder(theta)``) as float64 tensors; their
backward pass applies the score-function Jacobian of the same run
(``derivatives.score_estimate``), so that a loss built from them
back-propagates into the model parameters theta.  Entries of dG/dtheta that
the estimator does not provide (NaN: a theta that changes F(tau) other than by
a factor) contribute only where their incoming gradient is non-zero.
"""

from __future__ import annotations

import numpy as np
import torch

from .ctqmc import CtqmcGpu


class ScoreRun:
    """Batched sampling with score sums at a parameter point: ``run(theta)``
    builds the problem (``builder(theta) -> (Problem, InputDerivatives)``),
    warms ``nchains`` chains up for ``nwarm`` moves, measures ``nmeas`` times
    and returns ``CtqmcGpu.score_estimate``.  The last run's chain
    configurations stay in ``configs`` (warm starts for finite differences)."""

    def __init__(self, builder, nchains: int, seed: int, nwarm: int, nmeas: int,
                 taudiffmax: float | None = None, capacity: int = 256, ngroups: int = 64, device: int = 0):
        self.builder, self.nchains, self.seed = builder, int(nchains), int(seed)
        self.nwarm, self.nmeas, self.taudiffmax = int(nwarm), int(nmeas), taudiffmax
        self.capacity, self.ngroups, self.device = int(capacity), int(ngroups), int(device)
        self.configs = None
        self.last = None

    def __call__(self, theta) -> dict:
        prob, deriv = self.builder(np.asarray(theta, float))
        gpu = CtqmcGpu.from_problem(prob, nchains=self.nchains, seed=self.seed, capacity=self.capacity,
                                    device=self.device)
        try:
            if self.taudiffmax is not None:
                gpu.set_taudiffmax(self.taudiffmax)
            gpu.set_derivatives(deriv)
            gpu.warmup(self.nwarm)
            self.configs = [gpu.get_config(c) for c in range(self.nchains)]
            gpu.measure(self.nmeas)
            self.last = gpu.score_estimate(self.ngroups)
        finally:
            gpu.close()
        return self.last


class CtqmcObservables(torch.autograd.Function):
    """theta -> (G(tau), double occupancies) with score-function gradients."""

    @staticmethod
    def forward(ctx, theta, run):
        est = run(theta.detach().cpu().double().numpy())
        ctx.save_for_backward(torch.as_tensor(np.asarray(est["dgtau"]), dtype=torch.float64),
                              torch.as_tensor(np.asarray(est["ddouble_occ"]), dtype=torch.float64))
        ctx.theta_device = theta.device
        g = torch.as_tensor(np.asarray(est["gtau"]), dtype=torch.float64, device=theta.device)
        d = torch.as_tensor(np.asarray(est["double_occ"]), dtype=torch.float64, device=theta.device)
        return g, d

    @staticmethod
    def backward(ctx, grad_g, grad_d):
        dg, dd = ctx.saved_tensors
        nt = dg.shape[0]
        out = torch.zeros(nt, dtype=torch.float64)
        if grad_g is not None:
            gg = grad_g.detach().cpu().double()
            prod = torch.where(gg[None] == 0, torch.zeros_like(dg), dg * gg[None])
            out += prod.reshape(nt, -1).sum(1)
        if grad_d is not None:
            out += dd @ grad_d.detach().cpu().double()
        return out.to(ctx.theta_device), None


def observables(theta: torch.Tensor, run: ScoreRun):
    """(G(tau) [N][2][Ntau], double occupancies [N]) of ``run`` at ``theta``,
    differentiable with respect to theta."""
    return CtqmcObservables.apply(theta, run)
