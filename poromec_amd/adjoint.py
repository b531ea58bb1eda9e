"""Differentiable backward-Euler march: a ``torch.autograd.Function`` over the
device adjoint (``pm_adjoint``).

The reference marches the Biot u-p system forward only (the ``solv`` macro,
Main.f:876-1123).  Here the same march, from a fixed initial state, is a
function of the per-material log Young's modulus and log permeability
parameter (``log e(1)``, ``log e(3)`` of elmt17_pm) to a sensor misfit

    J = sum_n sum_i (u[node_i, dof_i] after step n - target[n, i])**2

whose gradient comes from one forward and one adjoint march on the GPU (the
adjoint states solve ``K lambda_n = dJ/ds_n + C^T lambda_{n+1}``).  Every
evaluation restarts from the initial state captured when the object is built.
"""
from __future__ import annotations

import numpy as np
import torch

from .problem import Deck
from .solver import MefSolver, SolverConfig


class _MarchMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_E, log_k, owner):  # noqa: D401 - torch API
        J, gE, gk = owner._evaluate(log_E.detach().cpu().double().numpy(),
                                    log_k.detach().cpu().double().numpy())
        ctx.save_for_backward(torch.from_numpy(gE).to(log_E), torch.from_numpy(gk).to(log_k))
        return torch.tensor(J, dtype=log_E.dtype, device=log_E.device)

    @staticmethod
    def backward(ctx, gJ):
        gE, gk = ctx.saved_tensors
        return gJ * gE, gJ * gk, None


class MarchMisfit:
    """``J(log_E, log_k)`` for a deck, sensors and targets.

    ``sensors`` are ``(node, dof)`` pairs of free dofs (dof 3 = pressure at a
    vertex); ``target`` has shape ``(nsteps, len(sensors))``.  The solver
    configuration should use a tight Krylov tolerance (the adjoint assumes
    one solve per step, exact to that tolerance)."""

    def __init__(self, deck: Deck, sensors, target, nsteps: int | None = None,
                 cfg: SolverConfig | None = None, ckpt_every: int = 8, device: int = 0):
        self.deck = deck
        self.nsteps = int(nsteps or deck.nsteps)
        self.sensors = np.asarray(sensors, dtype=np.int64).reshape(-1, 2)
        self.target = np.asarray(target, dtype=np.float64).reshape(self.nsteps, len(self.sensors))
        self.cfg = cfg or SolverConfig(tol=1e-13, maxit=200000)
        self.ckpt_every = int(ckpt_every)
        self.solver = MefSolver(deck, device=device)
        self.state0 = self.solver.state()
        self.e0 = np.array(deck.e, dtype=np.float64, copy=True)

    def params(self):
        """the deck's own (log E, log k) as float64 tensors"""
        return (torch.from_numpy(np.log(self.e0[:, 0]).copy()),
                torch.from_numpy(np.log(self.e0[:, 2]).copy()))

    def _evaluate(self, log_E: np.ndarray, log_k: np.ndarray):
        e = self.e0.copy()
        e[:, 0] = np.exp(log_E)
        e[:, 2] = np.exp(log_k)
        self.deck.e = e
        try:
            self.solver.set_materials(self.deck)
        finally:
            self.deck.e = self.e0
        s = self.state0
        self.solver.set_state(s["u"], s["u0"], s["dp"], s["porosity"], s["t"])
        return self.solver.adjoint(self.nsteps, self.sensors, self.target, self.cfg,
                                   ckpt_every=self.ckpt_every)

    def __call__(self, log_E: torch.Tensor, log_k: torch.Tensor) -> torch.Tensor:
        return _MarchMisfit.apply(log_E, log_k, self)

    def close(self):
        self.solver.close()
