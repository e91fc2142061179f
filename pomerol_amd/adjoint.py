"""Bath-parameter gradients of G(i w_n): the batched solver's hand-written adjoint behind a
``torch.autograd.Function``.

A parametric batch holds instances ``H_k = sum_p theta[k, p] * H_p`` (one operator list per
instance).  ``LehmannG.apply(theta, solver)`` returns G_k(i w_n) for n = 0..nfreq-1 as a complex128
tensor (instances x nfreq); its backward calls ``BatchSolver::gradient`` (C++), which runs the
device adjoint: Lehmann sums -> Boltzmann weights and rotated c^+ -> rotation (recomputed from
the eigenvectors) -> eigen-decomposition adjoint -> Hamiltonian entries -> theta.

PyTorch's convention for a real loss L of a complex tensor G is grad_G = dL/dRe G + i dL/dIm G,
which is exactly the ``Gbar`` the C++ adjoint takes.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _pomerol
from .presets import CoulombS, Hopping, Level
from .workloads import terms


def anderson_operators(n_bath: int, imp: str = "C", bath_prefix: str = "b") -> list:
    """Operator list of the Anderson impurity for theta = (eps_0..eps_{n-1}, V_0..V_{n-1}, U):
    ``sum_k eps_k Level(b_k) + sum_k V_k Hopping(imp, b_k) + U (n_up n_dn - (n_up + n_dn) / 2)``,
    i.e. ``CoulombS(imp, U, -mu)`` with mu = U/2 (the convention of the benchmark workloads)."""
    ops = [Level(f"{bath_prefix}{k}", 1.0) for k in range(n_bath)]
    ops += [Hopping(imp, f"{bath_prefix}{k}", 1.0) for k in range(n_bath)]
    ops.append(CoulombS(imp, 1.0, -0.5))
    return [terms(o) for o in ops]


def anderson_theta(cfg) -> np.ndarray:
    """theta of an ``AndersonConfig`` in the order of ``anderson_operators``."""
    return np.concatenate([np.asarray(cfg.eps, float), np.asarray(cfg.V, float), [float(cfg.U)]])


class GSolver:
    """Batched G(i w_n) solves of parametric Hamiltonians on the device, with their adjoint."""

    def __init__(self, operators: list, theta, betas, index=("C", 0, 1), nfreq: int = 128, coeff_tol: float = 1e-8,
                 use_graph: bool = True):
        theta = np.asarray(theta, dtype=float)
        if theta.ndim != 2 or theta.shape[0] != len(operators):
            raise ValueError("theta must be (instances, parameters) with one operator list per instance")
        self.batch = _pomerol.Batch(operators, theta.tolist(), [float(b) for b in betas], tuple(index), int(nfreq),
                                    float(coeff_tol))
        self.batch.prepare()
        self.use_graph = use_graph
        self.nfreq = int(nfreq)
        self.theta = None  # parameters of the last solve

    def forward(self, theta) -> np.ndarray:
        theta = np.ascontiguousarray(theta, dtype=float)
        self.batch.set_params(theta.tolist())
        self.batch.solve(self.use_graph)
        self.theta = theta.copy()
        return self.batch.G()

    def backward(self, gbar) -> np.ndarray:
        """dL/dtheta of the last solve for gbar = dL/dRe G + i dL/dIm G (instances x nfreq)."""
        return np.asarray(self.batch.gradient(np.ascontiguousarray(gbar, dtype=complex)), dtype=float)


class LehmannG(torch.autograd.Function):
    """G = LehmannG.apply(theta, solver): theta (instances x parameters, float64) -> G (complex128)."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, solver: GSolver) -> torch.Tensor:
        th = theta.detach().cpu().numpy().astype(float)
        G = solver.forward(th)
        ctx.solver = solver
        ctx.theta = th
        ctx.device = theta.device
        return torch.from_numpy(G).to(theta.device)

    @staticmethod
    def backward(ctx, grad_G: torch.Tensor):
        solver = ctx.solver
        if solver.theta is None or not np.array_equal(solver.theta, ctx.theta):
            solver.forward(ctx.theta)  # another forward ran in between: restore this one's state
        g = solver.backward(grad_G.detach().cpu().numpy())
        return torch.from_numpy(g).to(ctx.device), None


def bath_fit_loss(G: torch.Tensor, G_target: torch.Tensor) -> torch.Tensor:
    """sum_n |G(i w_n) - G_target(i w_n)|^2, per instance summed over the batch."""
    return ((G - G_target).abs() ** 2).sum()
