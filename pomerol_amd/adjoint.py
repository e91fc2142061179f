"""Bath-parameter gradients of G(i w_n): the batched solver's hand-written adjoint behind a
``torch.autograd.Function``; and the U-derivative of the up-down two-particle GF (forward-mode
tangent, ``TwoParticleU``).

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
from .presets import CoulombS, Hopping, Level, anderson
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


class AndersonTwoParticle:
    """chi_{u d u d}(w1, w2; w3) = <c_up c_dn c+_up c+_dn> of the Anderson impurity with bath
    (eps, V) at inverse temperature beta on a triplet list, as a function of U (mu = U/2), with
    dchi/dU from the device tangent (TwoParticleGF::computeTangent along
    dH/dU = n_up n_dn - (n_up + n_dn)/2)."""

    def __init__(self, eps, V, beta: float, triplets, pole_res: float = 1e-8, coeff_tol: float = 1e-16,
                 imp: str = "C"):
        self.eps = np.asarray(eps, dtype=float)
        self.V = np.asarray(V, dtype=float)
        self.beta = float(beta)
        self.triplets = np.ascontiguousarray(triplets, dtype=complex)
        self.pole_res = float(pole_res)
        self.coeff_tol = float(coeff_tol)
        self.imp = imp
        self.direction = terms(CoulombS(imp, 1.0, -0.5))

    def value_and_tangent(self, U: float):
        H = anderson(self.eps, self.V, U, U / 2, imp=self.imp)
        idx = H.indices()
        u, d = idx.index((self.imp, 0, 1)), idx.index((self.imp, 0, 0))
        e = _pomerol.ED(terms(H), self.beta)
        r = e.g2_tangent(u, d, u, d, self.triplets, self.direction, self.pole_res, self.coeff_tol)
        return np.asarray(r["values"]), np.asarray(r["tangent"])


class TwoParticleU(torch.autograd.Function):
    """chi = TwoParticleU.apply(U, model): U a 0-d float64 tensor, model an AndersonTwoParticle,
    chi complex128 over the model's triplets.  The forward pass evaluates the value and the tangent
    dchi/dU together (one parameter, so forward mode is the cheap direction); backward returns
    Re sum conj(grad_chi) dchi/dU, with grad_chi = dL/dRe chi + i dL/dIm chi."""

    @staticmethod
    def forward(ctx, U: torch.Tensor, model: AndersonTwoParticle) -> torch.Tensor:
        chi, dchi = model.value_and_tangent(float(U.detach().cpu()))
        ctx.dchi = torch.from_numpy(dchi)
        ctx.shape, ctx.dtype, ctx.device = U.shape, U.dtype, U.device
        return torch.from_numpy(chi).to(U.device)

    @staticmethod
    def backward(ctx, grad_chi: torch.Tensor):
        g = torch.real(torch.sum(torch.conj(grad_chi.detach().cpu()) * ctx.dchi))
        return g.to(dtype=ctx.dtype, device=ctx.device).reshape(ctx.shape), None
