"""Seeded model and frequency-grid definitions of the benchmark workloads.

* Anderson impurity (single orbital "C" + bath sites "b0".."b{n-1}"): ``CoulombS(C, U, -mu)
  + sum_k Level(b_k, eps_k) + Hopping(C, b_k, V_k)`` with eps_k and V_k drawn uniformly from a
  ``numpy.random.default_rng(seed)`` stream (first all eps_k, then all V_k), mu = U/2.
* 3x3 periodic Hubbard cluster with nearest (t) and diagonal (t') hopping, half filling.
* 2PGF grid as in ``prog/quantum_model.cpp:220-231``: for every bosonic nu_m (outer), fermionic
  w_n3, fermionic w_n2 (inner) the triplet (nu + w3, w2, w3); fermionic indices run over
  [-N/2, N/2), bosonic over [0, M).

The drawn parameters are also recorded in ``tests/golden/configs.json``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .presets import Expr, anderson, hubbard_cluster

SEED = 20260901


@dataclass
class AndersonConfig:
    n_bath: int
    U: float
    beta: float
    eps_range: tuple = (-1.0, 1.0)
    V_range: tuple = (0.2, 0.6)
    seed: int = SEED
    eps: np.ndarray = field(init=False)
    V: np.ndarray = field(init=False)

    def __post_init__(self):
        rng = np.random.default_rng(self.seed)
        self.eps = rng.uniform(*self.eps_range, self.n_bath)
        self.V = rng.uniform(*self.V_range, self.n_bath)

    @property
    def mu(self) -> float:
        return self.U / 2

    def expr(self) -> Expr:
        return anderson(self.eps, self.V, self.U, self.mu)


def cfg1() -> AndersonConfig:
    """Plain, CPU-runnable: 1+3 bath, 256 states, U=2, beta=10."""
    return AndersonConfig(3, 2.0, 10.0)


def cfg2() -> AndersonConfig:
    """Plain on GPU: 1+5 bath, 4096 states (largest block 400), U=3, beta=20."""
    return AndersonConfig(5, 3.0, 20.0)


def cfg3_expr(seed: int = SEED) -> tuple[Expr, float, float]:
    """3x3 periodic Hubbard cluster, t=1, t' ~ U[-0.3, 0], U=4, mu=2, beta=5."""
    tp = float(np.random.default_rng(seed).uniform(-0.3, 0.0))
    return hubbard_cluster(3, 3, 1.0, tp, 4.0, 2.0), 5.0, tp


def batch_configs(n: int = 512, seed: int = SEED) -> list[AndersonConfig]:
    """Batch of Anderson instances: bath size in {3,4,5,6}, eps ~ U[-1.5,1.5], V ~ U[0.1,0.8],
    U ~ U[1,5], beta ~ U[5,50]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nb = int(rng.integers(3, 7))
        U = float(rng.uniform(1.0, 5.0))
        beta = float(rng.uniform(5.0, 50.0))
        out.append(AndersonConfig(nb, U, beta, (-1.5, 1.5), (0.1, 0.8), seed=seed + 1 + k))
    return out


def adjoint_configs(n: int = 64, seed: int = SEED) -> tuple[list[AndersonConfig], list[AndersonConfig]]:
    """Bath-fit gradient workload: n Anderson instances with 5 bath sites (eps ~ U[-1,1],
    V ~ U[0.2,0.6], U ~ U[2,4], beta = 20) and, for each, a target parameter set drawn the same
    way from an independent seed (G_target is that target's G on the same frequencies)."""
    rng = np.random.default_rng(seed + 7919)
    Us = rng.uniform(2.0, 4.0, 2 * n)
    cfgs = [AndersonConfig(5, float(Us[k]), 20.0, seed=seed + 10_000 + k) for k in range(n)]
    targets = [AndersonConfig(5, float(Us[n + k]), 20.0, seed=seed + 20_000 + k) for k in range(n)]
    return cfgs, targets


def fermionic(n: np.ndarray | int, beta: float) -> np.ndarray:
    return 1j * np.pi * (2 * np.asarray(n) + 1) / beta


def bosonic(m: np.ndarray | int, beta: float) -> np.ndarray:
    return 1j * np.pi * 2 * np.asarray(m) / beta


def g1_grid(n_freq: int, beta: float) -> np.ndarray:
    return fermionic(np.arange(n_freq), beta)


def g2_grid(n_fermionic: int, n_bosonic: int, beta: float) -> np.ndarray:
    """(N, 3) complex triplets, nu outermost, w3, then w2 innermost."""
    f = fermionic(np.arange(-n_fermionic // 2, n_fermionic // 2), beta)
    b = bosonic(np.arange(n_bosonic), beta)
    out = np.empty((len(b), len(f), len(f), 3), dtype=complex)
    out[..., 0] = b[:, None, None] + f[None, :, None]
    out[..., 1] = f[None, None, :]
    out[..., 2] = f[None, :, None]
    return out.reshape(-1, 3)


def spin_orbital(H: Expr, label: str, spin: int, orbital: int = 0) -> int:
    """ParticleIndex of (label, orbital, spin) in the expression's sorted index list."""
    return H.indices().index((label, orbital, spin))


def terms(H: Expr) -> list:
    """Term list in the form the C++ binding takes: [(coeff, [(dag, (label, orb, spin)), ...])]."""
    return [(c, [(bool(d), i) for d, i in m]) for c, m in H.terms]
