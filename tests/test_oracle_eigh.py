"""The oracle's two dense eigensolvers (CPU only): cyclic Jacobi sweeps and the algorithm of the
reference's Eigen SelfAdjointEigenSolver (HamiltonianPart.cpp:98: Householder tridiagonalisation
and implicit symmetric QR with the Wilkinson shift), pinned against LAPACK (numpy) on random,
degenerate, split and clustered matrices, and against each other on the whole ED of a model."""
import numpy as np
import pytest
from helpers import dimer_model

from pomerol_amd import workloads as W


def _check(oracle, A, how, tol=1e-13):
    w, V = oracle.ED.eigh(A, how)
    n = A.shape[0]
    scale = max(1.0, np.abs(A).max() * n)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - np.linalg.eigvalsh(A)).max() <= tol * scale
    assert np.abs(A @ V - V * w).max() <= tol * scale
    assert np.abs(V.T @ V - np.eye(n)).max() <= tol * n


@pytest.mark.parametrize("how", ["qr", "jacobi"])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 40, 150])
def test_random_symmetric(oracle, how, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    _check(oracle, A + A.T, how)


@pytest.mark.parametrize("how", ["qr", "jacobi"])
def test_structured(oracle, how):
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((12, 12)))
    degenerate = Q @ np.diag([1.0, 1, 1, 2, 2, 3, 3, 3, 3, 4, 5, 5]) @ Q.T
    clustered = Q @ np.diag(1.0 + 1e-9 * np.arange(12)) @ Q.T
    split = np.diag(np.ones(11), 1) + np.diag(np.ones(11), -1)
    split[5, 6] = split[6, 5] = 0.0
    integer = np.diag(np.arange(12.0)[::-1]) + np.diag(np.ones(11), 1) + np.diag(np.ones(11), -1)
    for A in (degenerate, clustered, split, integer, np.zeros((4, 4)), np.eye(5)):
        _check(oracle, (A + A.T) / 2, how)


def test_ed_qr_matches_jacobi(oracle):
    """Both eigensolvers give the same spectrum and gauge-invariant G on config 1 and the
    HamiltonianTest dimer (E0 = -2.8860009, HamiltonianTest.cpp:70-89 golden)."""
    c = W.cfg1()
    H = c.expr()
    idx = H.indices()
    a = oracle.ED(len(idx), H.mode_terms(), c.beta)
    b = oracle.ED(len(idx), H.mode_terms(), c.beta, "qr")
    assert abs(a.E0 - b.E0) < 1e-12
    for k in range(a.n_blocks()):
        assert np.abs(a.evals(k) - b.evals(k)).max() < 1e-12
    u = idx.index(("C", 0, 1))
    z = W.g1_grid(32, c.beta)
    ga, gb = a.gf(u, u, z)["values"], b.gf(u, u, z)["values"]
    assert np.abs(ga - gb).max() < 1e-12 * np.abs(ga).max()
    d = dimer_model()
    e = oracle.ED(len(d.indices()), d.mode_terms(), 1.0, "qr")
    assert abs(e.E0 - (-2.8860009)) < 1e-7
