"""Host-side logic without a GPU: presets, index ordering, workloads and fixtures (CPU only)."""
import json
import os

import numpy as np
from helpers import GOLDEN_DIR

from pomerol_amd import workloads as W
from pomerol_amd.presets import CoulombS, Hopping, anderson


def test_index_order_matches_index_classification():
    """IndexClassification.hpp:133-140: indices are the sorted (label, orbital, spin) tuples."""
    H = anderson([0.1, 0.2], [0.3, 0.4], 2.0, 1.0)
    assert H.indices() == [("C", 0, 0), ("C", 0, 1), ("b0", 0, 0), ("b0", 0, 1), ("b1", 0, 0), ("b1", 0, 1)]


def test_hopping_is_hermitian():
    H = Hopping("A", "B", -0.7)
    assert sorted(H.terms) == sorted(H.hc().terms)


def test_config_fixture_matches_seeded_draws():
    with open(os.path.join(GOLDEN_DIR, "configs.json")) as f:
        cfg = json.load(f)
    for name, c in (("cfg1", W.cfg1()), ("cfg2", W.cfg2())):
        assert np.allclose(cfg[name]["eps"], c.eps, rtol=0, atol=0)
        assert np.allclose(cfg[name]["V"], c.V, rtol=0, atol=0)
        assert cfg[name]["n_bath"] == c.n_bath and cfg[name]["U"] == c.U and cfg[name]["beta"] == c.beta
    for k, c in enumerate(W.batch_configs(8)):
        assert cfg["batch_head"][k]["n_bath"] == c.n_bath
        assert np.array_equal(cfg["batch_head"][k]["eps"], c.eps)


def test_config_ranges():
    c = W.cfg2()
    assert c.n_bath == 5 and np.all((-1 <= c.eps) & (c.eps <= 1)) and np.all((0.2 <= c.V) & (c.V <= 0.6))
    b = W.batch_configs(64)
    assert {x.n_bath for x in b} == {3, 4, 5, 6}
    assert all(1 <= x.U <= 5 and 5 <= x.beta <= 50 for x in b)


def test_g2_grid_layout():
    beta = 20.0
    g = W.g2_grid(32, 8, beta)
    assert g.shape == (32 * 32 * 8, 3)
    # nu outermost, w3 middle, w2 innermost; first component nu + w3
    m, n3, n2 = 3, 5, 7
    row = g[(m * 32 + n3) * 32 + n2]
    f = W.fermionic(np.arange(-16, 16), beta)
    assert np.isclose(row[1], f[n2]) and np.isclose(row[2], f[n3]) and np.isclose(row[0], W.bosonic(m, beta) + f[n3])


def test_hubbard_cluster_bonds():
    _, _, tp = W.cfg3_expr()
    assert -0.3 <= tp <= 0.0
    H, beta, _ = W.cfg3_expr()
    assert len(H.indices()) == 18 and beta == 5.0


def test_lehmann_g_autograd_convention():
    """LehmannG hands the solver dL/dRe G + i dL/dIm G and returns its real gradient unchanged:
    checked with a linear stand-in solver G = theta @ M against torch's own autograd."""
    import torch

    from pomerol_amd.adjoint import LehmannG

    rng = np.random.default_rng(3)
    M = rng.standard_normal((4, 6)) + 1j * rng.standard_normal((4, 6))
    Gt = rng.standard_normal((2, 6)) + 1j * rng.standard_normal((2, 6))

    class Linear:
        theta = None

        def forward(self, th):
            self.theta = th.copy()
            return th @ M

        def backward(self, gbar):  # dL/dtheta = Re(conj(gbar) M^T)
            return np.real(np.conj(gbar) @ M.T)

    th0 = rng.standard_normal((2, 4))
    a = torch.tensor(th0, requires_grad=True)
    L = ((LehmannG.apply(a, Linear()) - torch.from_numpy(Gt)).abs() ** 2).sum()
    L.backward()
    b = torch.tensor(th0, requires_grad=True)
    L2 = ((b.to(torch.complex128) @ torch.from_numpy(M) - torch.from_numpy(Gt)).abs() ** 2).sum()
    L2.backward()
    assert np.allclose(a.grad.numpy(), b.grad.numpy(), rtol=1e-12, atol=1e-12)
