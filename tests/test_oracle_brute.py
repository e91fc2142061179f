"""The oracle against the full-space numpy brute force and structural invariants (CPU only).

These pin the parts no golden covers: the sector partition, the rotation, G/chi/2PGF at
tight tolerances (no merging slack), on systems small enough for the brute force."""
import itertools

import numpy as np
import pytest
from helpers import bosonic, dimer_model, fermionic, hubbard_atom, rel_err

import brute
from pomerol_amd.presets import anderson


def small_anderson(nb=2, seed=3, U=2.0, beta=8.0):
    rng = np.random.default_rng(seed)
    return anderson(rng.uniform(-1, 1, nb), rng.uniform(0.2, 0.6, nb), U, U / 2), beta


@pytest.mark.parametrize("nb", [1, 2, 3])
def test_partition_is_n_up_n_down_sectors(oracle, nb):
    """Connected bath models: the partition is exactly the (N_up, N_down) sectors."""
    H, beta = small_anderson(nb)
    idx = H.indices()
    ed = oracle.ED(len(idx), H.mode_terms(), beta)
    up = [k for k, t in enumerate(idx) if t[2] == 1]
    dn = [k for k, t in enumerate(idx) if t[2] == 0]
    key = lambda s: (sum((s >> k) & 1 for k in up), sum((s >> k) & 1 for k in dn))
    blocks = ed.blocks()
    assert len(blocks) == (len(up) + 1) * (len(dn) + 1)
    assert sorted(s for b in blocks for s in b) == list(range(1 << len(idx)))
    for b in blocks:
        assert len({key(s) for s in b}) == 1 and b == sorted(b)
    assert [b[0] for b in blocks] == sorted(b[0] for b in blocks)  # numbered by smallest state


def test_partition_atom_phase2(oracle):
    """Hubbard atom: Phase I gives singletons (H diagonal) and Phase II keeps them."""
    ed = oracle.ED(2, hubbard_atom(1.0, 0.4).mode_terms(), 10.0)
    assert ed.blocks() == [[0], [1], [2], [3]]


def test_partition_operators_map_block_to_block(oracle):
    H, beta = small_anderson(2)
    n = len(H.indices())
    ed = oracle.ED(n, H.mode_terms(), beta)
    sb = ed.state_block()
    for i in range(n):
        for dag in (True, False):
            M = brute.op_matrix(n, [(dag, i)])
            targets = {}
            for s in range(1 << n):
                t = np.nonzero(M[:, s])[0]
                if len(t):
                    targets.setdefault(sb[s], set()).add(sb[t[0]])
            assert all(len(v) == 1 for v in targets.values())


def test_hamiltonian_block_diagonal_and_spectrum(oracle):
    H, beta = small_anderson(2)
    n = len(H.indices())
    ed = oracle.ED(n, H.mode_terms(), beta)
    full = brute.hamiltonian(n, H.mode_terms())
    sb = np.array(ed.state_block())
    assert np.all(full[sb[:, None] != sb[None, :]] == 0)
    ev = np.sort(np.concatenate([ed.evals(b) for b in range(ed.n_blocks())]))
    assert np.abs(ev - np.linalg.eigvalsh(full)).max() < 1e-12


def test_rotation_gauge_invariants(oracle):
    """Sum over blocks of |<i|c^+|j>|^2 weighted by w_i + w_j equals <{c, c^+}> = 1 (sum rule),
    and the c part is the transpose of the c^+ part."""
    H, beta = small_anderson(2)
    n = len(H.indices())
    ed = oracle.ED(n, H.mode_terms(), beta)
    for i in range(n):
        parts = ed.rotate([(True, i)])
        s = 0.0
        for p in parts:
            w_to, w_from = ed.weights(p["to"]), ed.weights(p["from"])
            s += np.sum(p["m"] ** 2 * (w_to[:, None] + w_from[None, :]))
        assert abs(s - 1.0) < 1e-12
        cparts = {(q["from"], q["to"]): q["m"] for q in ed.rotate([(False, i)])}
        for p in parts:
            assert np.abs(cparts[(p["to"], p["from"])] - p["m"].T).max() < 1e-12


@pytest.mark.parametrize("model", ["dimer", "anderson2"])
def test_gf_vs_brute(oracle, model):
    if model == "dimer":
        H, beta = dimer_model(), 3.0
    else:
        H, beta = small_anderson(2)
    n = len(H.indices())
    ed = oracle.ED(n, H.mode_terms(), beta)
    fe = brute.FullED(n, H.mode_terms(), beta)
    z = fermionic(np.arange(-5, 30), beta)
    for i, j in [(0, 0), (1, 1), (0, 2)]:
        v = ed.gf(i, j, z, 1e-8, 0.0)["values"]
        assert rel_err(v, fe.gf(i, j, z)) < 1e-12 or np.abs(fe.gf(i, j, z)).max() < 1e-14


def test_chi_vs_brute(oracle):
    H, beta = small_anderson(2)
    n = len(H.indices())
    ed = oracle.ED(n, H.mode_terms(), beta)
    fe = brute.FullED(n, H.mode_terms(), beta)
    z = bosonic(np.arange(0, 12), beta)
    for A, B in [([(True, 1), (False, 0)], [(True, 0), (False, 1)]), ([(True, 1), (False, 1)], [(True, 0), (False, 0)])]:
        v = ed.chi(A, B, z, 1e-8, 0.0, False)["values"]
        assert rel_err(v, fe.chi(A, B, z)) < 1e-11


@pytest.mark.parametrize("ijkl", [(1, 1, 1, 1), (1, 0, 1, 0), (0, 1, 1, 0), (1, 0, 0, 1)])
def test_2pgf_vs_brute_dimer(oracle, ijkl):
    """doc/gamma4.tex:38-70 evaluated by direct quadruple sums on the 16-state dimer,
    including bosonic transfer 0 (resonant terms) and tight tolerances (no merge slack)."""
    H, beta = dimer_model(), 2.0
    ed = oracle.ED(4, H.mode_terms(), beta)
    fe = brute.FullED(4, H.mode_terms(), beta)
    from pomerol_amd.workloads import g2_grid
    zs = g2_grid(4, 2, beta)
    v = ed.g2(*ijkl, zs.ravel(), 1e-8, 0.0)["values"]
    assert rel_err(v, fe.g2(*ijkl, zs)) < 1e-11


def test_planted_sign_error_is_detected(oracle):
    """A mistake planted in a copy of the brute-force reference (a dropped resonant term) must
    make the 2PGF comparison fail: the check is sensitive to the resonant part."""
    H, beta = dimer_model(), 2.0
    ed = oracle.ED(4, H.mode_terms(), beta)
    fe = brute.FullED(4, H.mode_terms(), beta)
    from pomerol_amd.workloads import g2_grid
    zs = g2_grid(4, 2, beta)
    good = fe.g2(1, 0, 1, 0, zs)
    fe.beta = 0.0  # plants: beta*w_i resonant numerators vanish
    bad = fe.g2(1, 0, 1, 0, zs)
    v = ed.g2(1, 0, 1, 0, zs.ravel(), 1e-8, 0.0)["values"]
    assert rel_err(v, good) < 1e-11 and rel_err(v, bad) > 1e-6


def _gauss_legendre_grid(beta, panels=60, order=24):
    """Composite Gauss-Legendre nodes and weights on [0, beta] (imaginary-time quadrature)."""
    x, w = np.polynomial.legendre.leggauss(order)
    edges = np.linspace(0.0, beta, panels + 1)
    nodes = np.concatenate([(a + b) / 2 + (b - a) / 2 * x for a, b in zip(edges[:-1], edges[1:])])
    weights = np.concatenate([(b - a) / 2 * w for a, b in zip(edges[:-1], edges[1:])])
    return nodes, weights


@pytest.mark.parametrize("model", ["dimer", "anderson2"])
def test_gf_tau_fourier_and_sum_rule(oracle, model):
    """G(tau) of the oracle: its Fourier integral int_0^beta e^{i w_n tau} G(tau) dtau (numerical
    quadrature) reproduces G(i w_n), and G(0) + G(beta) = -<{c, c^+}> = -1."""
    H, beta = (dimer_model(), 3.0) if model == "dimer" else small_anderson(2)
    ed = oracle.ED(len(H.indices()), H.mode_terms(), beta)
    tau, wq = _gauss_legendre_grid(beta)
    z = fermionic(np.arange(-4, 12), beta)
    for i in (0, 1):
        r = ed.gf(i, i, z, 1e-8, 0.0, list(tau) + [0.0, beta])
        g_tau, ends = r["tau_values"][:-2], r["tau_values"][-2:]
        fourier = np.array([np.sum(wq * np.exp(zz * tau) * g_tau) for zz in z])
        assert rel_err(fourier, r["values"]) < 1e-10
        assert abs(ends[0] + ends[1] + 1.0) < 1e-12


def test_chi_tau_fourier_and_symmetry(oracle):
    """chi(tau) of the oracle: its Fourier integral reproduces chi(i nu_m) (zero-pole weight and the
    disconnected part included), and chi(tau) = chi(beta - tau) for <n_up(tau) n_dn>."""
    H, beta = small_anderson(2)
    ed = oracle.ED(len(H.indices()), H.mode_terms(), beta)
    tau, wq = _gauss_legendre_grid(beta)
    z = bosonic(np.arange(0, 10), beta)
    nu, nd = [(True, 1), (False, 1)], [(True, 0), (False, 0)]
    sp, sm = [(True, 1), (False, 0)], [(True, 0), (False, 1)]
    for A, B in [(nu, nd), (sp, sm)]:
        for sub in (False, True):
            r = ed.chi(A, B, z, 1e-8, 0.0, sub, list(tau))
            fourier = np.array([np.sum(wq * np.exp(zz * tau) * r["tau_values"]) for zz in z])
            assert rel_err(fourier, r["values"]) < 1e-10
    t = np.linspace(0.0, beta, 9)
    r = ed.chi(nu, nd, z[:1], 1e-8, 0.0, False, list(t))
    assert np.abs(r["tau_values"] - r["tau_values"][::-1]).max() < 1e-12


def _block_diag_ed(ed, H, n, beta):
    """Full-space eigensystem assembled block by block on the oracle's partition (numpy eigh per
    block), with the block of every eigenstate and the Boltzmann weights of the whole space."""
    Hf = brute.hamiltonian(n, H.mode_terms())
    dim = 1 << n
    U = np.zeros((dim, dim))
    E = np.zeros(dim)
    blk = np.zeros(dim, int)
    at = 0
    for b, states in enumerate(ed.blocks()):
        Eb, Vb = np.linalg.eigh(Hf[np.ix_(states, states)])
        cols = np.arange(at, at + len(states))
        U[np.ix_(states, cols)] = Vb
        E[cols], blk[cols] = Eb, b
        at += len(states)
    w = np.exp(-beta * (E - E.min()))
    return U, E, blk, w / w.sum()


@pytest.mark.parametrize("tol,changes_g", [(1e-1, True), (5e-3, False), (2e-6, False)])
def test_truncate_blocks_vs_brute(oracle, tol, changes_g):
    """DensityMatrix::truncateBlocks (DensityMatrixPart.cpp:41-50): a block is retained iff one of
    its Boltzmann weights exceeds tol, the weights themselves unchanged.  Then G keeps the block
    pairs with either block retained (GreensFunction.cpp:62) and the 2PGF the block quadruples with
    any block retained (TwoParticleGF.cpp:87-95).  Pinned against weights, part counts and sums
    restricted to those block pairs and quadruples, all from numpy eigensystems of the blocks."""
    H, beta = small_anderson(2, seed=11, U=2.5, beta=4.0)
    n = len(H.indices())
    ed = oracle.ED(n, H.mode_terms(), beta)
    U, E, blk, w = _block_diag_ed(ed, H, n, beta)
    nb = ed.n_blocks()
    want = [bool(w[blk == b].max() > tol) for b in range(nb)]
    assert any(want) and not all(want)  # the cut falls inside the spectrum
    ed.truncate_blocks(tol)
    assert list(ed.retained()) == want
    for b in range(nb):
        assert np.abs(ed.weights(b) - w[blk == b]).max() < 1e-12  # both in ascending energy
    ret = np.array(want)[blk]  # retained flag of every eigenstate's block
    z = fermionic(np.arange(-4, 12), beta)
    i = H.indices().index(("C", 0, 1))
    r = ed.gf(i, i, z, 1e-8, 0.0)
    # parts: connected (outer, inner) block pairs of c_i / c_i^+ with either block retained
    pairs = {(p["to"], p["from"]) for p in ed.rotate([(False, i)])}
    assert r["n_parts"] == sum(1 for a, b in pairs if want[a] or want[b])
    A = U.T @ brute.op_matrix(n, [(False, i)]) @ U
    B = U.T @ brute.op_matrix(n, [(True, i)]) @ U
    R = A * B.T * (w[:, None] + w[None, :])  # R[n, m] = <n|c|m><m|c+|n>(w_n + w_m)
    P = E[None, :] - E[:, None]
    keep = (ret[:, None] | ret[None, :]) & (np.abs(R) > 0)
    gb = np.array([np.sum(R[keep] / (zz - P[keep])) for zz in z])
    assert rel_err(r["values"], gb) < 1e-12
    full = np.array([np.sum(R[np.abs(R) > 0] / (zz - P[np.abs(R) > 0])) for zz in z])
    if changes_g:  # the cut drops pairs that carry weight: a real test of the part selection
        assert rel_err(gb, full) > 1e-6


def test_truncated_2pgf_vs_brute(oracle):
    """The 2PGF part list and values under truncateBlocks: quadruples with any block retained."""
    H, beta = dimer_model(), 2.0
    ed = oracle.ED(4, H.mode_terms(), beta)
    U, E, blk, w = _block_diag_ed(ed, H, 4, beta)
    tol = 1e-2
    want = [bool(w[blk == b].max() > tol) for b in range(ed.n_blocks())]
    assert any(want) and not all(want)
    ed.truncate_blocks(tol)
    assert list(ed.retained()) == want
    ret = np.array(want)[blk]
    from pomerol_amd.workloads import g2_grid

    zs = g2_grid(2, 2, beta)
    for ijkl in [(1, 0, 1, 0), (1, 1, 1, 1)]:
        v = ed.g2(*ijkl, zs.ravel(), 1e-8, 0.0)["values"]
        fe = brute.FullED(4, H.mode_terms(), beta)
        fe.E, fe.U, fe.w = E, U, w
        keep4 = ret[:, None, None, None] | ret[None, :, None, None] | ret[None, None, :, None] | ret[None, None, None, :]
        assert rel_err(v, fe.g2(*ijkl, zs, mask=keep4)) < 1e-11
