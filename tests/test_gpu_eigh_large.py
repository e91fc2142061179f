"""Large-block eigensolver (blocked Householder + divide and conquer + WY back-transformation)
through the C ABI, checked against numpy/LAPACK: eigenvalues, residuals, orthogonality."""
import ctypes

import numpy as np
import pytest
from test_gpu_cabi import Ctx, sym_with_spectrum

pytestmark = pytest.mark.gpu


@pytest.fixture
def cabi():
    c = Ctx()
    yield c
    c.close()


def _check(A, w, V, tol):
    n = A.shape[0]
    scale = max(1.0, np.max(np.abs(w)))
    wref = np.linalg.eigvalsh(A)
    ev = np.max(np.abs(w - wref)) / scale
    res = np.max(np.abs(A @ V - V * w)) / scale
    orth = np.max(np.abs(V.T @ V - np.eye(n)))
    print(f"[eigh-large] n={n}: eigenvalue err {ev:.2e}, residual {res:.2e}, orthogonality {orth:.2e}", flush=True)
    assert np.all(np.diff(w) >= 0)
    assert ev <= tol and res <= tol and orth <= tol


def _tridiag(cabi, d, e):
    n = len(d)
    d = np.ascontiguousarray(d, dtype=float)
    e = np.ascontiguousarray(e, dtype=float)
    d_Z = cabi.put(np.zeros((n, n)))
    w = np.empty(n)
    cabi.ok(cabi.lib.pomx_eigh_tridiag(cabi.ctx, n, d.ctypes.data_as(ctypes.c_void_p), e.ctypes.data_as(ctypes.c_void_p),
                                       d_Z, w.ctypes.data_as(ctypes.c_void_p)))
    return w, cabi.get(d_Z, np.zeros((n, n)))


@pytest.mark.parametrize("kind", ["random", "split", "clustered", "integer"])
def test_tridiagonal_divide_and_conquer(cabi, kind):
    rng = np.random.default_rng({"random": 1, "split": 2, "clustered": 3, "integer": 4}[kind])
    n = 1000
    d = rng.standard_normal(n)
    e = rng.standard_normal(n - 1)
    if kind == "split":
        e[::37] = 0.0
    elif kind == "clustered":
        d = 1.0 + 1e-12 * rng.standard_normal(n)
        e = 1e-9 * rng.standard_normal(n - 1)
    elif kind == "integer":  # many exactly degenerate eigenvalues
        d = np.round(rng.standard_normal(n))
        e = np.where(rng.random(n - 1) < 0.4, 0.0, 1.0)
    print(f"[eigh-large] tridiagonal {kind} n={n}", flush=True)
    w, V = _tridiag(cabi, d, e)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    _check(T, w, V, 1e-12)


def _dense(cabi, mats):
    sizes = [m.shape[0] for m in mats]
    mo = np.cumsum([0] + [n * n for n in sizes])[:-1].astype(np.int64)
    vo = np.cumsum([0] + sizes)[:-1].astype(np.int64)
    flat = np.concatenate([m.ravel() for m in mats])
    d_m = cabi.put(flat)
    d_e = cabi.put(np.zeros(sum(sizes)))
    n_arr = np.array(sizes, dtype=np.int32)
    cabi.ok(cabi.lib.pomx_eigh_batched(cabi.ctx, len(sizes), n_arr.ctypes.data_as(ctypes.c_void_p),
                                       mo.ctypes.data_as(ctypes.c_void_p), vo.ctypes.data_as(ctypes.c_void_p), d_m, d_e))
    V_all = cabi.get(d_m, flat)
    E_all = cabi.get(d_e, np.zeros(sum(sizes)))
    return [(E_all[vo[k]:vo[k] + n], V_all[mo[k]:mo[k] + n * n].reshape(n, n)) for k, n in enumerate(sizes)]


@pytest.mark.parametrize("degenerate", [False, True])
def test_dense_large_blocks(cabi, degenerate):
    rng = np.random.default_rng(17)
    sizes = [513, 700, 31, 1500, 200]
    mats = [sym_with_spectrum(rng, n, degenerate) for n in sizes]
    print(f"[eigh-large] dense sizes {sizes} degenerate={degenerate}", flush=True)
    for A, (w, V) in zip(mats, _dense(cabi, mats)):
        _check(A, w, V, 1e-11)


@pytest.mark.parametrize("two_stage", [False, True])
def test_dense_hamiltonian_block(cabi, pa, monkeypatch, two_stage):
    """The largest block of a 6-bath Anderson model (1225 states) through the ED path, with the
    one-stage and (forced) two-stage reductions."""
    from pomerol_amd import workloads as W

    if two_stage:
        monkeypatch.setenv("POMX_EIGH_TWO_STAGE_MIN", "1000")

    cfg = W.AndersonConfig(6, 2.5, 10.0, seed=W.SEED + 77)
    e = pa.ED(W.terms(cfg.expr()), cfg.beta, True)
    sizes = [len(b) for b in e.blocks()]
    b = int(np.argmax(sizes))
    print(f"[eigh-large] Anderson 1+6 bath, block {b} of size {sizes[b]}, two-stage {two_stage}", flush=True)
    H = e.hmat(b)
    _check(H, e.evals(b), e.evecs(b), 1e-11)


def _sparse_integer(rng, n):
    """Sparse symmetric matrix with small integer entries: rank-deficient panels (reflectors with
    nothing to annihilate) and exactly degenerate eigenvalues, as in Hamiltonian blocks."""
    A = np.zeros((n, n))
    nz = rng.random((n, n)) < 4.0 / n
    A[nz] = rng.integers(-2, 3, int(nz.sum()))
    A = np.triu(A)
    A = A + np.triu(A, 1).T
    A[np.diag_indices(n)] = rng.integers(-3, 4, n)
    return A


@pytest.mark.parametrize("kind", ["random", "degenerate", "sparse"])
def test_two_stage_reduction(cabi, monkeypatch, kind):
    """The two-stage path (dense -> band by panel LQ and MFMA products, bulge chase to tridiagonal,
    Q2 and Q1 back-transformations) forced on blocks of at least 100 rows: sizes just above the
    band width, partial last panels, and blocks of different sizes in one plan."""
    monkeypatch.setenv("POMX_EIGH_TWO_STAGE_MIN", "100")
    rng = np.random.default_rng({"random": 31, "degenerate": 32, "sparse": 33}[kind])
    sizes = [513, 700, 31, 1500, 200, 130, 161]
    if kind == "sparse":
        mats = [_sparse_integer(rng, n) for n in sizes]
    else:
        mats = [sym_with_spectrum(rng, n, kind == "degenerate") for n in sizes]
    print(f"[eigh-large] two-stage {kind} sizes {sizes}", flush=True)
    for A, (w, V) in zip(mats, _dense(cabi, mats)):
        _check(A, w, V, 1e-11)


def test_two_stage_and_one_stage_in_one_solve(cabi, monkeypatch):
    """Blocks above the two-stage threshold and blocks below it in one call (two chained plans)."""
    monkeypatch.setenv("POMX_EIGH_TWO_STAGE_MIN", "600")
    rng = np.random.default_rng(41)
    sizes = [513, 700, 1500, 200, 64]
    mats = [sym_with_spectrum(rng, n, False) for n in sizes]
    print(f"[eigh-large] mixed one-/two-stage sizes {sizes}", flush=True)
    for A, (w, V) in zip(mats, _dense(cabi, mats)):
        _check(A, w, V, 1e-11)


def test_two_stage_default_threshold(cabi):
    """A 2600-row block takes the two-stage path with the default threshold (2048)."""
    rng = np.random.default_rng(43)
    A = sym_with_spectrum(rng, 2600, False)
    print("[eigh-large] two-stage by default, n=2600", flush=True)
    (w, V), = _dense(cabi, [A])
    _check(A, w, V, 1e-11)
