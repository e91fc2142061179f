"""Direct calls through the C ABI (include/pomx_c.h) with ctypes, checked against numpy."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

c_int64_p = ctypes.POINTER(ctypes.c_int64)


class Ctx:
    def __init__(self):
        import pomerol_amd

        self.lib = ctypes.CDLL(pomerol_amd.POMX_LIB)
        self.lib.pomx_last_error.restype = ctypes.c_char_p
        self.ctx = ctypes.c_void_p()
        assert self.lib.pomx_ctx_create(0, ctypes.byref(self.ctx)) == 0
        self.bufs = []

    def ok(self, rc):
        assert rc == 0, self.lib.pomx_last_error(self.ctx).decode()

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self.ok(self.lib.pomx_malloc(self.ctx, ctypes.c_size_t(max(a.nbytes, 8)), ctypes.byref(p)))
        self.ok(self.lib.pomx_h2d(self.ctx, p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes)))
        self.bufs.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        self.ok(self.lib.pomx_d2h(self.ctx, out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes)))
        return out

    def close(self):
        for p in self.bufs:
            self.lib.pomx_free(self.ctx, p)
        self.lib.pomx_ctx_destroy(self.ctx)


@pytest.fixture
def cabi():
    c = Ctx()
    yield c
    c.close()


def sym_with_spectrum(rng, n, degenerate):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    if degenerate:
        lam = np.repeat(rng.uniform(-3, 3, (n + 2) // 3), 3)[:n]
    else:
        lam = rng.uniform(-3, 3, n)
    A = (Q * lam) @ Q.T
    return (A + A.T) / 2


@pytest.mark.parametrize("degenerate", [False, True])
def test_eigh_batched(cabi, degenerate):
    rng = np.random.default_rng(7)
    sizes = [1, 2, 3, 5, 17, 33, 64, 65, 130, 257, 400]
    mats = [sym_with_spectrum(rng, n, degenerate) for n in sizes]
    print("[cabi] eigh sizes", sizes, "degenerate", degenerate, flush=True)
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
    for k, n in enumerate(sizes):
        A = mats[k]
        V = V_all[mo[k]:mo[k] + n * n].reshape(n, n)
        E = E_all[vo[k]:vo[k] + n]
        assert np.abs(E - np.linalg.eigvalsh(A)).max() < 1e-12 * max(1, np.abs(E).max())
        assert np.abs(A @ V - V * E).max() < 1e-12 * max(1, np.abs(E).max())
        assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12


def test_boltzmann(cabi):
    rng = np.random.default_rng(3)
    E = rng.uniform(-5, 5, 5000)
    d_E, d_w = cabi.put(E), cabi.put(np.zeros_like(E))
    E0, Z = ctypes.c_double(), ctypes.c_double()
    cabi.ok(cabi.lib.pomx_boltzmann(cabi.ctx, ctypes.c_int64(len(E)), d_E, ctypes.c_double(7.0), d_w,
                                    ctypes.byref(E0), ctypes.byref(Z)))
    w = cabi.get(d_w, E)
    ref = np.exp(-7.0 * (E - E.min()))
    assert E0.value == E.min()
    assert abs(Z.value - ref.sum()) < 1e-12 * ref.sum()
    assert np.abs(w - ref / ref.sum()).max() < 1e-15


class RotJob(ctypes.Structure):
    _fields_ = [("n_to", ctypes.c_int), ("n_from", ctypes.c_int), ("u_to_off", ctypes.c_int64),
                ("u_from_off", ctypes.c_int64), ("src_off", ctypes.c_int64), ("out_off", ctypes.c_int64)]


def test_rotate_batched(cabi):
    rng = np.random.default_rng(11)
    shapes = [(1, 1), (3, 5), (16, 16), (70, 33), (129, 200), (400, 375)]
    U, src, jobs, refs = [], [], [], []
    uoff = soff = ooff = 0
    for nt, nf in shapes:
        Ut, Uf = rng.standard_normal((nt, nt)), rng.standard_normal((nf, nf))
        s = np.zeros(nt, dtype=np.int32)
        pre = rng.choice(nf, size=min(nt, nf), replace=False)
        tgt = rng.choice(nt, size=len(pre), replace=False)
        for a, t in zip(pre, tgt):
            s[t] = (a + 1) * (1 if rng.random() < 0.5 else -1)
        X = np.zeros((nt, nf))
        for t in range(nt):
            if s[t]:
                X[t] = np.sign(s[t]) * Uf[abs(s[t]) - 1]
        refs.append((ooff, Ut.T @ X))
        jobs.append(RotJob(nt, nf, uoff, uoff + nt * nt, soff, ooff))
        U += [Ut.ravel(), Uf.ravel()]
        src.append(s)
        uoff += nt * nt + nf * nf
        soff += nt
        ooff += nt * nf
    d_U, d_s, d_o = cabi.put(np.concatenate(U)), cabi.put(np.concatenate(src)), cabi.put(np.zeros(ooff))
    arr = (RotJob * len(jobs))(*jobs)
    cabi.ok(cabi.lib.pomx_rotate_batched(cabi.ctx, len(jobs), arr, d_U, d_s, d_o, ctypes.c_double(1e-20)))
    out = cabi.get(d_o, np.zeros(ooff))
    for (off, ref), (nt, nf) in zip(refs, shapes):
        got = out[off:off + nt * nf].reshape(nt, nf)
        assert np.abs(got - ref).max() < 1e-12 * max(1, np.abs(ref).max())


def test_rotate_batched_empty_and_full_maps(cabi):
    """Gather maps with no mapped row (the product is zero), with every row mapped, and with
    mapped rows only past the first 16-row slab, in one batch (the kernel runs over the
    compacted map of each job)."""
    rng = np.random.default_rng(12)
    cases = [(70, 40, "none"), (130, 130, "all"), (200, 90, "tail")]
    U, src, jobs, refs = [], [], [], []
    uoff = soff = ooff = 0
    for nt, nf, kind in cases:
        Ut, Uf = rng.standard_normal((nt, nt)), rng.standard_normal((nf, nf))
        s = np.zeros(nt, dtype=np.int32)
        if kind == "all":  # nt == nf: a signed permutation
            s[:] = (rng.permutation(nf) + 1) * np.where(rng.random(nt) < 0.5, 1, -1)
        elif kind == "tail":
            tgt = np.arange(150, 200, 3)
            s[tgt] = (rng.choice(nf, size=len(tgt), replace=False) + 1) * -1
        X = np.zeros((nt, nf))
        for t in range(nt):
            if s[t]:
                X[t] = np.sign(s[t]) * Uf[abs(s[t]) - 1]
        refs.append((ooff, Ut.T @ X))
        jobs.append(RotJob(nt, nf, uoff, uoff + nt * nt, soff, ooff))
        U += [Ut.ravel(), Uf.ravel()]
        src.append(s)
        uoff += nt * nt + nf * nf
        soff += nt
        ooff += nt * nf
    d_U, d_s = cabi.put(np.concatenate(U)), cabi.put(np.concatenate(src))
    d_o = cabi.put(np.full(ooff, np.nan))
    arr = (RotJob * len(jobs))(*jobs)
    cabi.ok(cabi.lib.pomx_rotate_batched(cabi.ctx, len(jobs), arr, d_U, d_s, d_o, ctypes.c_double(1e-20)))
    out = cabi.get(d_o, np.zeros(ooff))
    for (off, ref), (nt, nf, kind) in zip(refs, cases):
        got = out[off:off + nt * nf].reshape(nt, nf)
        if kind == "none":
            assert np.all(got == 0.0)
        else:
            assert np.abs(got - ref).max() < 1e-12 * max(1, np.abs(ref).max())


class View(ctypes.Structure):
    _fields_ = [("base", ctypes.c_void_p), ("rs", ctypes.c_int64), ("cs", ctypes.c_int64)]


class PairPart(ctypes.Structure):
    _fields_ = [("n_out", ctypes.c_int), ("n_in", ctypes.c_int), ("A", View), ("B", View),
                ("E_out", ctypes.c_void_p), ("E_in", ctypes.c_void_p), ("w_out", ctypes.c_void_p),
                ("w_in", ctypes.c_void_p)]


@pytest.mark.parametrize("kind", [0, 1])
def test_pair_terms_view_layouts(cabi, kind):
    """pomx_pair_terms reads the operator views along whichever index they are contiguous in:
    the same parts given with row-major and with transposed storage give bit-identical term lists,
    and each part's residues add up to the brute-force sum over its candidates (tol 0: nothing is
    erased; GreensFunctionPart.cpp:84-94, SusceptibilityPart.cpp:93-116)."""
    rng = np.random.default_rng(21 + kind)
    shapes = [(150, 100), (64, 1), (3, 200)]
    res = 1e-8
    out = {}
    for layout in ("outer_fast", "inner_fast"):
        parts, sums, zps = [], [], []
        for n_out, n_in in shapes:
            A = rng.standard_normal((n_out, n_in)) * (rng.random((n_out, n_in)) < 0.3)
            B = rng.standard_normal((n_in, n_out)) * (rng.random((n_in, n_out)) < 0.5)
            Eo = np.round(rng.uniform(-2, 2, n_out), 2)  # repeated levels: shared pole bins
            Ei = np.round(rng.uniform(-2, 2, n_in), 2)
            wo, wi = rng.uniform(0, 1, n_out), rng.uniform(0, 1, n_in)
            ab = A * B.T
            P = Ei[None, :] - Eo[:, None]
            if kind == 0:
                sums.append((ab * (wo[:, None] + wi[None, :])).sum())
                zps.append(0.0)
            else:
                zero = np.abs(P) < res
                sums.append((ab * (wo[:, None] - wi[None, :]))[~zero].sum())
                zps.append((ab * wo[:, None])[zero].sum())
            if layout == "outer_fast":  # A stored transposed, B row-major
                vA = View(cabi.put(A.T.copy()).value, 1, n_out)
                vB = View(cabi.put(B).value, n_out, 1)
            else:  # A row-major, B stored transposed
                vA = View(cabi.put(A).value, n_in, 1)
                vB = View(cabi.put(B.T.copy()).value, 1, n_in)
            parts.append(PairPart(n_out, n_in, vA, vB, *(cabi.put(x).value for x in (Eo, Ei, wo, wi))))
        # (the draws repeat for the second layout: same seed order per layout)
        rng = np.random.default_rng(21 + kind)
        arr = (PairPart * len(parts))(*parts)
        d_r, d_p = ctypes.c_void_p(), ctypes.c_void_p()
        cnt = (ctypes.c_int64 * len(parts))()
        zp = (ctypes.c_double * len(parts))()
        cabi.ok(cabi.lib.pomx_pair_terms(cabi.ctx, kind, len(parts), arr, ctypes.c_double(res), ctypes.c_double(0.0),
                                         ctypes.byref(d_r), ctypes.byref(d_p), cnt, zp))
        n = sum(cnt)
        R = cabi.get(d_r, np.zeros(max(n, 1)))[:n]
        Pl = cabi.get(d_p, np.zeros(max(n, 1)))[:n]
        cabi.lib.pomx_free(cabi.ctx, d_r)
        cabi.lib.pomx_free(cabi.ctx, d_p)
        off = np.concatenate([[0], np.cumsum(list(cnt))])
        for k in range(len(parts)):
            assert abs(R[off[k]:off[k + 1]].sum() - sums[k]) < 1e-12 * max(1.0, abs(sums[k]))
            if kind == 1:
                assert abs(zp[k] - zps[k]) < 1e-12 * max(1.0, abs(zps[k]))
        out[layout] = (list(cnt), R, Pl, list(zp))
    a, b = out["outer_fast"], out["inner_fast"]
    assert a[0] == b[0] and a[3] == b[3]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.parametrize("layout", ["outer_fast", "inner_fast"])
def test_pair_terms_weight_bound(cabi, layout):
    """POMX_PAIR_UNIT_ELEMENTS (elements of magnitude <= 1): G candidates whose two states both
    have |w| <= tol / 2 are skipped unread; the term lists are bit-identical to the plain
    enumeration.  Boltzmann weights at a large beta leave few hot states (hot rows and columns
    scattered, not a prefix), plus parts with no hot state and with only hot states."""
    rng = np.random.default_rng(31)
    tol, res = 1e-8, 1e-8
    parts, keep = [], []
    for n_out, n_in, beta, shift in [(300, 200, 30.0, 0.0), (130, 257, 20.0, 0.0), (70, 90, 400.0, 1.5), (65, 40, 1e-3, 0.0),
                                     (1, 150, 20.0, 0.0)]:
        A = rng.uniform(-1, 1, (n_out, n_in)) * (rng.random((n_out, n_in)) < 0.4)
        B = rng.uniform(-1, 1, (n_in, n_out)) * (rng.random((n_in, n_out)) < 0.6)
        Eo, Ei = rng.permutation(np.linspace(0, 3, n_out)), rng.permutation(np.linspace(0.01, 3, n_in))
        E0 = min(Eo.min(), Ei.min())
        wo, wi = np.exp(-beta * (Eo - E0 + shift)), np.exp(-beta * (Ei - E0 + shift))
        Z = max(wo.sum() + wi.sum(), 1.0)
        wo, wi = wo / Z, wi / Z
        if layout == "outer_fast":
            vA, vB = View(cabi.put(A.T.copy()).value, 1, n_out), View(cabi.put(B).value, n_out, 1)
        else:
            vA, vB = View(cabi.put(A).value, n_in, 1), View(cabi.put(B.T.copy()).value, 1, n_in)
        parts.append(PairPart(n_out, n_in, vA, vB, *(cabi.put(x).value for x in (Eo, Ei, wo, wi))))
        keep.append(int(((np.abs(A * B.T * (wo[:, None] + wi[None, :]))) > tol).sum()))
    print("[cabi] pair weight bound: candidates kept per part", keep, flush=True)
    arr = (PairPart * len(parts))(*parts)
    out = []
    for flags in (0, 0x100):
        d_r, d_p = ctypes.c_void_p(), ctypes.c_void_p()
        cnt = (ctypes.c_int64 * len(parts))()
        cabi.ok(cabi.lib.pomx_pair_terms(cabi.ctx, flags, len(parts), arr, ctypes.c_double(res), ctypes.c_double(tol),
                                         ctypes.byref(d_r), ctypes.byref(d_p), cnt, None))
        n = sum(cnt)
        R = cabi.get(d_r, np.zeros(max(n, 1)))[:n]
        P = cabi.get(d_p, np.zeros(max(n, 1)))[:n]
        cabi.lib.pomx_free(cabi.ctx, d_r)
        cabi.lib.pomx_free(cabi.ctx, d_p)
        out.append((list(cnt), R, P))
    assert sum(out[0][0]) > 0 and out[0][0][2] == 0
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("kind", [0, 1])
def test_pair_eval(cabi, kind):
    rng = np.random.default_rng(5)
    n, nz = 100000, 300
    R, P = rng.standard_normal(n), rng.uniform(-4, 4, n)
    z = 1j * np.pi * (2 * np.arange(nz) + 1) / 13.0
    zi = np.stack([z.real, z.imag], 1).ravel()
    d_R, d_P, d_z = cabi.put(R), cabi.put(P), cabi.put(zi)
    d_acc = cabi.put(np.zeros(2 * nz))
    cabi.ok(cabi.lib.pomx_pair_eval(cabi.ctx, kind, ctypes.c_int64(n), d_R, d_P, nz, d_z, d_acc))
    acc = cabi.get(d_acc, np.zeros(2 * nz))
    got = acc[0::2] + 1j * acc[1::2]
    ref = (R[None, :] / (z[:, None] - P[None, :])).sum(1) * (1 if kind == 0 else -1)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


class GemmJob(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int), ("trans_a", ctypes.c_int),
                ("trans_b", ctypes.c_int), ("lda", ctypes.c_int), ("ldb", ctypes.c_int), ("ldc", ctypes.c_int),
                ("a_off", ctypes.c_int64), ("b_off", ctypes.c_int64), ("c_off", ctypes.c_int64),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double)]


def test_gemm_batched(cabi):
    """Mixed shapes, all four transposition patterns, padded leading dimensions, alpha/beta."""
    rng = np.random.default_rng(11)
    shapes = [(1, 1, 1), (5, 7, 3), (64, 64, 64), (65, 33, 129), (130, 70, 17), (17, 200, 64), (3, 5, 0), (257, 129, 300),
              # larger jobs, every transposition pattern
              (300, 260, 70), (513, 300, 129), (256, 385, 16), (290, 257, 1),
              # 128 x 128 tiles (m, n >= 256, k >= 64) with partial tiles, every transposition pattern
              (300, 300, 64), (257, 260, 65), (260, 384, 100), (300, 257, 80)]
    A_parts, B_parts, C_parts, jobs, refs = [], [], [], [], []
    ao = bo = co = 0
    for q, (m, n, k) in enumerate(shapes):
        ta, tb = q % 2, (q // 2) % 2
        pa_, pb_, pc_ = q % 3, (q + 1) % 3, q % 2  # row padding
        ra, ca = (k, m) if ta else (m, k)
        rb, cb = (n, k) if tb else (k, n)
        A = rng.standard_normal((ra, ca + pa_))
        B = rng.standard_normal((rb, cb + pb_))
        C = rng.standard_normal((m, n + pc_))
        alpha, beta = (1.0, 0.0) if q % 3 == 0 else (-0.5, 1.5)
        opA = (A[:, :ca].T if ta else A[:, :ca])
        opB = (B[:, :cb].T if tb else B[:, :cb])
        ref = C.copy()
        ref[:, :n] = alpha * (opA @ opB) + beta * C[:, :n]
        jobs.append(GemmJob(m, n, k, ta, tb, ca + pa_, cb + pb_, n + pc_, ao, bo, co, alpha, beta))
        A_parts.append(A.ravel())
        B_parts.append(B.ravel())
        C_parts.append(C.ravel())
        refs.append((co, ref))
        ao += A.size
        bo += B.size
        co += C.size
    print("[cabi] gemm shapes", shapes, flush=True)
    d_A = cabi.put(np.concatenate(A_parts))
    d_B = cabi.put(np.concatenate(B_parts))
    Cflat = np.concatenate(C_parts)
    d_C = cabi.put(Cflat)
    arr = (GemmJob * len(jobs))(*jobs)
    cabi.ok(cabi.lib.pomx_gemm_batched(cabi.ctx, len(jobs), arr, d_A, d_B, d_C))
    out = cabi.get(d_C, Cflat)
    for off, ref in refs:
        got = out[off:off + ref.size].reshape(ref.shape)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    bad = GemmJob(4, 4, 4, 0, 0, 2, 4, 4, 0, 0, 0, 1.0, 0.0)  # lda < k
    assert cabi.lib.pomx_gemm_batched(cabi.ctx, 1, ctypes.byref(bad), d_A, d_B, d_C) == 1


def _aligned_eigh(H, ref=None):
    E, U = np.linalg.eigh(H)
    if ref is not None:
        U = U * np.sign(np.sum(U * ref, axis=0))
    return E, U


def test_eigh_tangent(cabi):
    """pomx_eigh_tangent against central differences of numpy's eigendecomposition (h = 1e-6):
    dE and dU = U K; K vanishes between exactly degenerate levels."""
    rng = np.random.default_rng(21)
    sizes = [1, 2, 7, 64, 130, 300]
    h = 1e-6
    blocks = []
    for n in sizes:
        H = rng.standard_normal((n, n))
        H = (H + H.T) / 2
        dH = rng.standard_normal((n, n))
        blocks.append((H, (dH + dH.T) / 2))
    # one more block with an exactly degenerate pair (eigenvalue 0.5 twice)
    Q, _ = np.linalg.qr(rng.standard_normal((9, 9)))
    lam = np.array([-2.0, -1.0, 0.5, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0])
    Hd = (Q * lam) @ Q.T
    dHd = rng.standard_normal((9, 9))
    blocks.append(((Hd + Hd.T) / 2, (dHd + dHd.T) / 2))
    sizes.append(9)
    mo = np.cumsum([0] + [n * n for n in sizes])[:-1].astype(np.int64)
    vo = np.cumsum([0] + sizes)[:-1].astype(np.int64)
    Es, Us = zip(*[_aligned_eigh(H) for H, _ in blocks])
    d_U = cabi.put(np.concatenate([U.ravel() for U in Us]))
    d_E = cabi.put(np.concatenate(Es))
    d_dH = cabi.put(np.concatenate([dH.ravel() for _, dH in blocks]))
    d_K = cabi.put(np.zeros(int(mo[-1] + sizes[-1] ** 2)))
    d_dE = cabi.put(np.zeros(sum(sizes)))
    n_arr = np.array(sizes, dtype=np.int32)
    cabi.ok(cabi.lib.pomx_eigh_tangent(cabi.ctx, len(sizes), n_arr.ctypes.data_as(ctypes.c_void_p),
                                       mo.ctypes.data_as(ctypes.c_void_p), vo.ctypes.data_as(ctypes.c_void_p), d_U, d_E,
                                       d_dH, d_K, d_dE))
    K_all = cabi.get(d_K, np.zeros(int(mo[-1] + sizes[-1] ** 2)))
    dE_all = cabi.get(d_dE, np.zeros(sum(sizes)))
    for b, n in enumerate(sizes[:-1]):
        H, dH = blocks[b]
        K = K_all[mo[b]:mo[b] + n * n].reshape(n, n)
        dE = dE_all[vo[b]:vo[b] + n]
        Ep, Up = _aligned_eigh(H + h * dH, Us[b])
        Em, Um = _aligned_eigh(H - h * dH, Us[b])
        fdE, fdU = (Ep - Em) / (2 * h), (Up - Um) / (2 * h)
        errE = np.abs(dE - fdE).max()
        errU = np.abs(Us[b] @ K - fdU).max() / max(1.0, np.abs(fdU).max())
        print(f"[cabi] eigh tangent n={n}: |dE - FD| {errE:.1e}, |U K - dU_FD| {errU:.1e}", flush=True)
        assert errE <= 1e-7 and errU <= 1e-6
        assert np.abs(K + K.T).max() <= 1e-12 * max(1.0, np.abs(K).max())
    K = K_all[mo[-1]:].reshape(9, 9)
    assert K[2, 3] == 0.0 and K[3, 2] == 0.0  # the degenerate pair carries no rotation
    assert K[2, 4] != 0.0 and K[1, 3] != 0.0


def test_boltzmann_tangent(cabi):
    rng = np.random.default_rng(23)
    E, dE, beta, h = rng.uniform(-3, 3, 4000), rng.standard_normal(4000), 4.0, 1e-6

    def weights(x):
        w = np.exp(-beta * (x - x.min()))
        return w / w.sum()

    w = weights(E)
    d_w, d_dE, d_dw = cabi.put(w), cabi.put(dE), cabi.put(np.zeros_like(E))
    cabi.ok(cabi.lib.pomx_boltzmann_tangent(cabi.ctx, ctypes.c_int64(len(E)), d_w, d_dE, ctypes.c_double(beta), d_dw))
    dw = cabi.get(d_dw, E)
    fd = (weights(E + h * dE) - weights(E - h * dE)) / (2 * h)
    assert np.abs(dw - fd).max() <= 1e-8 * max(1.0, np.abs(fd).max())
    assert abs(dw.sum()) <= 1e-12
