"""pomx_pair_eval_noalloc (the Lehmann sum with its partial sums in the library's static device
array, no allocation) against pomx_pair_eval (partial sums from the context's pool), through the
C ABI, on two calls in a row, the second smaller; and its size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _sum(c, fn, kind, R, P, z, acc0):
    dR, dP, dz, dacc = c.put(R), c.put(P), c.put(z.view(np.float64)), c.put(acc0.view(np.float64))
    c.ok(fn(c.ctx, ctypes.c_int(kind), ctypes.c_int64(len(R)), dR, dP, ctypes.c_int(len(z)), dz, dacc))
    return c.get(dacc, acc0.view(np.float64)).view(np.complex128)


@pytest.mark.parametrize("kind", [0, 1])
def test_noalloc_matches_pool_path_twice(kind):
    """Same values as pomx_pair_eval (accumulated onto a non-zero d_acc) on a large call that fills
    the scratch array with many slices, then on a smaller one that reuses it; both against numpy."""
    c = Ctx()
    try:
        rng = np.random.default_rng(7 + kind)
        for nterms, nz in ((200_000, 1000), (3000, 64)):
            R = rng.standard_normal(nterms)
            P = rng.uniform(-4, 4, nterms)
            z = 1j * np.pi * (2 * np.arange(nz) + 1) / 10.0
            acc0 = rng.standard_normal(nz) + 1j * rng.standard_normal(nz)
            a = _sum(c, c.lib.pomx_pair_eval, kind, R, P, z, acc0)
            b = _sum(c, c.lib.pomx_pair_eval_noalloc, kind, R, P, z, acc0)
            sgn = 1.0 if kind == 0 else -1.0
            ref = acc0.copy()
            for t0 in range(0, nterms, 20_000):  # numpy in chunks of terms
                ref += sgn * (R[None, t0:t0 + 20_000] / (z[:, None] - P[None, t0:t0 + 20_000])).sum(axis=1)
            scale = np.abs(ref).max()
            print(f"[no-alloc] kind {kind} {nterms} terms x {nz}: |noalloc - pool| {np.abs(a - b).max():.1e}, "
                  f"|noalloc - numpy| / max {np.abs(b - ref).max() / scale:.1e}", flush=True)
            assert np.abs(a - b).max() <= 1e-14 * scale
            assert np.abs(b - ref).max() <= 1e-10 * scale
    finally:
        c.close()


def test_noalloc_rejects_oversize_frequency_list():
    c = Ctx()
    try:
        z = np.zeros(1_048_577, np.complex128)
        R, P = np.ones(4), np.ones(4)
        dR, dP, dz, dacc = c.put(R), c.put(P), c.put(z.view(np.float64)), c.put(z.view(np.float64))
        rc = c.lib.pomx_pair_eval_noalloc(c.ctx, ctypes.c_int(0), ctypes.c_int64(4), dR, dP, ctypes.c_int(len(z)), dz,
                                          dacc)
        assert rc == 3  # POMX_ERR_RANGE
        assert b"1048576" in c.lib.pomx_last_error(c.ctx)
    finally:
        c.close()
