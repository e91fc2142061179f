"""Three-point susceptibility on the GPU (C++ API -> C ABI -> gfx950 kernels) against the reference
tests' closed form and golden values and against the CPU oracle on the same inputs."""
import numpy as np
import pytest
from helpers import (anderson_test_model, chi3_atom_closed_form, chi3_grid, cluster3_model, goldens, hubbard_atom,
                     rel_err)

from pomerol_amd import workloads as W

pytestmark = pytest.mark.gpu

CHANNELS = {0: "PP", 1: "PH", 2: "xPH"}


def inject(e, o):
    """Oracle copy fed the GPU's spectrum: isolates the term pipeline from eigensolver roundoff."""
    for b in range(e.n_blocks()):
        o.set_spectrum(b, e.evals(b), e.evecs(b).ravel(), e.weights(b))
    return o


@pytest.mark.parametrize("h", [0.0, 0.01])
def test_chi3_hubbard_atom_closed_form(pa, h):
    """3PSusc1siteTest.cpp:62-212 at its 1e-14, through compute(), operator() and evaluate()."""
    U, mu, beta = 1.0, 0.4, 10.0
    print(f"[gpu-chi3] Hubbard atom h={h}", flush=True)
    e = pa.ED(W.terms(hubbard_atom(U, mu, -h)), beta)
    ns = np.arange(-20, 20)
    zs = chi3_grid(ns, beta)
    for channel in (0, 1, 2):
        for i1 in (1, 0):
            for i2 in (1, 0):
                r = e.chi3(channel, i1, i1, i2, i2, zs, keep=True)
                if channel == 0 and i1 == i2:
                    assert r["vanishing"] and np.all(r["values"] == 0)
                    continue
                ref = chi3_atom_closed_form(channel, i1, i2, U, mu, h, beta, ns)
                for key in ("values", "values_bulk"):
                    assert np.abs(r[key] - ref).max() < 1e-14, (CHANNELS[channel], i1, i2, key)
                # point-by-point operator() on the first 64 pairs
                assert np.abs(r["values_lazy"][:64] - ref[:64]).max() < 1e-14
                assert abs(r["part0_from_terms"] - r["part0"]) < 1e-14


def test_chi3_hubbard_atom_crossing_symmetry(pa):
    """3PSusc1siteTest.cpp:215-250: chi3_xPH(i1, i4, i3, i2) = -chi3_PH(i1, i2, i3, i4)."""
    beta = 10.0
    e = pa.ED(W.terms(hubbard_atom(1.0, 0.4, -0.01)), beta)
    zs = chi3_grid(np.arange(-20, 20), beta)
    for i1, i2, i3, i4 in np.ndindex(2, 2, 2, 2):
        ph = e.chi3(1, i1, i2, i3, i4, zs)["values"]
        xph = e.chi3(2, i1, i4, i3, i2, zs)["values"]
        assert np.abs(xph + ph).max() < 1e-14, (i1, i2, i3, i4)


def test_chi3_3site_golden(pa):
    """3PSusc3siteTest.cpp:72-156 at its 1e-10."""
    g = goldens()["3PSusc3siteTest"]
    H, beta = cluster3_model()
    idx = H.indices()
    e = pa.ED(W.terms(H), beta)
    zs = chi3_grid([-1, 0, 1], beta)
    spin = {"up": 1, "dn": 0}
    for channel, name in ((0, "PP"), (1, "PH"), (2, "xPH")):
        ii = [idx.index((site, 0, spin[s])) for site, s in g[name]["indices"]]
        ref = np.array([complex(a, b) for row in g[name]["chi3"] for a, b in row])
        v = e.chi3(channel, *ii, zs)["values"]
        assert np.abs(v - ref).max() < g["tolerance"], name


MODELS = {
    "anderson_test": anderson_test_model,
    "cfg1": lambda: (W.cfg1().expr(), W.cfg1().beta),
}


@pytest.mark.parametrize("name", list(MODELS))
@pytest.mark.parametrize("channel", [0, 1, 2])
def test_chi3_matches_oracle(pa, oracle, name, channel):
    """Values (Matsubara grid and off-axis pairs) at 1e-10 relative against the oracle; with the
    GPU's spectrum injected into the oracle, identical parts and merged-term counts."""
    H, beta = MODELS[name]()
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    b0 = next(k for k, t in enumerate(idx) if t[0] != "C" and t[2] == 1)  # a spin-up bath mode
    print(f"[gpu-chi3] {name} channel {CHANNELS[channel]}", flush=True)
    e = pa.ED(W.terms(H), beta)
    o = oracle.ED(len(idx), H.mode_terms(), beta)
    rng = np.random.default_rng(7)
    grid = chi3_grid(np.arange(-6, 6), beta)
    off = rng.normal(size=(16, 2)) + 1j * rng.normal(size=(16, 2))
    zs = np.concatenate([grid, off])
    for ijkl in [(u, u, d, d), (u, d, d, u), (d, u, u, d), (u, b0, b0, u)]:
        r = e.chi3(channel, *ijkl, zs)
        ref = o.chi3(channel, *ijkl, zs.ravel())
        assert r["n_parts"] == ref["n_parts"], ijkl
        assert list(r["part_blocks"]) == list(ref["part_blocks"]), ijkl
        if ref["n_parts"] == 0:
            assert np.all(r["values"] == 0)
            continue
        assert rel_err(r["values"], ref["values"]) < 1e-10, ijkl
        same = inject(e, oracle.ED(len(idx), H.mode_terms(), beta)).chi3(channel, *ijkl, zs.ravel())
        assert rel_err(r["values"], same["values"]) < 1e-12, ijkl
        assert list(r["part_ff"]) == list(same["part_ff"]), ijkl
        assert list(r["part_fb"]) == list(same["part_fb"]), ijkl
        assert list(r["part_res"]) == list(same["part_res"]), ijkl


def test_chi3_container(pa):
    """ThreePointSusceptibilityContainer::computeAll (ThreePointSusceptibilityContainer.cpp:19-48):
    every requested element equals the directly computed one; an element's permutation images
    (IndexContainer4.hpp:190-237) receive the empty list of an already computed element."""
    H, beta = anderson_test_model()
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    e = pa.ED(W.terms(H), beta)
    zs = chi3_grid(np.arange(-4, 4), beta)
    want = [(u, u, d, d), (d, d, u, u)]
    out = e.chi3_container(1, want, zs)
    for ijkl in want:
        assert np.abs(out[ijkl] - e.chi3(1, *ijkl, zs)["values"]).max() < 1e-14
    # (u, d, u, d) brings images (d, u, u, d), (u, d, d, u), (d, u, d, u)
    out = e.chi3_container(1, [(u, d, u, d)], zs)
    first = min(out)
    assert len(out) == 4
    for k, v in out.items():
        assert len(v) == (len(zs) if k == first else 0), k
    assert np.abs(out[first] - e.chi3(1, u, d, u, d, zs)["values"]).max() < 1e-14


def test_chi3_config2_matches_oracle(pa, oracle):
    """The workload of the bench's 3-point susceptibility leg (Anderson 1+5 bath, 4096 states,
    truncateBlocks(1e-12), CoefficientTolerance 1e-12, PH up-up-dn-dn): values on 8 x 8 pairs at
    1e-10 relative and every part's merged-term counts equal to the oracle's (whose merge runs the
    reference's sequential TermList algorithm), erasures included."""
    c = W.cfg2()
    H = c.expr()
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    zs = chi3_grid(np.arange(-4, 4), c.beta)
    print("[gpu-chi3] config 2 on the device", flush=True)
    e = pa.ED(W.terms(H), c.beta)
    e.truncate_blocks(1e-12)
    r = e.chi3(1, u, u, d, d, zs, 1e-8, 1e-12)
    print("[gpu-chi3] config 2 on the oracle (about 25 s)", flush=True)
    o = oracle.ED(len(idx), H.mode_terms(), c.beta)
    o.truncate_blocks(1e-12)
    ref = o.chi3(1, u, u, d, d, zs.ravel(), 1e-8, 1e-12)
    assert r["n_parts"] == ref["n_parts"] and list(r["part_blocks"]) == list(ref["part_blocks"])
    assert rel_err(r["values"], ref["values"]) < 1e-10
    for key in ("part_ff", "part_fb", "part_res"):
        assert list(r[key]) == list(ref[key]), key
