"""The oracle against the reference tests' golden values and closed forms (CPU only)."""
import numpy as np
import pytest
from helpers import (anderson2pgf_freqs, anderson2pgf_model, anderson_test_model, cplx, dimer_model, fermionic,
                     bosonic, gf2site_model, gf4site_model, goldens, hubbard_atom)


def test_hamiltonian_ground_energy_and_eigensystem(oracle):
    """HamiltonianTest.cpp:70-89: E0 and eigen-residuals."""
    g = goldens()["HamiltonianTest"]
    H = dimer_model()
    ed = oracle.ED(4, H.mode_terms(), 1.0)
    assert abs(ed.E0 - g["E0"]) < g["tolerance"]
    for b in range(ed.n_blocks()):
        hm, V, w = ed.hmat(b), ed.evecs(b), ed.evals(b)
        assert np.abs(hm @ V - V * w).max() < 1e-10
        assert np.all(np.diff(w) >= 0)


def test_anderson_gf_golden(oracle):
    """AndersonTest.cpp:40-64, 1e-8."""
    g = goldens()["AndersonTest"]
    H, beta = anderson_test_model()
    idx = H.indices()
    ed = oracle.ED(len(idx), H.mode_terms(), beta)
    z = fermionic(np.arange(10), beta)
    for key, spin in (("G_up", 1), ("G_down", 0)):
        i = idx.index(("C", 0, spin))
        v = ed.gf(i, i, z)["values"]
        assert np.abs(v - cplx(g[key])).max() < g["tolerance"]


def test_gf2site_golden(oracle):
    g = goldens()["GF2siteTest"]
    H, beta = gf2site_model()
    idx = H.indices()
    ed = oracle.ED(len(idx), H.mode_terms(), beta)
    i = idx.index(("A", 0, 0))
    v = ed.gf(i, i, fermionic(np.arange(10), beta))["values"]
    assert np.abs(v - 1j * np.array(g["G_im"])).max() < g["tolerance"]


def test_gf4site_golden(oracle):
    g = goldens()["GF4siteTest"]
    H, beta = gf4site_model()
    idx = H.indices()
    ed = oracle.ED(len(idx), H.mode_terms(), beta)
    i = idx.index(("A", 0, 0))
    v = ed.gf(i, i, fermionic(np.arange(10), beta))["values"]
    assert np.abs(v - cplx(g["G"])).max() < g["tolerance"]


def test_hubbard_atom_gf_closed_form(oracle):
    """GF1siteTest.cpp:40-52,98-102 (1e-14), n in [-100, 100)."""
    U, mu, beta = 1.0, 0.4, 10.0
    H = hubbard_atom(U, mu)
    ed = oracle.ED(2, H.mode_terms(), beta)
    n = np.arange(-100, 100)
    w = np.pi * (2 * n + 1) / beta
    w0, w1, w2 = 1.0, np.exp(beta * mu), np.exp(-beta * (-2 * mu + U))
    Z = w0 + 2 * w1 + w2
    ref = (w0 + w1) / Z / (1j * w + mu) + (w1 + w2) / Z / (1j * w + mu - U)
    for i in range(2):
        assert np.abs(ed.gf(i, i, 1j * w)["values"] - ref).max() < 1e-14
    assert np.abs(ed.gf(0, 1, 1j * w)["values"]).max() < 1e-14


def test_anderson_2pgf_golden(oracle):
    """Anderson2PGFTest.cpp:46-57: chi_uuuu = chi_dddd at (w+W, w_p, w), PoleRes 1e-5, CoeffTol 1e-8."""
    g = goldens()["Anderson2PGFTest"]
    H, beta = anderson2pgf_model()
    idx = H.indices()
    ed = oracle.ED(len(idx), H.mode_terms(), beta)
    zs = anderson2pgf_freqs(beta)
    for spin in (1, 0):
        i = idx.index(("C", 0, spin))
        v = ed.g2(i, i, i, i, zs.ravel(), g["params"]["pole_resolution"], g["params"]["coefficient_tolerance"])["values"]
        assert np.abs(v - np.array(g["chi_uuuu"])).max() < g["tolerance"]


def test_vertex4_hubbard_atom_udud_closed_form(oracle):
    """Vertex4Test.cpp:59-79,152-167: chi_udud = Gamma G G G G - beta G G delta(n1,n3), 1e-10."""
    U, beta, n_iw = 1.0, 40.0, 2
    H = hubbard_atom(U, U / 2)
    ed = oracle.ED(2, H.mode_terms(), beta)
    up, dn = 1, 0
    w = lambda n: np.pi * (2 * n + 1) / beta
    ns = np.arange(-n_iw, n_iw)
    G = {n: ed.gf(up, up, np.array([1j * w(n)]))["values"][0] for n in range(-3 * n_iw, 3 * n_iw)}
    pts, ref = [], []
    for n1 in ns:
        for n2 in ns:
            for n3 in ns:
                n4 = n1 + n2 - n3
                o1, o2, o3, o4 = w(n1), w(n2), w(n3), w(n4)
                x = 1.0 / (1.0 + np.exp(beta * 0.5 * U))
                d = lambda a, b: 1.0 if a == b else 0.0
                dm = lambda a, b: 1.0 if a + b == -1 else 0.0
                sq = lambda v: v * v
                Gam = U - 0.125 * U ** 3 * (o1 ** 2 + o2 ** 2 + o3 ** 2 + o4 ** 2) / (o1 * o2 * o3 * o4)
                Gam += -0.1875 * U ** 5 / (o1 * o2 * o3 * o4)
                Gam += -beta * (2 * dm(n1, n2) + d(n1, n3)) * x * sq(0.5 * U) * (1 + sq(0.5 * U / o2)) * (1 + sq(0.5 * U / o3))
                Gam += beta * (2 * d(n2, n3) + d(n1, n3)) * (1 - x) * sq(0.5 * U) * (1 + sq(0.5 * U / o1)) * (1 + sq(0.5 * U / o2))
                ref.append(Gam * G[n1] * G[n2] * G[n3] * G[n4] - beta * G[n1] * G[n2] * d(n1, n3))
                pts.append([1j * o1, 1j * o2, 1j * o3])
    # the reference element evaluated in the test is (down, up, down, up) = (up, down, up, down) by spin symmetry
    v = ed.g2(dn, up, dn, up, np.array(pts).ravel(), 1e-4, 1e-12)["values"]
    assert np.abs(v - np.array(ref)).max() < 1e-10


def test_susceptibility_closed_forms(oracle):
    """SusceptibilityTest.cpp:33-163 (1e-14)."""
    U, mu, h, beta = 1.0, 0.4, 0.01, 10.0
    H = hubbard_atom(U, mu, -h)
    ed = oracle.ED(2, H.mode_terms(), beta)
    up, dn = 1, 0
    wts = np.array([1.0, np.exp(-beta * (-mu - h)), np.exp(-beta * (-mu + h)), np.exp(-beta * (-2 * mu + U))])
    wts /= wts.sum()
    wu, wd, w2 = wts[1], wts[2], wts[3]
    z = bosonic(np.arange(20), beta)
    n_up, n_dn = [(True, up), (False, up)], [(True, dn), (False, dn)]
    s_plus, s_minus = [(True, up), (False, dn)], [(True, dn), (False, up)]
    assert abs(ed.ensemble_average(n_up) - (wu + w2)) < 1e-14
    assert abs(ed.ensemble_average(s_plus)) < 1e-14
    v = ed.chi(s_plus, s_minus, z, 1e-8, 1e-8, True)["values"]
    ref = -(wu - wd) / (z - 2 * h)
    assert np.abs(v - ref).max() < 1e-14
    v = ed.chi(n_up, n_up, z, 1e-8, 1e-8, True)["values"]
    ref = np.where(np.arange(20) == 0, (wu + w2) * (1 - wu - w2) * beta, 0)
    assert np.abs(v - ref).max() < 1e-14
    v = ed.chi(n_up, n_dn, z, 1e-8, 1e-8, True)["values"]
    ref = np.where(np.arange(20) == 0, (w2 - (wu + w2) * (wd + w2)) * beta, 0)
    assert np.abs(v - ref).max() < 1e-14


@pytest.mark.parametrize("h", [0.0, 0.01])
def test_chi3_hubbard_atom_closed_form(oracle, h):
    """3PSusc1siteTest.cpp:62-212: every channel and spin pair on n1, n2 in [-20, 20) at 1e-14."""
    from helpers import chi3_atom_closed_form, chi3_grid
    U, mu, beta = 1.0, 0.4, 10.0
    ed = oracle.ED(2, hubbard_atom(U, mu, -h).mode_terms(), beta)
    ns = np.arange(-20, 20)
    zs = chi3_grid(ns, beta).ravel()
    for channel in (0, 1, 2):
        for i1 in (1, 0):
            for i2 in (1, 0):
                r = ed.chi3(channel, i1, i1, i2, i2, zs)
                ref = chi3_atom_closed_form(channel, i1, i2, U, mu, h, beta, ns)
                assert np.abs(r["values"] - ref).max() < 1e-14, (channel, i1, i2)
                if channel == 0 and i1 == i2:
                    assert r["n_parts"] == 0  # isVanishing()


def test_chi3_hubbard_atom_crossing_symmetry(oracle):
    """3PSusc1siteTest.cpp:215-250: chi3_xPH(i1, i4, i3, i2) = -chi3_PH(i1, i2, i3, i4)."""
    from helpers import chi3_grid
    beta = 10.0
    ed = oracle.ED(2, hubbard_atom(1.0, 0.4, -0.01).mode_terms(), beta)
    zs = chi3_grid(np.arange(-20, 20), beta).ravel()
    for i1, i2, i3, i4 in np.ndindex(2, 2, 2, 2):
        ph = ed.chi3(1, i1, i2, i3, i4, zs)["values"]
        xph = ed.chi3(2, i1, i4, i3, i2, zs)["values"]
        assert np.abs(xph + ph).max() < 1e-14


def test_chi3_3site_golden(oracle):
    """3PSusc3siteTest.cpp:72-156: PP, PH and xPH on n1, n2 in -1..1 at 1e-10."""
    from helpers import chi3_grid, cluster3_model
    g = goldens()["3PSusc3siteTest"]
    H, beta = cluster3_model()
    idx = H.indices()
    ed = oracle.ED(len(idx), H.mode_terms(), beta)
    zs = chi3_grid([-1, 0, 1], beta).ravel()
    spin = {"up": 1, "dn": 0}
    for channel, name in ((0, "PP"), (1, "PH"), (2, "xPH")):
        ii = [idx.index((site, 0, spin[s])) for site, s in g[name]["indices"]]
        ref = np.array([complex(a, b) for row in g[name]["chi3"] for a, b in row])
        v = ed.chi3(channel, *ii, zs)["values"]
        assert np.abs(v - ref).max() < g["tolerance"], name
