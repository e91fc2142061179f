"""GPU path (C++ API -> C ABI -> gfx950 kernels) against the CPU oracle on the same inputs."""
import numpy as np
import pytest
from helpers import (anderson2pgf_freqs, anderson2pgf_model, anderson_test_model, bosonic, cplx, dimer_model, fermionic,
                     gf2site_model, gf4site_model, goldens, hubbard_atom, rel_err)

from pomerol_amd import workloads as W

pytestmark = pytest.mark.gpu


def build(pa, oracle, H, beta, keep_hmat=True):
    print(f"[gpu-parity] model with {len(H.indices())} modes, beta={beta}", flush=True)
    e = pa.ED(W.terms(H), beta, keep_hmat)
    o = oracle.ED(len(H.indices()), H.mode_terms(), beta)
    return e, o


def inject(e, o):
    """Oracle copy fed the GPU's spectrum: isolates the term pipelines from eigensolver roundoff."""
    for b in range(e.n_blocks()):
        o.set_spectrum(b, e.evals(b), e.evecs(b).ravel(), e.weights(b))
    return o


MODELS = {
    "dimer": lambda: (dimer_model(), 1.0),
    "anderson_test": anderson_test_model,
    "gf4site": gf4site_model,
    "anderson2pgf": anderson2pgf_model,
    "cfg1": lambda: (W.cfg1().expr(), W.cfg1().beta),
}


@pytest.mark.parametrize("name", list(MODELS))
def test_sectors_and_eigensystem(pa, oracle, name):
    H, beta = MODELS[name]()
    e, o = build(pa, oracle, H, beta)
    assert e.blocks() == o.blocks()  # bit-exact sector partition and state order
    assert abs(e.E0 - o.E0) < 1e-12
    for b in range(e.n_blocks()):
        hm, V, w = e.hmat(b), e.evecs(b), e.evals(b)
        # same sparsity; values are sums of the same inputs, in another order (the operator
        # expression is normal-ordered), so an exact zero may come out as a rounding residue
        tiny = 1e-14 * max(1.0, np.abs(hm).max())
        assert np.array_equal(np.abs(hm) > tiny, np.abs(o.hmat(b)) > tiny)
        assert np.abs(hm - o.hmat(b)).max() < 1e-14 * max(1.0, np.abs(hm).max())
        assert np.all(np.diff(w) >= 0)
        scale = max(1.0, np.abs(w).max())
        assert np.abs(w - o.evals(b)).max() < 1e-12 * scale
        assert np.abs(hm @ V - V * w).max() < 1e-12 * scale
        assert np.abs(V.T @ V - np.eye(len(w))).max() < 1e-12
        assert rel_err(e.weights(b), o.weights(b)) < 1e-10 if o.weights(b).max() > 1e-300 else True


@pytest.mark.parametrize("name", ["anderson_test", "gf4site", "cfg1"])
def test_rotation_matches_reference_formula(pa, oracle, name):
    """O = U_L^T (M U_R) computed on the GPU equals the same product in numpy (GPU eigenvectors),
    and the connected block pairs equal the oracle's find_connections."""
    H, beta = MODELS[name]()
    e, o = build(pa, oracle, H, beta)
    blocks = e.blocks()
    n = len(H.indices())
    for i in range(n):
        gparts = e.field_operator(i, True)
        assert [(p["from"], p["to"]) for p in gparts] == [(p["from"], p["to"]) for p in o.rotate([(True, i)])]
        for p in gparts:
            fb, tb = p["from"], p["to"]
            Uf, Ut = e.evecs(fb), e.evecs(tb)
            X = np.zeros((len(blocks[tb]), len(blocks[fb])))
            pos = {s: k for k, s in enumerate(blocks[tb])}
            for a, s in enumerate(blocks[fb]):
                if (s >> i) & 1:
                    continue
                sign = -1.0 if bin(s & ((1 << i) - 1)).count("1") % 2 else 1.0
                X[pos[s | (1 << i)], :] += sign * Uf[a, :]
            ref = Ut.T @ X
            ref[np.abs(ref) <= 1e-20] = 0.0
            assert np.abs(p["m"] - ref).max() < 1e-13
        cparts = e.field_operator(i, False)  # adjoint view
        cmap = {(p["from"], p["to"]): p["m"] for p in cparts}
        for p in gparts:
            assert np.array_equal(cmap[(p["to"], p["from"])], p["m"].T)


@pytest.mark.parametrize("name", list(MODELS))
def test_greens_function(pa, oracle, name):
    H, beta = MODELS[name]()
    e, o = build(pa, oracle, H, beta)
    z = fermionic(np.arange(-8, 64), beta)
    oi = inject(e, oracle.ED(len(H.indices()), H.mode_terms(), beta))
    for i in range(min(4, len(H.indices()))):
        g = e.gf(i, i, z, 1e-8, 1e-8)
        ref = o.gf(i, i, z, 1e-8, 1e-8)
        assert rel_err(g["values"], ref["values"]) < 1e-10
        same = oi.gf(i, i, z, 1e-8, 1e-8)
        assert g["n_parts"] == same["n_parts"] and g["n_terms"] == same["n_terms"]
        assert rel_err(g["values"], same["values"]) < 1e-12


def test_goldens_on_gpu(pa, oracle):
    g = goldens()
    H, beta = anderson_test_model()
    e = pa.ED(W.terms(H), beta)
    idx = H.indices()
    z = fermionic(np.arange(10), beta)
    for key, spin in (("G_up", 1), ("G_down", 0)):
        i = idx.index(("C", 0, spin))
        assert np.abs(e.gf(i, i, z)["values"] - cplx(g["AndersonTest"][key])).max() < 1e-8
        # the per-frequency operator()(long n) path of GreensFunction
        assert np.abs(e.gf_pointwise(i, i, list(range(10))) - cplx(g["AndersonTest"][key])).max() < 1e-8
    H, beta = anderson2pgf_model()
    e = pa.ED(W.terms(H), beta)
    idx = H.indices()
    p = g["Anderson2PGFTest"]["params"]
    for spin in (1, 0):
        i = idx.index(("C", 0, spin))
        r = e.g2(i, i, i, i, anderson2pgf_freqs(beta), p["pole_resolution"], p["coefficient_tolerance"], None, True)
        ref = np.array(g["Anderson2PGFTest"]["chi_uuuu"])
        assert np.abs(r["values"] - ref).max() < 1e-6
        assert np.abs(r["values_lazy"] - ref).max() < 1e-6  # TwoParticleGF::operator() on kept terms
    e = pa.ED(W.terms(dimer_model()), 1.0)
    assert abs(e.E0 - g["HamiltonianTest"]["E0"]) < 1e-7


def test_hubbard_atom_closed_forms_on_gpu(pa):
    U, mu, beta = 1.0, 0.4, 10.0
    e = pa.ED(W.terms(hubbard_atom(U, mu)), beta)
    n = np.arange(-100, 100)
    z = fermionic(n, beta)
    w0, w1, w2 = 1.0, np.exp(beta * mu), np.exp(-beta * (-2 * mu + U))
    Z = w0 + 2 * w1 + w2
    ref = (w0 + w1) / Z / (z + mu) + (w1 + w2) / Z / (z + mu - U)
    assert np.abs(e.gf(0, 0, z)["values"] - ref).max() < 1e-14
    assert np.abs(e.gf(0, 1, z)["values"]).max() < 1e-14


def test_susceptibility_on_gpu(pa, oracle):
    U, mu, h, beta = 1.0, 0.4, 0.01, 10.0
    H = hubbard_atom(U, mu, -h)
    e, o = build(pa, oracle, H, beta)
    z = bosonic(np.arange(20), beta)
    up, dn = 1, 0
    ops = {"s+": [(True, up), (False, dn)], "s-": [(True, dn), (False, up)], "nu": [(True, up), (False, up)],
           "nd": [(True, dn), (False, dn)]}
    for a, b in [("s+", "s-"), ("nu", "nu"), ("nu", "nd")]:
        v = e.chi(ops[a], ops[b], z, 1e-8, 1e-8, True)
        r = o.chi(ops[a], ops[b], z, 1e-8, 1e-8, True)
        assert np.abs(v["values"] - r["values"]).max() < 1e-14 and v["n_terms"] == r["n_terms"]
    assert abs(e.ensemble_average(ops["nu"]) - o.ensemble_average(ops["nu"])) < 1e-14
    # a bath model: spin and charge susceptibilities
    H, beta = anderson_test_model()
    e, o = build(pa, oracle, H, beta)
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    oi = inject(e, oracle.ED(len(idx), H.mode_terms(), beta))
    for A, B in [([(True, u), (False, d)], [(True, d), (False, u)]), ([(True, u), (False, u)], [(True, d), (False, d)])]:
        v = e.chi(A, B, z, 1e-8, 1e-8, True)
        r = o.chi(A, B, z, 1e-8, 1e-8, True)
        assert rel_err(v["values"], r["values"]) < 1e-10
        for tol in (1e-8, 1e-16):  # merged-term counts, TermList erasures included
            assert e.chi(A, B, z, 1e-8, tol, True)["n_terms"] == oi.chi(A, B, z, 1e-8, tol, True)["n_terms"]


@pytest.mark.parametrize("name", ["anderson2pgf", "cfg1"])
@pytest.mark.parametrize("comp", ["uuuu", "udud", "duud"])
def test_two_particle_gf(pa, oracle, name, comp):
    H, beta = MODELS[name]()
    e, o = build(pa, oracle, H, beta)
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    ijkl = {"uuuu": (u, u, u, u), "udud": (u, d, u, d), "duud": (d, u, u, d)}[comp]
    zs = W.g2_grid(8, 2, beta) if name == "cfg1" else W.g2_grid(6, 3, beta)
    r = e.g2(*ijkl, zs, 1e-8, 1e-16)
    ref = o.g2(*ijkl, zs.ravel(), 1e-8, 1e-16)
    assert r["n_parts"] == ref["n_parts"] and list(r["part_blocks"]) == list(ref["part_blocks"])
    assert rel_err(r["values"], ref["values"]) < 1e-10
    oi = inject(e, oracle.ED(len(idx), H.mode_terms(), beta))
    same = oi.g2(*ijkl, zs.ravel(), 1e-8, 1e-16)
    assert rel_err(r["values"], same["values"]) < 1e-12
    # merged-term counts of every part, erasures by TermList's tol/(size+1) included: exact
    assert list(r["part_nonres"]) == list(same["part_nonres"])
    assert list(r["part_res"]) == list(same["part_res"])


def hub2x2():
    """Periodic 2x2 Hubbard cluster (t = 1, t' = -0.1, U = 4, half filling, beta = 5): a degenerate
    spectrum whose 2PGF merges many terms and erases some by TermList's tol/(size+1) rule."""
    from pomerol_amd.presets import hubbard_cluster

    return hubbard_cluster(2, 2, 1.0, -0.1, 4.0, 2.0), 5.0


@pytest.mark.parametrize("direct", ["1", "0"])
@pytest.mark.parametrize("name,tol,budget", [("hub2x2", 1e-12, None), ("hub2x2", 1e-12, "3000"),
                                             ("hub2x2", 1e-6, None), ("hub2x2", 1e-6, "20000"),
                                             ("anderson2pgf", 1e-12, "500")])
def test_2pgf_erase_rule_exact(pa, oracle, monkeypatch, name, tol, budget, direct):
    """Merged-term counts of every part equal the oracle's (the reference's unordered_set TermList:
    merge on equal pole bins in visiting order, erase when |C| <= tol/(size+1)), with the GPU's
    spectrum injected, for parts whose sorted terms are generated in one chunk and in many
    (POMX_G2_CHUNK_TERMS, a test hook, sets the chunk budget in raw terms), with the terms of
    unique pole bins written directly and with every term sorted (POMX_G2_DIRECT=0, a test hook);
    values agree to 1e-12."""
    monkeypatch.setenv("POMX_G2_DIRECT", direct)
    H, beta = hub2x2() if name == "hub2x2" else MODELS[name]()
    e = pa.ED(W.terms(H), beta)
    idx = H.indices()
    site = "s00" if name == "hub2x2" else "C"
    u, d = idx.index((site, 0, 1)), idx.index((site, 0, 0))
    zs = W.g2_grid(6, 2, beta)
    same = inject(e, oracle.ED(len(idx), H.mode_terms(), beta)).g2(u, d, u, d, zs.ravel(), 1e-8, tol)
    if name == "hub2x2":
        assert same["n_erased"] > 0  # the rule is exercised
    pa.eval_timing(1)  # counts the sorted chunks (g2_sort launches)
    one = e.g2(u, d, u, d, zs, 1e-8, tol)
    l1 = pa.kernel_stats()["g2_sort"]["launches"]
    pa.eval_timing(1)
    if budget:
        monkeypatch.setenv("POMX_G2_CHUNK_TERMS", budget)
    r = e.g2(u, d, u, d, zs, 1e-8, tol)
    if budget:
        monkeypatch.delenv("POMX_G2_CHUNK_TERMS")
    lm = pa.kernel_stats()["g2_sort"]["launches"]
    pa.eval_timing(0)
    print(f"[erase-rule] {name} tol {tol} budget {budget} direct {direct}: sorted chunks {l1} -> {lm}, merged "
          f"{sum(r['part_nonres']) + sum(r['part_res'])}, oracle erased {same['n_erased']}", flush=True)
    if budget and (name == "hub2x2" or direct == "0"):
        assert lm > l1  # several chunks (hub2x2 sorts most of its terms even with direct terms)
    assert list(r["part_blocks"]) == list(same["part_blocks"])
    assert list(r["part_nonres"]) == list(same["part_nonres"])
    assert list(r["part_res"]) == list(same["part_res"])
    assert rel_err(r["values"], same["values"]) < 1e-12
    assert rel_err(r["values"], one["values"]) < 1e-12


def test_2pgf_multichunk_matches_oracle(pa, oracle, monkeypatch):
    """Config 1's parts generated in several chunks (a chunk budget of 20000 raw terms, so its
    large parts take the multi-chunk generation and merge) give the oracle's values and, with the
    GPU's spectrum injected, its merged-term counts part by part."""
    H, beta = MODELS["cfg1"]()
    e, o = build(pa, oracle, H, beta)
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    zs = W.g2_grid(8, 2, beta)
    ref = o.g2(u, d, u, d, zs.ravel(), 1e-8, 1e-16)
    same = inject(e, oracle.ED(len(idx), H.mode_terms(), beta)).g2(u, d, u, d, zs.ravel(), 1e-8, 1e-16)
    monkeypatch.setenv("POMX_G2_DIRECT", "0")  # every term through the sort and its chunks
    pa.eval_timing(1)  # counts the sorted chunks (g2_sort launches)
    one = e.g2(u, d, u, d, zs, 1e-8, 1e-16)
    l1 = pa.kernel_stats()["g2_sort"]["launches"]
    pa.eval_timing(1)
    monkeypatch.setenv("POMX_G2_CHUNK_TERMS", "20000")
    many = e.g2(u, d, u, d, zs, 1e-8, 1e-16)
    monkeypatch.delenv("POMX_G2_CHUNK_TERMS")
    lm = pa.kernel_stats()["g2_sort"]["launches"]
    pa.eval_timing(0)
    monkeypatch.delenv("POMX_G2_DIRECT")
    print(f"[multichunk] sorted chunks {l1} -> {lm}", flush=True)
    assert lm > l1
    assert rel_err(many["values"], ref["values"]) < 1e-10
    assert rel_err(many["values"], one["values"]) < 1e-12
    assert list(many["part_nonres"]) == list(same["part_nonres"])
    assert list(many["part_res"]) == list(same["part_res"])


@pytest.mark.parametrize("name", ["anderson2pgf", "anderson_test"])
def test_2pgf_container_images_match_oracle(pa, oracle, name):
    """TwoParticleGFContainer over all spin combinations of the impurity: every element the
    container returns, computed ones and their IndexContainer4 images (permutations4[1], [6], [7]
    with their frequency permutations and signs, IndexContainer4.hpp:190-237), equals the oracle's
    direct computation of that index combination at the same Matsubara triplets; computeAll
    returns the computed elements on the grid (TwoParticleGFContainer.cpp:33-136)."""
    H, beta = MODELS[name]()
    e = pa.ED(W.terms(H), beta)
    o = oracle.ED(len(H.indices()), H.mode_terms(), beta)
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    combos = [(i, j, k, l) for i in (d, u) for j in (d, u) for k in (d, u) for l in (d, u) if i <= j and k <= l]
    box = pa.G2Container(e, combos, 1e-8, 1e-16)
    ns = [(a, b, c) for a in range(-3, 2) for b in range(-2, 2) for c in range(-2, 3)]
    zs = np.array([[fermionic(a, beta), fermionic(b, beta), fermionic(c, beta)] for a, b, c in ns])
    grid = W.g2_grid(4, 2, beta)
    full = box.compute_all(grid, False)
    assert sorted(full) == sorted(box.elements())
    every = [(i, j, k, l) for i in (d, u) for j in (d, u) for k in (d, u) for l in (d, u)]
    for ijkl in every:
        both = o.g2(*ijkl, np.concatenate([zs, grid]).ravel(), 1e-8, 1e-16)["values"]
        direct, ref = both[:len(zs)], both[len(zs):]
        got = box.value(ijkl, ns)
        assert np.abs(got - direct).max() <= 1e-10 * max(1.0, np.abs(direct).max()), ijkl
        if ijkl in full:
            if len(full[ijkl]) == 0:  # a vanishing element (no parts) comes back empty, as in the reference
                assert np.abs(ref).max() == 0, ijkl
            else:
                assert rel_err(full[ijkl], ref) < 1e-10, ijkl
    # an image is its computed element with permuted frequencies and a sign
    perm, sign = box.permutation((u, d, d, u))
    assert sorted(perm[:3] + [perm[3]]) == [0, 1, 2, 3] and sign in (1, -1)


@pytest.mark.parametrize("name", ["anderson_test", "cfg1"])
def test_imaginary_time(pa, oracle, name):
    """G(tau) and chi(tau) (GreensFunction::of_tau, Susceptibility::of_tau and the parts') on the
    device against the oracle, including both ends of [0, beta]."""
    H, beta = MODELS[name]()
    e, o = build(pa, oracle, H, beta)
    tau = list(np.linspace(0.0, beta, 41))
    z = fermionic(np.arange(4), beta)
    for i in range(2):
        g = e.gf(i, i, z, 1e-8, 1e-8, tau)
        ref = o.gf(i, i, z, 1e-8, 1e-8, tau)
        assert rel_err(g["tau_values"], ref["tau_values"]) < 1e-10
        assert abs(g["tau0_parts"] - g["tau0"]) < 1e-13 and abs(g["tau0"] - ref["tau_values"][0]) < 1e-10
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    zb = bosonic(np.arange(2), beta)
    for A, B in [([(True, u), (False, d)], [(True, d), (False, u)]), ([(True, u), (False, u)], [(True, d), (False, d)])]:
        for sub in (False, True):
            v = e.chi(A, B, zb, 1e-8, 1e-8, sub, tau)
            r = o.chi(A, B, zb, 1e-8, 1e-8, sub, tau)
            assert rel_err(v["tau_values"], r["tau_values"]) < 1e-10
            if not sub:
                assert abs(v["tau0_parts"] - v["tau0"]) < 1e-13


@pytest.mark.parametrize("grid", [(16, 4), (8, 8), (20, 2), (32, 1)])
@pytest.mark.parametrize("comp", ["uuuu", "udud", "duud", "uddu"])
def test_2pgf_grouped_grid_matches_general_path(pa, comp, grid):
    """The Matsubara grid takes the matrix-core kernels (separable and diagonal, the latter with a
    partial second block of 16 rows at 20 fermionic frequencies and two full blocks at 32) and the
    grouped evaluation kernels (aligned groups sharing (z1, z3); with 8 bosonic frequencies also
    triplets a stride apart sharing (z2, z3)); the same triplets in a shuffled order take the
    general kernel.  Both must agree."""
    c = W.cfg1()
    H = c.expr()
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    ijkl = {"uuuu": (u, u, u, u), "udud": (u, d, u, d), "duud": (d, u, u, d), "uddu": (u, d, d, u)}[comp]
    e = pa.ED(W.terms(H), c.beta)
    zs = W.g2_grid(*grid, c.beta)
    order = np.random.default_rng(3).permutation(len(zs))
    grouped = e.g2(*ijkl, zs, 1e-8, 1e-16)["values"]
    general = np.empty_like(grouped)
    general[order] = e.g2(*ijkl, zs[order], 1e-8, 1e-16)["values"]
    assert rel_err(grouped, general) < 1e-12


def test_vertex4_hubbard_atom_closed_form(pa):
    """Vertex4 (Gamma4 = chi - chi0, MatsubaraContainer4) of a Hubbard atom at half filling
    against the closed form Gamma_uuuu G^4 of test/Vertex4Test.cpp:52-58 (absolute 1e-10, as
    there), for every stored triplet with |n| <= 4, plus triplets outside the container."""
    from pomerol_amd.presets import CoulombS

    U, beta, n_iw = 1.0, 40.0, 4
    H = CoulombS("A", U, -U / 2)
    idx = H.indices()
    up = idx.index(("A", 0, 1))
    e = pa.ED(W.terms(H), beta)
    q = [(n1, n2, n3) for n1 in range(-n_iw, n_iw) for n2 in range(-n_iw, n_iw) for n3 in range(-n_iw, n_iw)]
    q_out = [(n_iw + 1, 0, n_iw + 1), (-n_iw - 2, 1, 0)]
    r = e.vertex4(up, up, up, up, n_iw, 1e-4, 1e-12, q + q_out)
    w = lambda n: np.pi * (2 * n + 1) / beta
    G = e.gf(up, up, fermionic(np.arange(-3 * n_iw, 3 * n_iw), beta), 1e-8, 1e-12)["values"]
    g = lambda n: G[n + 3 * n_iw]

    def gamma_ref(n1, n2, n3):
        d = float(n1 == n3) - float(n2 == n3)
        return (-beta * d * (0.5 * U) ** 2 * (1 + (0.5 * U / w(n1)) ** 2) * (1 + (0.5 * U / w(n2)) ** 2)
                * g(n1) * g(n2) * g(n3) * g(n1 + n2 - n3))

    ref = np.array([gamma_ref(*t) for t in q])
    got = r["gamma"][:len(q)]
    print(f"[vertex4] max |Gamma - closed form| {np.abs(got - ref).max():.2e} over {len(q)} triplets", flush=True)
    assert np.abs(got - ref).max() < 1e-10
    assert np.abs(r["gamma"] - r["gamma_value"]).max() < 1e-12  # container == point-by-point value()


@pytest.mark.parametrize("name", ["anderson_test", "cfg1"])
def test_quartic_and_quadratic_rotation(pa, oracle, name):
    """QuarticOperator c+_i c+_j c_k c_l and QuadraticOperator c+_i c_j rotated on the device
    (MonomialOperator.hpp:166-200, MonomialOperatorPart.cpp:48-92) against the oracle's rotation with
    the same eigenvectors: identical block connectivity, elements to 1e-12; and the gauge-invariant
    ensemble average <c+_u c+_d c_d c_u> = <n_u n_d> against the oracle's own eigensolver."""
    H, beta = MODELS[name]()
    e = pa.ED(W.terms(H), beta)
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    o = inject(e, oracle.ED(len(idx), H.mode_terms(), beta))
    quartic = [(True, u), (True, d), (False, d), (False, u)]
    quadratic = [(True, u), (False, d)]
    for op in (quartic, quadratic):
        got = e.rotate(op, 1e-8)
        ref = o.rotate(op, 1e-8)
        assert sorted((p["from"], p["to"]) for p in got) == sorted((p["from"], p["to"]) for p in ref)
        rm = {(p["from"], p["to"]): p["m"] for p in ref}
        for p in got:
            m = rm[(p["from"], p["to"])]
            assert np.abs(p["m"] - m).max() <= 1e-12 * max(1.0, np.abs(m).max())
    own = oracle.ED(len(idx), H.mode_terms(), beta)
    nn = e.ensemble_average(quartic)
    assert abs(nn - own.ensemble_average(quartic)) < 1e-12
    assert abs(nn - own.ensemble_average([(True, u), (False, u), (True, d), (False, d)])) < 1e-12
