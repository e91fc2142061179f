"""Batched ED + G(i w_n) solves (one captured hipGraph) against the plain per-solve API path."""
import numpy as np
import pytest
from helpers import rel_err

from pomerol_amd import workloads as W

pytestmark = pytest.mark.gpu


def instances(n, seed=5):
    cfgs = W.batch_configs(n, seed=seed)
    return cfgs, [W.terms(c.expr()) for c in cfgs], [c.beta for c in cfgs]


def plain_g(pa, terms, beta, nfreq):
    e = pa.ED(terms, beta)
    idx = e.indices()
    i = [tuple(t) for t in idx].index(("C", 0, 1))
    return e.gf(i, i, W.g1_grid(nfreq, beta), 1e-8, 1e-8)["values"], e.E0


def test_batch_matches_plain_path(pa):
    cfgs, terms, betas = instances(10)
    print("[batch] bath sizes", [c.n_bath for c in cfgs], flush=True)
    b = pa.Batch(terms, betas, ("C", 0, 1), 128, 1e-8)
    b.solve(True)
    G, E0 = b.G(), b.E0()
    assert G.shape == (10, 128)
    for k in range(10):
        g_ref, e0_ref = plain_g(pa, terms[k], betas[k], 128)
        assert abs(E0[k] - e0_ref) < 1e-12 * max(1, abs(e0_ref))
        assert rel_err(G[k], g_ref) < 1e-10, k
    # eager launch sequence = graph replay
    b.solve(False)
    assert np.array_equal(b.G(), G)


def test_batch_graph_replays_and_autotune(pa):
    cfgs = [c for c in W.batch_configs(64, seed=9) if c.n_bath <= 4][:16]
    terms, betas = [W.terms(c.expr()) for c in cfgs], [c.beta for c in cfgs]
    print("[batch] replay/autotune test, bath sizes", [c.n_bath for c in cfgs], flush=True)
    b = pa.Batch(terms, betas, ("C", 0, 1), 64, 1e-8)
    b.solve(True)
    first = b.G().copy()
    for _ in range(3):  # replays with eager work in between
        plain_g(pa, terms[0], betas[0], 8)
        b.solve(True)
        assert np.array_equal(b.G(), first)
    report = b.autotune()
    print("[batch] autotune:", report, flush=True)
    assert "g_terms_per_item" in report
    b.solve(True)
    assert rel_err(b.G(), first) < 1e-12  # only the summation grouping changed
    assert b.n_structures() == len({c.n_bath for c in cfgs})


def test_batch_config4_instances_match_oracle(pa, oracle):
    """Instances of the 512-instance workload itself (bath sizes 3-6, CoefficientTolerance 1e-8,
    128 frequencies) from one batched solve against the oracle (Householder + implicit QR per
    block, the algorithm of the reference's Eigen solver; GreensFunction terms and TermList merge)."""
    cfgs = W.batch_configs(512)
    pick = []
    for nb in (3, 4, 5, 6):
        pick += [k for k, c in enumerate(cfgs) if c.n_bath == nb][:2 if nb < 6 else 1]
    sub = [cfgs[k] for k in pick]
    print("[batch] config-4 instances", pick, "bath sizes", [c.n_bath for c in sub], flush=True)
    b = pa.Batch([W.terms(c.expr()) for c in sub], [c.beta for c in sub], ("C", 0, 1), 128, 1e-8)
    b.solve(True)
    G, E0 = b.G(), b.E0()
    for k, c in enumerate(sub):
        H = c.expr()
        idx = H.indices()
        i = idx.index(("C", 0, 1))
        o = oracle.ED(len(idx), H.mode_terms(), c.beta, "qr")
        g = o.gf(i, i, W.g1_grid(128, c.beta), 1e-8, 1e-8)["values"]
        err = rel_err(G[k], g)
        print(f"[batch] instance {pick[k]} (bath {c.n_bath}, beta {c.beta:.1f}, U {c.U:.2f}): rel err {err:.1e}", flush=True)
        assert abs(E0[k] - o.E0) < 1e-12 * max(1.0, abs(o.E0))
        assert err < 1e-10
