"""Derivative of the up-down two-particle GF chi_{u d u d}(w1, w2; w3) with respect to the
Hubbard U of the Anderson impurity (mu = U/2 tied to it, so the direction is
dH/dU = n_up n_dn - (n_up + n_dn)/2), checked against central finite differences, h = 1e-6:

* spectrum: dE of every block against differences of the device eigenvalues, dw against
  differences of the Boltzmann weights;
* 1+2 bath model against differences of the oracle's 2PGF;
* 1+3 bath model (U = 2, beta = 10) on the 16x16x4 grid against differences of the device 2PGF
  (which matches the oracle at rel 1e-10 in the parity tests), with the tangent's value equal to
  the plain 2PGF;
* a 1+5 bath instance of the bath-fit workload (U ~ U[2,4], beta = 20) on the 16x16x4 grid,
  against differences of the device 2PGF, with the cost of the tangent bounded by 3x the
  forward's.

The finite-difference checks on the small models use CoefficientTolerance 0: with a per-term
cutoff the 2PGF jumps wherever a term crosses it, and a difference quotient across a jump
measures the jump.
"""
import time

import numpy as np
import pytest
from helpers import rel_err

from pomerol_amd import workloads as W
from pomerol_amd.presets import CoulombS, anderson

pytestmark = pytest.mark.gpu

H_STEP = 1e-6
DIRECTION = W.terms(CoulombS("C", 1.0, -0.5))  # dH/dU with mu = U/2


def _model(cfg, U):
    return anderson(cfg.eps, cfg.V, U, U / 2)


def _ud(H):
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    return (u, d, u, d)


def _fd_device(pa, cfg, zs, pr=1e-8):
    vals = []
    for s in (1, -1):
        H = _model(cfg, cfg.U + s * H_STEP)
        e = pa.ED(W.terms(H), cfg.beta)
        vals.append(e.g2(*_ud(H), zs, pr, 0.0)["values"])
    return (vals[0] - vals[1]) / (2 * H_STEP)


def _check(name, der, fd):
    err = np.abs(der - fd) / max(1.0, np.max(np.abs(fd)))
    print(f"[g2-tangent] {name}: max |d chi/dU| {np.max(np.abs(fd)):.4e}, max err vs FD {err.max():.2e}", flush=True)
    assert err.max() <= 1e-6


def test_spectrum_tangent_matches_finite_differences(pa):
    cfg = W.cfg1()
    H = _model(cfg, cfg.U)
    e = pa.ED(W.terms(H), cfg.beta)
    tan = e.spectrum_tangent(DIRECTION)
    ep = pa.ED(W.terms(_model(cfg, cfg.U + H_STEP)), cfg.beta)
    em = pa.ED(W.terms(_model(cfg, cfg.U - H_STEP)), cfg.beta)
    err_e = err_w = 0.0
    for b, (dE, dw) in enumerate(tan):
        fdE = (ep.evals(b) - em.evals(b)) / (2 * H_STEP)
        fdw = (ep.weights(b) - em.weights(b)) / (2 * H_STEP)
        err_e = max(err_e, np.max(np.abs(np.sort(dE) - np.sort(fdE))))
        err_w = max(err_w, np.max(np.abs(dw - fdw)))
    print(f"[g2-tangent] spectrum of 1+3 bath: max |dE - FD| {err_e:.2e}, max |dw - FD| {err_w:.2e}", flush=True)
    assert err_e <= 1e-7 and err_w <= 1e-7


def test_g2_tangent_matches_oracle_finite_differences(pa, oracle):
    cfg = W.AndersonConfig(2, 2.5, 8.0, seed=W.SEED + 501)
    zs = W.g2_grid(8, 3, cfg.beta)
    H = _model(cfg, cfg.U)
    e = pa.ED(W.terms(H), cfg.beta)
    r = e.g2_tangent(*_ud(H), zs, DIRECTION, 1e-8, 0.0)
    vals = []
    for s in (1, -1):
        Hs = _model(cfg, cfg.U + s * H_STEP)
        o = oracle.ED(len(Hs.indices()), Hs.mode_terms(Hs.indices()), cfg.beta)
        vals.append(o.g2(*_ud(Hs), zs.ravel(), 1e-8, 0.0)["values"])
    o0 = oracle.ED(len(H.indices()), H.mode_terms(H.indices()), cfg.beta)
    ref = o0.g2(*_ud(H), zs.ravel(), 1e-8, 0.0)["values"]
    assert rel_err(r["values"], ref) <= 1e-10
    _check("1+2 bath vs oracle FD", r["tangent"], (vals[0] - vals[1]) / (2 * H_STEP))


def test_two_particle_u_autograd(pa):
    """torch.autograd.Function of the tangent: dL/dU of L = sum |chi(U) - chi_t|^2 against central
    differences of L from the device 2PGF."""
    import torch

    from pomerol_amd.adjoint import AndersonTwoParticle, TwoParticleU

    cfg = W.AndersonConfig(2, 2.5, 8.0, seed=W.SEED + 502)
    zs = W.g2_grid(8, 3, cfg.beta)
    model = AndersonTwoParticle(cfg.eps, cfg.V, cfg.beta, zs, 1e-8, 0.0)
    target = torch.from_numpy(model.value_and_tangent(3.0)[0])
    U = torch.tensor(cfg.U, dtype=torch.float64, requires_grad=True)
    L = ((TwoParticleU.apply(U, model) - target).abs() ** 2).sum()
    L.backward()

    def loss(u):
        return float(np.sum(np.abs(model.value_and_tangent(u)[0] - target.numpy()) ** 2))

    fd = (loss(cfg.U + H_STEP) - loss(cfg.U - H_STEP)) / (2 * H_STEP)
    print(f"[g2-tangent] autograd dL/dU {float(U.grad):.10e}, FD {fd:.10e}", flush=True)
    assert abs(float(U.grad) - fd) <= 1e-6 * max(1.0, abs(fd))


def test_g2_tangent_config_1plus3_bath(pa):
    cfg = W.cfg1()
    zs = W.g2_grid(16, 4, cfg.beta)
    H = _model(cfg, cfg.U)
    e = pa.ED(W.terms(H), cfg.beta)
    r = e.g2_tangent(*_ud(H), zs, DIRECTION, 1e-8, 0.0)
    plain = e.g2(*_ud(H), zs, 1e-8, 0.0)["values"]
    assert rel_err(r["values"], plain) <= 1e-12
    print(f"[g2-tangent] 1+3 bath: {r['n_nonres']} + {r['n_res']} merged terms, tangent {r['seconds']:.3f}s", flush=True)
    _check("1+3 bath 16x16x4 vs device FD", r["tangent"], _fd_device(pa, cfg, zs))


@pytest.mark.timeout(900)
def test_g2_tangent_5bath_instance_and_cost(pa):
    """First instance of the bath-fit workload.  The cost ratio is measured at CoefficientTolerance
    1e-12 (2.2e9 raw terms), the finite differences at 1e-16 (1.1e10 raw terms): at 1e-12 a few
    thousand terms cross the cutoff within +-h and their jumps dominate the difference quotient
    (its error against the tangent grows as h shrinks, 3e-6 at h = 1e-4 to 1e-5 at 1e-6)."""
    cfgs, _ = W.adjoint_configs(64)
    cfg = cfgs[0]
    zs = W.g2_grid(16, 4, cfg.beta)
    H = _model(cfg, cfg.U)
    e = pa.ED(W.terms(H), cfg.beta)
    print(f"[g2-tangent] 1+5 bath U={cfg.U:.3f}: warm-up forward", flush=True)
    e.g2(*_ud(H), zs, 1e-8, 1e-12)  # operators rotated, kernels loaded
    t0 = time.perf_counter()
    plain = e.g2(*_ud(H), zs, 1e-8, 1e-12)["values"]
    t_fwd = time.perf_counter() - t0
    print(f"[g2-tangent] forward {t_fwd:.3f}s; tangent", flush=True)
    t0 = time.perf_counter()
    r = e.g2_tangent(*_ud(H), zs, DIRECTION, 1e-8, 1e-12)
    t_tan = time.perf_counter() - t0
    print(f"[g2-tangent] 1+5 bath, cutoff 1e-12: {r['n_raw']} raw terms, forward {t_fwd:.3f}s, value+tangent "
          f"{t_tan:.3f}s, ratio {t_tan / t_fwd:.2f}", flush=True)
    assert rel_err(r["values"], plain) <= 1e-12
    assert t_tan <= 3.0 * t_fwd
    r = e.g2_tangent(*_ud(H), zs, DIRECTION, 1e-8, 1e-16)
    print(f"[g2-tangent] 1+5 bath, cutoff 1e-16: {r['n_raw']} raw terms, value+tangent {r['seconds']:.3f}s", flush=True)
    fd = []
    for s in (1, -1):
        Hs = _model(cfg, cfg.U + s * H_STEP)
        es = pa.ED(W.terms(Hs), cfg.beta)
        fd.append(es.g2(*_ud(Hs), zs, 1e-8, 1e-16)["values"])
        print(f"[g2-tangent] forward at U {'+' if s > 0 else '-'} h done", flush=True)
    _check(f"1+5 bath U={cfg.U:.3f} vs device FD", r["tangent"], (fd[0] - fd[1]) / (2 * H_STEP))
