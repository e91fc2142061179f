"""Adjoint of the batched G(i w_n) solve (bath-fit gradients), checked three ways:

* against torch.linalg.eigh autograd of an independent block-diagonal torch restatement of the
  same forward (tiny instances, every parameter);
* against central finite differences (h = 1e-6) of the CPU oracle's G (4 bath sites);
* against central finite differences of the device forward for 64 instances with 5 bath sites,
  with the cost of the backward pass bounded by 3x the forward pass.

The finite-difference checks use CoefficientTolerance 0: with the reference's per-term cutoff
|R| > 1e-8 the loss jumps wherever a term crosses the cutoff (at beta = 20 and ~7e5 terms per
instance some term does so within +-h for a few of the 704 parameters), and a difference
quotient across a jump measures the jump, not the derivative.  The torch check keeps the cutoff:
both sides drop the same terms.
"""
import time

import numpy as np
import pytest
import torch

import brute
from pomerol_amd import workloads as W
from pomerol_amd.presets import anderson

pytestmark = pytest.mark.gpu

IDX = ("C", 0, 1)


def _solver(cfgs, nfreq, tol=1e-8):
    from pomerol_amd.adjoint import GSolver, anderson_operators, anderson_theta

    ops = [anderson_operators(c.n_bath) for c in cfgs]
    theta = np.array([anderson_theta(c) for c in cfgs])
    return GSolver(ops, theta, [c.beta for c in cfgs], IDX, nfreq, tol), theta


def _loss_and_grad(solver, theta, Gt):
    from pomerol_amd.adjoint import LehmannG, bath_fit_loss

    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    G = LehmannG.apply(th, solver)
    L = bath_fit_loss(G, torch.from_numpy(Gt))
    L.backward()
    return float(L), th.grad.numpy().copy(), G.detach().numpy()


def _torch_reference(n_bath, theta, beta, nfreq, tol, Gt):
    """Block-diagonal torch restatement of the forward; returns (G, dL/dtheta) by autograd."""
    from pomerol_amd.adjoint import anderson_operators  # noqa: F401 (same operator convention)
    from pomerol_amd.presets import CoulombS, Hopping, Level

    ops = [Level(f"b{k}", 1.0) for k in range(n_bath)] + [Hopping("C", f"b{k}", 1.0) for k in range(n_bath)]
    ops.append(CoulombS("C", 1.0, -0.5))
    full = anderson(np.ones(n_bath), np.ones(n_bath), 1.0, 0.5)
    idx = full.indices()
    nm = len(idx)
    Hp = [torch.from_numpy(brute.hamiltonian(nm, o.mode_terms(idx))) for o in ops]
    cd = torch.from_numpy(brute.op_matrix(nm, [(True, idx.index(IDX))]))
    up = [k for k, t in enumerate(idx) if t[2] == 1]
    dn = [k for k, t in enumerate(idx) if t[2] == 0]
    sector = {}
    for s in range(1 << nm):
        key = (sum((s >> k) & 1 for k in up), sum((s >> k) & 1 for k in dn))
        sector.setdefault(key, []).append(s)
    keys = sorted(sector)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    H = sum(th[p] * Hp[p] for p in range(len(ops)))
    E, U = {}, {}
    for key in keys:
        ii = torch.tensor(sector[key])
        E[key], U[key] = torch.linalg.eigh(H[ii][:, ii])
    Eall = torch.cat([E[k] for k in keys])
    wall = torch.softmax(-beta * Eall, 0)
    w, at = {}, 0
    for key in keys:
        w[key] = wall[at:at + len(sector[key])]
        at += len(sector[key])
    z = torch.from_numpy(W.g1_grid(nfreq, beta))
    G = torch.zeros(nfreq, dtype=torch.complex128)
    for (nu, nd) in keys:  # c^+_up: (nu, nd) -> (nu + 1, nd)
        t = (nu + 1, nd)
        if t not in sector:
            continue
        f = (nu, nd)
        M = cd[torch.tensor(sector[t])][:, torch.tensor(sector[f])]
        A = U[t].T @ M @ U[f]
        R = A * A * (w[t][:, None] + w[f][None, :])
        P = E[t][:, None] - E[f][None, :]
        keep = ((A != 0) & (R.abs() > tol)).detach()
        Rk, Pk = R[keep], P[keep]
        G = G + (Rk[None, :].to(torch.complex128) / (z[:, None] - Pk[None, :])).sum(1)
    L = ((G - torch.from_numpy(Gt)).abs() ** 2).sum()
    L.backward()
    return G.detach().numpy(), th.grad.numpy()


def test_adjoint_matches_torch_eigh_autograd(pa):
    cfgs = [W.AndersonConfig(2, U, beta, seed=W.SEED + 300 + k) for k, (U, beta) in enumerate([(2.0, 10.0), (3.5, 25.0),
                                                                                               (1.2, 6.0)])]
    nfreq = 32
    solver, theta = _solver(cfgs, nfreq)
    rng = np.random.default_rng(5)
    Gt = solver.forward(theta + 0.05 * rng.standard_normal(theta.shape))
    L, g, G = _loss_and_grad(solver, theta, Gt)
    print(f"[adjoint] tiny: L={L:.6e}", flush=True)
    for k, c in enumerate(cfgs):
        Gref, gref = _torch_reference(c.n_bath, theta[k], c.beta, nfreq, 1e-8, Gt[k])
        assert np.max(np.abs(G[k] - Gref)) <= 1e-12 * np.max(np.abs(Gref))
        err = np.max(np.abs(g[k] - gref)) / max(1.0, np.max(np.abs(gref)))
        print(f"[adjoint] instance {k}: grad {g[k][:3]} ... rel err vs torch eigh autograd {err:.2e}", flush=True)
        assert err <= 1e-9


def _oracle_loss(oracle, cfg, theta, Gt, nfreq):
    nb = cfg.n_bath
    H = anderson(theta[:nb], theta[nb:2 * nb], theta[2 * nb], theta[2 * nb] / 2)
    idx = H.indices()
    e = oracle.ED(len(idx), H.mode_terms(idx), cfg.beta)
    i = idx.index(IDX)
    G = e.gf(i, i, W.g1_grid(nfreq, cfg.beta), 1e-8, 0.0)["values"]
    return float(np.sum(np.abs(G - Gt) ** 2))


def test_adjoint_matches_oracle_finite_differences(pa, oracle):
    cfg = W.AndersonConfig(4, 2.5, 15.0, seed=W.SEED + 401)
    tgt = W.AndersonConfig(4, 3.0, 15.0, seed=W.SEED + 402)
    nfreq, h = 128, 1e-6
    solver, theta = _solver([cfg], nfreq, tol=0.0)
    from pomerol_amd.adjoint import anderson_theta

    Gt = solver.forward(anderson_theta(tgt)[None, :])
    L, g, _ = _loss_and_grad(solver, theta, Gt)
    L0 = _oracle_loss(oracle, cfg, theta[0], Gt[0], nfreq)
    assert abs(L - L0) <= 1e-10 * max(1.0, abs(L0))
    fd = np.empty(theta.shape[1])
    for p in range(theta.shape[1]):
        tp, tm = theta[0].copy(), theta[0].copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (_oracle_loss(oracle, cfg, tp, Gt[0], nfreq) - _oracle_loss(oracle, cfg, tm, Gt[0], nfreq)) / (2 * h)
    err = np.abs(g[0] - fd)
    print(f"[adjoint] oracle FD: grad {g[0]}\n[adjoint] fd   {fd}\n[adjoint] max abs err {err.max():.2e}", flush=True)
    assert np.all(err <= 1e-6 * np.maximum(1.0, np.abs(fd)))


def test_adjoint_batch_of_64_finite_differences_and_cost(pa):
    from pomerol_amd.adjoint import anderson_theta

    cfgs, targets = W.adjoint_configs(64)
    nfreq, h = 128, 1e-6
    solver, theta = _solver(cfgs, nfreq, tol=0.0)
    Gt = solver.forward(np.array([anderson_theta(t) for t in targets]))
    L, g, _ = _loss_and_grad(solver, theta, Gt)  # warm (graph captured)
    # cost: forward (parameters, graph replay, G read) vs backward (adjoint)
    t0 = time.perf_counter()
    for _ in range(3):
        G = solver.forward(theta)
    t_fwd = (time.perf_counter() - t0) / 3
    gbar = 2.0 * (G - Gt)
    t0 = time.perf_counter()
    for _ in range(3):
        g2 = solver.backward(gbar)
    t_bwd = (time.perf_counter() - t0) / 3
    print(f"[adjoint] 64 x 5-bath: forward {t_fwd * 1e3:.1f} ms, backward {t_bwd * 1e3:.1f} ms, "
          f"ratio {t_bwd / t_fwd:.2f}", flush=True)
    assert np.max(np.abs(g - g2)) <= 1e-12 * max(1.0, np.max(np.abs(g)))
    assert t_bwd <= 3.0 * t_fwd
    # central differences of the device forward, one parameter of every instance at a time
    # (instances are independent, so the per-instance losses separate)
    def per_inst_loss(th):
        Gx = solver.forward(th)
        return np.sum(np.abs(Gx - Gt) ** 2, axis=1)

    fd = np.empty_like(theta)
    for p in range(theta.shape[1]):
        tp, tm = theta.copy(), theta.copy()
        tp[:, p] += h
        tm[:, p] -= h
        fd[:, p] = (per_inst_loss(tp) - per_inst_loss(tm)) / (2 * h)
    err = np.abs(g - fd) / np.maximum(1.0, np.abs(fd))
    print(f"[adjoint] 64 x 5-bath FD: max rel err {err.max():.2e} (param {np.unravel_index(err.argmax(), err.shape)})",
          flush=True)
    assert err.max() <= 1e-6


def test_adjoint_hubbard_cluster_hoppings(pa):
    """Gradients with respect to hoppings: 2x2 periodic Hubbard clusters with theta = (t, t', U)
    (mu = U/2), G_00(i w_n) of the up spin, against central differences of the device forward."""
    from pomerol_amd.adjoint import GSolver, LehmannG, bath_fit_loss
    from pomerol_amd.presets import CoulombS, Expr, hubbard_cluster

    u_op = Expr()
    for s in range(4):
        u_op += CoulombS(f"s{s:02d}", 1.0, -0.5)
    ops = [W.terms(hubbard_cluster(2, 2, 1.0, 0.0, 0.0, 0.0)), W.terms(hubbard_cluster(2, 2, 0.0, 1.0, 0.0, 0.0)),
           W.terms(u_op)]
    rng = np.random.default_rng(31)
    n, nfreq, h, beta = 4, 64, 1e-6, 5.0
    theta = np.stack([rng.uniform(0.8, 1.2, n), rng.uniform(-0.3, 0.0, n), rng.uniform(2.0, 5.0, n)], 1)
    solver = GSolver([ops] * n, theta, [beta] * n, ("s00", 0, 1), nfreq, 0.0)
    Gt = solver.forward(theta + 0.1 * rng.standard_normal(theta.shape))
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    L = bath_fit_loss(LehmannG.apply(th, solver), torch.from_numpy(Gt))
    L.backward()
    g = th.grad.numpy()

    def per_inst_loss(x):
        return np.sum(np.abs(solver.forward(x) - Gt) ** 2, axis=1)

    fd = np.empty_like(theta)
    for p in range(3):
        tp, tm = theta.copy(), theta.copy()
        tp[:, p] += h
        tm[:, p] -= h
        fd[:, p] = (per_inst_loss(tp) - per_inst_loss(tm)) / (2 * h)
    err = np.abs(g - fd) / np.maximum(1.0, np.abs(fd))
    print(f"[adjoint] 2x2 Hubbard (t, t', U): grad {g[0]}, fd {fd[0]}, max rel err {err.max():.2e}", flush=True)
    assert err.max() <= 1e-6
