"""Benchmark: two-particle GF Lehmann terms x frequencies per second on one MI355X.

Workload (config 2 of the benchmark set): single-orbital Anderson impurity with 5 seeded bath
sites (12 spin-orbitals, 4096 states, largest N/Sz block 400), U = 3, beta = 20, Boltzmann
cutoff 1e-12 (DensityMatrix::truncateBlocks(1e-12) and CoefficientTolerance = 1e-12,
PoleResolution 1e-8), 2PGF <c_up c_dn c^+_up c^+_dn> on the 32 x 32 x 8 Matsubara grid
(8192 frequency triplets), all FP64.

One step = one pass of the TwoParticleGF hot path (term generation, pole merge, Lehmann sum on
all 8192 triplets, TwoParticleGFPart::compute + frequency fill) over a fixed batch of parts of
that component.  The full component holds 206 parts and 6.5e9 merged terms (5e13 terms x
frequencies), so a step is a deterministic sample of its parts (every --stride-th part, which
includes the 400-state blocks); the metric counts exactly the merged terms the timed region
processes, times 8192.

Setup (eigensolve, rotation, part enumeration) is outside the timed region; inputs are resident
in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_NONRES = 28  # complex division 11, products 12, pole subtractions 3, accumulate 2
FLOP_RES = 36
PEAK_FP64_TFLOPS = 78.6
# HBM bytes per evaluation-kernel dispatch of the default step (--stride 8), measured with
# rocprofv3 --pmc: FETCH_SIZE x 2 + WRITE_SIZE x 1 (the factors are what the same counters read
# for a copy kernel of known size); the counters cannot be read from inside this process.
EVAL_TRAFFIC_BYTES = 1.84e9
# Hardware FP64 FLOP per term x frequency of the default step (FMA = 2, one MFMA MOPS unit =
# 512 FLOP), measured with rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 and
# SQ_INSTS_VALU_MFMA_MOPS_F64 over its evaluation dispatches: the matrix-core and grouped kernels
# evaluate fewer operations than the 28/36-FLOP direct formula the metric counts.
EVAL_HW_FLOP_PER_TERM_FREQ = 9.28
DEFAULT_STRIDE = 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def select_parts(parts, stride: int):
    """Deterministic batch: every stride-th part in TwoParticleGF::prepare order (a mix of the
    component's small and 400-state blocks)."""
    return list(range(0, len(parts), stride))


def cpu_sample(batch, counts, budget_terms: float):
    """Parts of the batch for the CPU baseline: smallest non-empty parts first until the merged
    term count reaches budget_terms (about 10-30 s of oracle work at 8192 frequencies)."""
    order = sorted((c, k) for k, c in zip(batch, counts) if c > 0)
    out, acc = [], 0
    for c, k in order:
        if acc + c > budget_terms and out:
            break
        out.append(k)
        acc += c
    return sorted(out)


def cpu_baseline(e, H, beta, zs, ijkl, parts, sample_parts):
    """The oracle (pomerol's algorithm restated on the CPU, OpenMP over frequencies) on a
    bounded sample of parts, fed the same spectrum as the GPU."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pomerol_oracle

    o = pomerol_oracle.ED(len(H.indices()), H.mode_terms(), beta)
    for b in range(e.n_blocks()):
        o.set_spectrum(b, e.evals(b), e.evecs(b).ravel(), e.weights(b))
    o.truncate_blocks(1e-12)
    t0 = time.perf_counter()
    r = o.g2(*ijkl, zs.ravel(), 1e-8, 1e-12, -1, sample_parts)
    secs = time.perf_counter() - t0
    terms = r["n_nonres"] + r["n_res"]
    return terms * len(zs) / secs, secs, terms


def run_batch(args):
    """Batched ED + G(i w_n) solves/s (512 Anderson instances, bath sizes 3-6, 128 frequencies,
    one captured hipGraph, autotuned) against the plain per-solve path (Hamiltonian ->
    FieldOperatorContainer -> GreensFunction through the C++ API, one instance at a time) on a
    sample of the same instances."""
    import pomerol_amd
    from pomerol_amd import workloads as W

    cfgs = W.batch_configs(args.batch_inst)
    terms = [W.terms(c.expr()) for c in cfgs]
    betas = [c.beta for c in cfgs]
    t0 = time.perf_counter()
    b = pomerol_amd.Batch(terms, betas, ("C", 0, 1), 128, 1e-8)
    b.prepare()
    setup = time.perf_counter() - t0
    tune = b.autotune()
    graph = not args.batch_eager
    b.solve(graph)  # capture + warm
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        b.solve(graph)
    secs = (time.perf_counter() - t0) / args.batch_steps
    batched = len(cfgs) / secs
    log(f"[bench] batch: {len(cfgs)} instances, {b.n_blocks()} blocks, setup {setup:.1f}s, {secs * 1e3:.1f} ms/solve-batch, "
        f"tune: {tune}")
    # plain path: a few instances of every bath size, weighted by the batch's own mix of sizes
    per_size = {}
    for k, c in enumerate(cfgs):
        per_size.setdefault(c.n_bath, [])
        if len(per_size[c.n_bath]) < args.plain_sample:
            per_size[c.n_bath].append(k)
    t_size = {}
    for nb, ks in sorted(per_size.items()):
        t0 = time.perf_counter()
        for k in ks:
            e = pomerol_amd.ED(terms[k], betas[k])
            i = [tuple(t) for t in e.indices()].index(("C", 0, 1))
            e.gf(i, i, W.g1_grid(128, betas[k]), 1e-8, 1e-8)
        t_size[nb] = (time.perf_counter() - t0) / len(ks)
    mix = {nb: sum(c.n_bath == nb for c in cfgs) / len(cfgs) for nb in t_size}
    plain_secs = sum(mix[nb] * t_size[nb] for nb in t_size)
    plain = 1.0 / plain_secs
    ns = sum(len(v) for v in per_size.values())
    log(f"[bench] plain path s/solve by bath size: { {nb: round(t, 4) for nb, t in t_size.items()} }, mix {mix}")
    return {"instances": len(cfgs), "freqs": 128, "solves_per_s": batched, "ms_per_batch": secs * 1e3,
            "plain_sample": ns, "plain_solves_per_s": plain, "speedup_vs_plain": batched / plain,
            "plain_s_per_solve_by_bath": t_size, "bath_mix": mix, "autotune": tune, "graph": graph}


def run_dense(args):
    """FP64 eigensolve and GEMM rates through the C ABI on dense device matrices.

    * eigensolve (pomx_eigh_batched, large-block path): one GOE matrix of size --eigh-n (default
      15876, the largest block of the 3x3 Hubbard workload; a random spectrum deflates less than
      the cluster's degenerate one), one warm-up solve then one timed solve; algorithmic work
      (10/3) n^3 (tridiagonalisation 4/3 n^3 + back-transformation 2 n^3, divide-and-conquer
      merges not counted); residual and orthogonality checked on 32 eigenvectors;
    * GEMM (pomx_gemm_batched): C = A B with n = --gemm-n, 2 n^3 FLOP, best of 4.
    Matrices are drawn on the host (numpy) and uploaded; the timed calls start with them in HBM."""
    import ctypes

    import numpy as np

    import pomerol_amd

    lib = ctypes.CDLL(pomerol_amd.POMX_LIB)
    lib.pomx_last_error.restype = ctypes.c_char_p
    ctx = ctypes.c_void_p()
    if lib.pomx_ctx_create(0, ctypes.byref(ctx)) != 0:
        raise RuntimeError("pomx_ctx_create failed")

    def check(rc):
        if rc != 0:
            raise RuntimeError(lib.pomx_last_error(ctx).decode())

    def alloc(nbytes):
        p = ctypes.c_void_p()
        check(lib.pomx_malloc(ctx, ctypes.c_size_t(nbytes), ctypes.byref(p)))
        return p

    def h2d(p, a, off=0):
        check(lib.pomx_h2d(ctx, ctypes.c_void_p(p.value + off), a.ctypes.data_as(ctypes.c_void_p),
                           ctypes.c_size_t(a.nbytes)))

    def d2h(a, p):
        check(lib.pomx_d2h(ctx, a.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(a.nbytes)))

    out = {}
    n = args.eigh_n
    if n > 0:
        rng = np.random.default_rng(12345)
        A0 = rng.standard_normal((n, n))
        A0 += A0.T
        A0 *= 0.5
        d_A, d_w = alloc(A0.nbytes), alloc(8 * n)
        sizes = (ctypes.c_int * 1)(n)
        zero = (ctypes.c_int64 * 1)(0)
        secs = []
        for _ in range(2):
            h2d(d_A, A0)
            t0 = time.perf_counter()
            check(lib.pomx_eigh_batched(ctx, 1, sizes, zero, zero, d_A, d_w))
            secs.append(time.perf_counter() - t0)
        w = np.empty(n)
        d2h(w, d_w)
        V = np.empty((n, n))
        d2h(V, d_A)
        cols = np.arange(0, n, max(1, n // 32))[:32]
        Vs, ws = V[:, cols], w[cols]
        scale = float(np.abs(w).max())
        res = float(np.abs(A0 @ Vs - Vs * ws).max()) / scale
        orth = float(np.abs(Vs.T @ Vs - np.eye(len(cols))).max())
        flops = 10.0 / 3.0 * n ** 3
        out["eigensolve"] = {"n": n, "matrix": "GOE (dense random symmetric)", "seconds": secs[-1],
                             "warmup_seconds": secs[0], "tflops_algorithmic": flops / secs[-1] / 1e12,
                             "flop_model": "(10/3) n^3", "residual": res, "orthogonality": orth,
                             "ascending": bool(np.all(np.diff(w) >= 0))}
        log(f"[bench] eigensolve n={n}: {secs[-1]:.3f}s ({flops / secs[-1] / 1e12:.2f} TFLOP/s), residual {res:.1e}, "
            f"orthogonality {orth:.1e}")
        check(lib.pomx_free(ctx, d_A))
        check(lib.pomx_free(ctx, d_w))
        del A0, V
    n = args.gemm_n
    if n > 0:
        rng = np.random.default_rng(54321)
        X = rng.standard_normal((2, n, n))
        d_X, d_C = alloc(X.nbytes), alloc(8 * n * n)
        h2d(d_X, X)

        class Job(ctypes.Structure):
            _fields_ = [("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int), ("trans_a", ctypes.c_int),
                        ("trans_b", ctypes.c_int), ("lda", ctypes.c_int), ("ldb", ctypes.c_int), ("ldc", ctypes.c_int),
                        ("a_off", ctypes.c_int64), ("b_off", ctypes.c_int64), ("c_off", ctypes.c_int64),
                        ("alpha", ctypes.c_double), ("beta", ctypes.c_double)]

        job = Job(n, n, n, 0, 0, n, n, n, 0, n * n, 0, 1.0, 0.0)
        best = float("inf")
        for _ in range(4):
            t0 = time.perf_counter()
            check(lib.pomx_gemm_batched(ctx, 1, ctypes.byref(job), d_X, d_X, d_C))
            best = min(best, time.perf_counter() - t0)
        C = np.empty((64, n))
        check(lib.pomx_d2h(ctx, C.ctypes.data_as(ctypes.c_void_p), d_C, ctypes.c_size_t(C.nbytes)))
        ref = X[0, :64] @ X[1]
        err = float(np.abs(C - ref).max()) / float(np.abs(ref).max())
        out["gemm"] = {"n": n, "seconds": best, "tflops": 2.0 * n ** 3 / best / 1e12, "rel_err_vs_numpy": err}
        log(f"[bench] gemm n={n}: {best * 1e3:.1f} ms ({2.0 * n ** 3 / best / 1e12:.2f} TFLOP/s), rel err {err:.1e}")
        check(lib.pomx_free(ctx, d_X))
        check(lib.pomx_free(ctx, d_C))
    lib.pomx_ctx_destroy(ctx)
    return out


def run_chi3(args):
    """Three-point susceptibility (PH channel, c+_up c_up c+_dn c_dn) of config 2 on an n x n grid of
    fermionic Matsubara pairs: merged terms x pairs per second of the whole device pipeline
    (generation, merge, evaluation), one warm-up; and the same quantity on config 1 for the device
    and the oracle (the CPU port, all host threads), on a 16 x 16 grid."""
    import numpy as np

    import pomerol_amd
    from pomerol_amd import workloads as W

    def grid(n, beta):
        w = 1j * np.pi * (2 * np.arange(-n // 2, n // 2) + 1) / beta
        return np.array([[a, b] for a in w for b in w])

    out = {}
    for name, c, n in (("config2", W.cfg2(), args.chi3_n), ("config1", W.cfg1(), 16)):
        H = c.expr()
        idx = H.indices()
        u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
        e = pomerol_amd.ED(W.terms(H), c.beta)
        if name == "config2":
            e.truncate_blocks(1e-12)
        tol = 1e-12 if name == "config2" else 1e-16
        zs = grid(n, c.beta)
        e.chi3(1, u, u, d, d, zs, 1e-8, tol)  # warm-up (operators, workspace)
        r = e.chi3(1, u, u, d, d, zs, 1e-8, tol)
        merged = r["n_ff"] + r["n_fb"] + r["n_res"]
        rec = {"pairs": len(zs), "parts": r["n_parts"], "raw_terms": r["n_raw"], "merged_terms": merged,
               "terms_ff_fb_res": [r["n_ff"], r["n_fb"], r["n_res"]], "seconds": r["seconds"],
               "terms_pairs_per_s": merged * len(zs) / r["seconds"], "coeff_tol": tol}
        if name == "config1" and not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pomerol_oracle

            o = pomerol_oracle.ED(len(idx), H.mode_terms(), c.beta)
            t0 = time.perf_counter()
            ro = o.chi3(1, u, u, d, d, zs.ravel(), 1e-8, tol)
            secs = time.perf_counter() - t0
            rec["oracle_seconds"] = secs
            rec["oracle_terms_pairs_per_s"] = (ro["n_ff"] + ro["n_fb"] + ro["n_res"]) * len(zs) / secs
            rec["oracle_cores"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))
            rec["max_rel_diff_vs_oracle"] = float(np.abs(r["values"] - ro["values"]).max() / np.abs(ro["values"]).max())
        log(f"[bench] chi3 {name}: {rec}")
        out[name] = rec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--stride", type=int, default=DEFAULT_STRIDE, help="every stride-th part of the component forms a step")
    ap.add_argument("--cpu-terms", type=float, default=1.5e6, help="merged-term budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--survey", action="store_true", help="print per-part term counts (no frequencies)")
    ap.add_argument("--batch-inst", type=int, default=512, help="instances of the batched ED+G solve (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--plain-sample", type=int, default=4, help="instances per bath size timed on the plain per-solve path")
    ap.add_argument("--eigh-n", type=int, default=15876, help="dense eigensolve size (0: skip)")
    ap.add_argument("--gemm-n", type=int, default=8192, help="dense GEMM size (0: skip)")
    ap.add_argument("--dense-only", action="store_true", help="only the dense eigensolve and GEMM legs")
    ap.add_argument("--batch-only", action="store_true", help="only the batched-solve leg")
    ap.add_argument("--batch-eager", action="store_true", help="batched solves launched eagerly (no graph replay)")
    ap.add_argument("--chi3-n", type=int, default=64, help="3-point susceptibility leg: n x n pairs (0: skip)")
    ap.add_argument("--chi3-only", action="store_true", help="only the 3-point susceptibility leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path's multi-GPU "
              "sharding is not built", file=sys.stderr)
        sys.exit(2)

    if args.dense_only:
        print(json.dumps({"dense": run_dense(args)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batch": run_batch(args)}), flush=True)
        return
    if args.chi3_only:
        print(json.dumps({"chi3": run_chi3(args)}), flush=True)
        return

    import numpy as np

    import pomerol_amd
    from pomerol_amd import workloads as W

    c = W.cfg2()
    H = c.expr()
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    ijkl = (u, d, u, d)
    t0 = time.perf_counter()
    e = pomerol_amd.ED(W.terms(H), c.beta)
    e.truncate_blocks(1e-12)
    zs = W.g2_grid(32, 8, c.beta)
    parts = e.g2_parts(*ijkl)
    log(f"[bench] setup {time.perf_counter() - t0:.2f}s: {e.n_blocks()} blocks, eigh {e.t_eigh:.3f}s, "
        f"rot {e.t_rot:.3f}s, {len(parts)} parts, {len(zs)} frequency triplets")

    if args.survey:
        for k, p in enumerate(parts):
            t = time.perf_counter()
            r = e.g2(*ijkl, zs[:0], 1e-8, 1e-12, [k], False)
            log(f"part {k} blocks {p[:4]} perm {p[4]} dims {p[5:]} raw {r['n_raw']} merged "
                f"{r['n_nonres'] + r['n_res']} ({time.perf_counter() - t:.2f}s)")
        return

    batch = select_parts(parts, args.stride)
    log(f"[bench] batch: {len(batch)} parts {batch}")

    def step():
        return e.g2(*ijkl, zs, 1e-8, 1e-12, batch, False)

    for _ in range(args.warmup):
        step()
    pomerol_amd.eval_timing(1)
    t0 = time.perf_counter()
    terms_nr = terms_r = 0
    for _ in range(args.steps):
        r = step()
        terms_nr += r["n_nonres"]
        terms_r += r["n_res"]
        log(f"[bench] step: {r['n_nonres'] + r['n_res']} merged terms, {r['seconds']:.3f}s")
    wall = time.perf_counter() - t0
    eval_ms, launches = pomerol_amd.eval_timing(0)
    nfreq = len(zs)
    term_freqs = (terms_nr + terms_r) * nfreq
    value = term_freqs / wall
    flops = (FLOP_NONRES * terms_nr + FLOP_RES * terms_r) * nfreq
    achieved = flops / (eval_ms * 1e-3) / 1e12 if eval_ms > 0 else 0.0
    hw = (EVAL_HW_FLOP_PER_TERM_FREQ * term_freqs / (eval_ms * 1e-3) / 1e12
          if eval_ms > 0 and args.stride == DEFAULT_STRIDE else None)

    cpu = None
    if not args.no_cpu:
        counts = [a + b for a, b in zip(r["part_nonres"], r["part_res"])]
        small = cpu_sample(batch, counts, args.cpu_terms)
        log(f"[bench] cpu sample parts {small}")
        v, secs, terms = cpu_baseline(e, H, c.beta, zs, ijkl, parts, small)
        cpu = {"value": v, "unit": "terms*freqs/s", "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
               "kind": "port",
               "sample": f"oracle TwoParticleGFPart compute + fill of {len(small)} config-2 parts ({terms} merged "
                         f"terms x {nfreq} triplets, {secs:.1f}s); generation serial, fill OpenMP"}

    batched = run_batch(args) if args.batch_inst > 0 else None
    dense = run_dense(args) if (args.eigh_n > 0 or args.gemm_n > 0) else None
    chi3 = run_chi3(args) if args.chi3_n > 0 else None

    out = {
        "metric": "2PGF terms*freqs/s",
        "value": value,
        "unit": "terms*freqs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Anderson bath parameters)",
        "config": {"workload": "config 2: Anderson 1+5 bath, 4096 states, U=3, beta=20, 2PGF up-dn-up-dn on "
                               "32x32x8 Matsubara triplets, cutoff 1e-12",
                   "parts_per_step": len(batch), "merged_terms_per_step": (terms_nr + terms_r) // args.steps,
                   "freq_triplets": nfreq},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": EVAL_TRAFFIC_BYTES if args.stride == DEFAULT_STRIDE else None,
                     "traffic_unit": "bytes per launch (rocprofv3 PMC FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "kernel": "g2_eval_mfma_kernel / g2_eval_diag_kernel / g2_eval_s_kernel / g2_eval_kernel (FP64 Lehmann sum)", "eval_ms": eval_ms,
                     "eval_launches": launches,
                     "flop_model": "28 per non-resonant, 36 per resonant term x frequency (direct formula)",
                     "hw_achieved": hw, "hw_frac": hw / PEAK_FP64_TFLOPS if hw else None,
                     "hw_flop_model": f"{EVAL_HW_FLOP_PER_TERM_FREQ} FP64 FLOP per term x frequency (PMC, vector + matrix)"},
        "cpu_baseline": cpu,
        "batch": batched,
        "dense": dense,
        "chi3": chi3,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
