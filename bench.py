"""Benchmark: two-particle GF Lehmann terms x frequencies per second on one MI355X.

Workload (config 2 of the benchmark set): single-orbital Anderson impurity with 5 seeded bath
sites (12 spin-orbitals, 4096 states, largest N/Sz block 400), U = 3, beta = 20, Boltzmann
cutoff 1e-12 (DensityMatrix::truncateBlocks(1e-12) and CoefficientTolerance = 1e-12,
PoleResolution 1e-8); the full 2PGF for all spin combinations of the impurity: a
TwoParticleGFContainer over every (i <= j, k <= l) of {dn, up} (three computed elements
dn-dn-dn-dn, dn-up-dn-up, up-up-up-up with 582 parts and 1.97e10 merged terms; their images and
the six combinations that vanish by spin conservation), on the 32 x 32 x 8 Matsubara grid
(8192 frequency triplets (W + w3, w2, w3)), all FP64.

A step is one slice of that container: the 582 (element, part) pairs are dealt to --slices (20)
slices of equal size by the largest-first rule, and step i runs slice i mod 20 through the
container's elements (TwoParticleGF::computeParts: term generation, pole merge, Lehmann sum on all
8192 triplets, terms cleared after the fill).  20 consecutive steps are the whole container once
(the default and the driver's --steps 20); the metric counts exactly the merged terms of the
timed steps, times 8192.  --slices 1 makes every step the whole container.

Setup (eigensolve, rotation, part enumeration) is outside the timed region; inputs are resident
in HBM.  Prints one JSON line.  Further objects: the roofline of the dominant kernel from the
library's per-class work model and HIP-event timing, the CPU baseline (the oracle, pomerol's
algorithm restated on the CPU, on the largest affordable parts, extrapolated to the container),
parity of that sample against the oracle, the batched ED+G solve (config 4) against the plain
path and the oracle, the FP64 eigensolve and GEMM rates, the bath-fit gradient (config 5)
against the oracle's finite differences, and the 3-point susceptibility.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector and matrix (dense)
PEAK_HBM_GBS = 8000.0
DEFAULT_SLICES = 20
# HBM traffic per dispatch of the dominant kernel on this workload, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (copy-calibrated; the command and the factors are in the file)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "g2_traffic_cfg2.json")
EVAL_CLASSES = ("g2_sep", "g2_diag", "g2_grouped", "g2_general", "g2_reduce")
_FLOP_RULES = "work model of the algorithm the kernel runs (FMA 2, add/mul 1, v_rcp_f64 1, Newton reciprocal 5)"
FLOP_MODEL = {
    "g2_sep": _FLOP_RULES + ": per term and bosonic frequency 6 NR NC (complex GEMM from three real MFMA products) "
                            "+ row and column elements (about 18 NR + 13 NC) + 2",
    "g2_diag": _FLOP_RULES + ": per term, bosonic frequency and 16-row tile 6 x 16 x (NC + 15) (complex product "
                             "with the diagonal values, three real MFMA products) + 18 per row element + 13 per "
                             "diagonal value + 3",
    "g2_grouped": _FLOP_RULES + ": per term and frequency 13.5 (|V|^2, a quarter of a shared Newton reciprocal, "
                                "K conj(V) / |V|^2 accumulated) + 20-27 per term and thread of 8 frequencies",
    "g2_general": _FLOP_RULES + ": per term and frequency 20.75 on purely imaginary triplets (one shared "
                                "reciprocal per four), else 28 / 31 / 32 (the direct formula)",
}
KERNEL_OF_CLASS = {"g2_sep": "g2_eval_mfma_kernel", "g2_diag": "g2_eval_diag_kernel",
                   "g2_grouped": "g2_eval_s_kernel", "g2_general": "g2_eval_kernel",
                   "g2_reduce": "g2_eval_reduce_kernel"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cfg2_container(pa, W):
    c = W.cfg2()
    H = c.expr()
    idx = H.indices()
    u, d = idx.index(("C", 0, 1)), idx.index(("C", 0, 0))
    e = pa.ED(W.terms(H), c.beta)
    e.truncate_blocks(1e-12)
    combos = [(i, j, k, l) for i in (d, u) for j in (d, u) for k in (d, u) for l in (d, u) if i <= j and k <= l]
    box = pa.G2Container(e, combos, 1e-8, 1e-12)
    return c, H, e, box


def slice_plan(box, nslices):
    """(element, part) pairs dealt to nslices slices, largest n1 n2 n3 n4 first, each to the
    currently lightest slice (ties: lowest slice).  Returns [{element: [parts]}]."""
    work = []
    for k in range(len(box.elements())):
        for p, desc in enumerate(box.parts(k)):
            n1, n2, n3, n4 = desc[5:9]
            work.append((n1 * n2 * n3 * n4, k, p))
    work.sort(key=lambda x: (-x[0], x[1], x[2]))
    load = [0] * nslices
    slices = [dict() for _ in range(nslices)]
    for w, k, p in work:
        s = min(range(nslices), key=lambda i: (load[i], i))
        load[s] += w
        slices[s].setdefault(k, []).append(p)
    for sl in slices:
        for k in sl:
            sl[k].sort()
    return slices, len(work)


def run_slice(box, sl, zs, acc, counts):
    merged = 0
    for k, parts in sorted(sl.items()):
        r = box.compute_parts(k, parts, zs)
        if len(zs):
            acc[k] += r["values"]
        for p, a, b, raw in zip(parts, r["part_nonres"], r["part_res"], r["part_raw"]):
            counts[(k, p)] = (int(a) + int(b), int(raw))
            merged += int(a) + int(b)
    return merged


def load_traffic(kernel):
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    rec = t.get("kernels", {}).get(kernel)
    if not rec:
        return None
    return {"bytes_per_dispatch": rec["bytes_per_dispatch"], "fetch_bytes": rec.get("fetch_bytes"),
            "write_bytes": rec.get("write_bytes"), "source": os.path.relpath(TRAFFIC_FILE, ROOT),
            "measured": t.get("measured")}


def roofline(stats, wall_s):
    """The dominant evaluation kernel (most time) with its work-model FLOP rate, and every class."""
    classes = {}
    for name, s in stats.items():
        if s["launches"] == 0:
            continue
        secs = s["ms"] * 1e-3
        rec = {"launches": s["launches"], "seconds": secs, "share_of_timed": secs / wall_s, "flop": s["flop"],
               "term_freqs": s["term_freqs"], "bytes": s["bytes"]}
        if s["flop"] > 0 and secs > 0:
            rec["tflops"] = s["flop"] / secs / 1e12
            rec["frac_fp64_peak"] = rec["tflops"] / PEAK_FP64_TFLOPS
        if name not in ("g2_gen", "g2_sort") and s["term_freqs"] > 0 and secs > 0:
            rec["term_freqs_per_s"] = s["term_freqs"] / secs
        if s["bytes"] > 0 and secs > 0:
            rec["algorithmic_GBs"] = s["bytes"] / secs / 1e9
        classes[name] = rec
    evals = [n for n in EVAL_CLASSES if n in classes and n != "g2_reduce"]
    if not evals:
        return None, classes
    dom = max(evals, key=lambda n: classes[n]["seconds"])
    s = stats[dom]
    secs = s["ms"] * 1e-3
    achieved = s["flop"] / secs / 1e12
    traffic = load_traffic(KERNEL_OF_CLASS[dom])
    out = {"bound": "mfma" if dom in ("g2_sep", "g2_diag") else "valu", "achieved": achieved,
           "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
           "traffic": traffic["bytes_per_dispatch"] if traffic else None,
           "kernel": KERNEL_OF_CLASS[dom], "launches": s["launches"], "avg_launch_ms": s["ms"] / s["launches"],
           "flop_per_launch": s["flop"] / s["launches"],
           "algorithmic_bytes_per_launch": s["bytes"] / s["launches"],
           "flop_model": FLOP_MODEL[dom],
           "traffic_unit": "HBM bytes per dispatch (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; copy-calibrated)",
           "traffic_source": traffic}
    all_eval = sum(stats[n]["ms"] for n in EVAL_CLASSES) * 1e-3
    all_flop = sum(stats[n]["flop"] for n in EVAL_CLASSES)
    out["all_evaluation"] = {"seconds": all_eval, "share_of_timed": all_eval / wall_s,
                             "tflops": all_flop / all_eval / 1e12 if all_eval > 0 else None,
                             "frac_fp64_peak": all_flop / all_eval / 1e12 / PEAK_FP64_TFLOPS if all_eval > 0 else None}
    return out, classes


def cpu_baseline(pa, e, H, beta, box, zs, counts, args):
    """The oracle (pomerol's algorithm restated on the CPU: TermList merge, addMultiterm, OpenMP
    frequency fill), fed the device spectrum, on the largest parts of each computed element with at
    most --cpu-part-terms merged terms, on all 8192 triplets; the same parts on the device for
    parity; and the container's CPU time extrapolated from the sample: generation by raw-term
    count, the frequency fill by merged-term count."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pomerol_oracle

    o = pomerol_oracle.ED(len(H.indices()), H.mode_terms(), beta)
    for b in range(e.n_blocks()):
        o.set_spectrum(b, e.evals(b), e.evecs(b).ravel(), e.weights(b))
    o.truncate_blocks(1e-12)
    sample = {}
    for k, el in enumerate(box.elements()):
        cand = [(m, p) for (kk, p), (m, raw) in counts.items() if kk == k and m <= args.cpu_part_terms]
        if cand:
            sample[k] = [max(cand)[1]]
    gen_s = eval_s = 0.0
    s_merged = s_raw = 0
    max_rel = 0.0
    counts_equal = True
    for k, parts in sorted(sample.items()):
        el = box.elements()[k]
        r = o.g2(*el, zs.ravel(), 1e-8, 1e-12, -1, parts)
        g = box.compute_parts(k, parts, zs)
        gen_s += r["gen_seconds"]
        eval_s += r["eval_seconds"]
        s_merged += r["n_nonres"] + r["n_res"]
        s_raw += sum(counts[(k, p)][1] for p in parts)
        counts_equal = counts_equal and list(g["part_nonres"]) == list(r["part_nonres"]) and \
            list(g["part_res"]) == list(r["part_res"])
        max_rel = max(max_rel, float(np.abs(g["values"] - r["values"]).max() / np.abs(r["values"]).max()))
        log(f"[bench] cpu sample {el} parts {parts}: {r['n_nonres'] + r['n_res']} terms, gen {r['gen_seconds']:.2f}s, "
            f"fill {r['eval_seconds']:.2f}s")
    tot_merged = sum(m for m, raw in counts.values())
    tot_raw = sum(raw for m, raw in counts.values())
    nf = len(zs)
    t_serial = gen_s * tot_raw / s_raw + eval_s * tot_merged / s_merged
    nc = cores()
    # the reference farms parts over MPI ranks: with one rank per core, generation runs nc parts at
    # once (ideal scaling assumed) while each fill is single-threaded; the fill timed here used nc
    # OpenMP threads, so it already is the per-core rate times nc
    t_farm = gen_s * tot_raw / s_raw / nc + eval_s * tot_merged / s_merged
    return {"value": tot_merged * nf / t_farm, "unit": "terms*freqs/s", "cores": nc, "kind": "port",
            "sample": f"oracle on {len(sample)} parts ({s_merged} merged terms, {s_raw} raw) x {nf} triplets: "
                      f"generation {gen_s:.1f}s (one thread), fill {eval_s:.1f}s ({nc} OpenMP threads); "
                      f"extrapolated to the container ({tot_merged} merged, {tot_raw} raw terms) by raw terms "
                      f"(generation) and merged terms (fill), generation farmed over {nc} cores",
            "sample_rate": s_merged * nf / (gen_s + eval_s),
            "extrapolated_container_seconds": t_farm,
            "extrapolated_container_seconds_serial_generation": t_serial,
            "parity": {"parts": {str(box.elements()[k]): v for k, v in sample.items()},
                       "max_rel_diff_vs_oracle": max_rel, "merged_counts_equal": counts_equal}}


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
    out = {"instances": len(cfgs), "freqs": 128, "solves_per_s": batched, "ms_per_batch": secs * 1e3,
           "plain_sample": ns, "plain_solves_per_s": plain, "speedup_vs_plain": batched / plain,
           "plain_s_per_solve_by_bath": t_size, "bath_mix": mix, "autotune": tune, "graph": graph}
    if not args.no_cpu:
        # the oracle (Householder + implicit QR per block, the algorithm of the reference's Eigen
        # solver; blocks over the host cores; rotation and G terms as GreensFunctionPart) on one
        # instance of every bath size, weighted by the batch's mix; its G against the batch's
        import numpy as np

        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pomerol_oracle

        G = b.G()
        c_size, rel = {}, 0.0
        for nb, ks in sorted(per_size.items()):
            k = ks[0]
            H = cfgs[k].expr()
            idx = H.indices()
            i = idx.index(("C", 0, 1))
            t0 = time.perf_counter()
            o = pomerol_oracle.ED(len(idx), H.mode_terms(), betas[k], "qr")
            g = o.gf(i, i, W.g1_grid(128, betas[k]), 1e-8, 1e-8)["values"]
            c_size[nb] = time.perf_counter() - t0
            rel = max(rel, float(np.abs(G[k] - g).max() / np.abs(g).max()))
        cpu_secs = sum(mix[nb] * c_size[nb] for nb in c_size)
        out["cpu_baseline"] = {"solves_per_s": 1.0 / cpu_secs, "cores": cores(), "kind": "port",
                               "s_per_solve_by_bath": c_size,
                               "sample": "oracle ED (Householder + implicit QR per block, blocks over the host cores) "
                                         "+ G on 128 frequencies, one instance per bath size, weighted by the mix",
                               "max_rel_diff_batch_vs_oracle": rel}
        log(f"[bench] batch cpu baseline: {out['cpu_baseline']}")
    return out


def run_cfg3(args):
    """Config 3 end to end on the GPU: the 3x3 periodic Hubbard cluster (18 spin-orbitals,
    262144 states, 100 N/Sz blocks up to 15876), t = 1, seeded t', U = 4, half filling, beta = 5.
    Hamiltonian blocks and eigensolve (timed), truncateBlocks(1e-8), the rotated c / c^+ of the
    --cfg3-sites up-spin sites (the down-spin ones are their spin images), G_ij(i w_n) for every
    pair of them on 128 fermionic frequencies, and the transverse spin susceptibility
    <S+_0 ; S-_0>(i nu_m) on 32 bosonic frequencies (CoefficientTolerance 1e-8, PoleResolution
    1e-8 as the reference's defaults).

    CPU baseline (the oracle, kind "port"): its Householder + implicit-QR eigensolver
    (HamiltonianPart.cpp:98) timed on the leading n_s x n_s block of the largest Hamiltonian block
    and extrapolated to every block by n^3; its operator rotation (MonomialOperatorPart.cpp:48-92)
    timed on config 2's c^+ operators and extrapolated by the rotation FLOP count
    (2 n_L^2 n_R per connected block pair)."""
    import numpy as np

    import pomerol_amd as pa
    from pomerol_amd import workloads as W

    H, beta, tp = W.cfg3_expr()
    out = {"model": f"3x3 periodic Hubbard, t=1, t'={tp:.4f}, U=4, mu=2, beta={beta}", "states": 1 << 18}
    t0 = time.perf_counter()
    # (the largest Hamiltonian block is kept on the host for the CPU sample; that copy is not ED)
    e = pa.ED(W.terms(H), beta, "largest" if not args.no_cpu else False)
    out["ed_seconds"] = time.perf_counter() - t0 - e.t_keep
    out["ed_host_copy_seconds"] = e.t_keep
    out["eigensolve_seconds"] = e.t_eigh
    sizes = [len(b) for b in e.blocks()]
    out["blocks"] = len(sizes)
    out["largest_block"] = max(sizes)
    e.truncate_blocks(1e-8)
    out["retained_blocks"] = int(sum(e.retained()))
    log(f"[bench] cfg3: ED {out['ed_seconds']:.2f}s (eigensolve {e.t_eigh:.2f}s), {len(sizes)} blocks, "
        f"retained {out['retained_blocks']}")
    idx = [tuple(x) for x in e.indices()]
    ns = max(1, min(9, args.cfg3_sites))
    up = [idx.index((f"s{k:02d}", 0, 1)) for k in range(ns)]
    e.use_indices(up)
    zf = W.g1_grid(128, beta)
    t0 = time.perf_counter()
    G = np.empty((ns, ns, len(zf)), complex)
    nterms = 0
    for a, i in enumerate(up):
        for b, j in enumerate(up):
            r = e.gf(i, j, zf, 1e-8, 1e-8)
            G[a, b] = r["values"]
            nterms += int(r["n_terms"])
    out["g_seconds"] = time.perf_counter() - t0
    out["rotation_seconds"] = e.t_rot
    out["g_pairs"] = ns * ns
    out["g_terms"] = nterms
    out["g_symmetric_max_diff"] = float(np.abs(G - G.transpose(1, 0, 2)).max())
    d0 = idx.index(("s00", 0, 0))
    zb = W.bosonic(np.arange(32), beta)
    t0 = time.perf_counter()
    x = e.chi([(True, up[0]), (False, d0)], [(True, d0), (False, up[0])], zb, 1e-8, 1e-8)
    out["chi_seconds"] = time.perf_counter() - t0
    out["chi_terms"] = int(x["n_terms"])
    out["chi_static"] = float(np.real(x["values"][0]))
    log(f"[bench] cfg3: {ns} sites, G_ij ({ns * ns} pairs, {nterms} terms) {out['g_seconds']:.2f}s incl. rotation "
        f"{e.t_rot:.2f}s; chi S+S- {out['chi_seconds']:.2f}s ({out['chi_terms']} terms)")
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pomerol_oracle as o

        b = int(np.argmax(sizes))
        n_s = min(max(sizes), args.cfg3_cpu_n)
        A = np.asarray(e.hmat(b))[:n_s, :n_s].copy()
        t0 = time.perf_counter()
        o.ED.eigh(A, "qr")
        t_s = time.perf_counter() - t0
        t_eigh = t_s * sum((n / n_s) ** 3 for n in sizes)
        # rotation rate on config 2's c^+ operators (whole ED built first, not timed)
        c2 = W.cfg2()
        H2 = c2.expr()
        ed2 = o.ED(len(H2.indices()), H2.mode_terms(), c2.beta, "qr")
        flop = 0.0
        t0 = time.perf_counter()
        for i in range(len(H2.indices())):
            for part in ed2.rotate([(True, i)], 1e-8):
                nl, nr = part["m"].shape
                flop += 2.0 * nl * nl * nr
        t_rot2 = time.perf_counter() - t0
        rate = flop / t_rot2
        # the config-3 rotation of the ns c^+ operators: every (n_up, n_dn) block pair they connect
        spin = {k: t[2] for k, t in enumerate(idx)}
        sector = {}
        for bl in e.blocks():
            st = bl[0]
            nu = sum(1 for k in range(len(idx)) if (st >> k) & 1 and spin[k] == 1)
            sector[(nu, sum(1 for k in range(len(idx)) if (st >> k) & 1 and spin[k] == 0))] = len(bl)
        flop3 = 0.0
        for (nu, nd), n_r in sector.items():
            n_l = sector.get((nu + 1, nd), 0)
            flop3 += ns * 2.0 * n_l * n_l * n_r
        out["cpu_baseline"] = {
            "eigensolve_seconds_extrapolated": t_eigh, "rotation_seconds_extrapolated": flop3 / rate,
            "cores": 1, "kind": "port",
            "sample": f"oracle Householder + implicit QR on the leading {n_s} x {n_s} block of the largest "
                      f"({max(sizes)}) block: {t_s:.2f}s, scaled by n^3 over the {len(sizes)} blocks; oracle rotation "
                      f"of config 2's {len(H2.indices())} c^+ operators ({flop:.3g} FLOP, {t_rot2:.2f}s, "
                      f"{rate / 1e9:.2f} GFLOP/s) scaled to the {ns} config-3 operators' {flop3:.3g} FLOP",
            "gpu_vs_cpu_eigensolve": t_eigh / max(e.t_eigh, 1e-9)}
        log(f"[bench] cfg3 cpu: {out['cpu_baseline']}")
    return out


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


def run_grad(args):
    """Bath-fit gradient (config 5): dL/d(eps_k, V_k, U) of L = sum_n |G(i w_n) - G_target(i w_n)|^2
    over 128 frequencies for 64 Anderson instances with 5 bath sites, by the device adjoint, against
    the forward pass on the device and against central finite differences (h = 1e-6) of the oracle on
    the host (one instance, extrapolated to the batch).  CoefficientTolerance 0, as in the tests:
    with a per-term cutoff the loss jumps where a term crosses it, and a difference quotient across a
    jump is not a derivative."""
    import numpy as np

    from pomerol_amd import workloads as W
    from pomerol_amd.adjoint import GSolver, anderson_operators, anderson_theta

    cfgs, targets = W.adjoint_configs(args.grad_inst)
    nfreq, h = 128, 1e-6
    ops = [anderson_operators(c.n_bath) for c in cfgs]
    theta = np.array([anderson_theta(c) for c in cfgs])
    solver = GSolver(ops, theta, [c.beta for c in cfgs], ("C", 0, 1), nfreq, 0.0)
    Gt = solver.forward(np.array([anderson_theta(t) for t in targets]))
    G = solver.forward(theta)  # warm: graph captured
    t0 = time.perf_counter()
    for _ in range(3):
        G = solver.forward(theta)
    t_fwd = (time.perf_counter() - t0) / 3
    gbar = 2.0 * (G - Gt)
    g = solver.backward(gbar)
    t0 = time.perf_counter()
    for _ in range(3):
        g = solver.backward(gbar)
    t_bwd = (time.perf_counter() - t0) / 3
    out = {"instances": len(cfgs), "parameters_per_instance": theta.shape[1], "freqs": nfreq,
           "forward_s": t_fwd, "backward_s": t_bwd, "backward_over_forward": t_bwd / t_fwd}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pomerol_oracle

        from pomerol_amd.presets import anderson

        c0 = cfgs[0]
        z = W.g1_grid(nfreq, c0.beta)

        def loss(th):
            nb = c0.n_bath
            Hx = anderson(th[:nb], th[nb:2 * nb], th[-1], th[-1] / 2)
            idx = Hx.indices()
            i = idx.index(("C", 0, 1))
            o = pomerol_oracle.ED(len(idx), Hx.mode_terms(), c0.beta, "qr")
            gx = o.gf(i, i, z, 1e-8, 0.0)["values"]
            return float(np.sum(np.abs(gx - Gt[0]) ** 2))

        t0 = time.perf_counter()
        fd = np.empty(theta.shape[1])
        for p in range(theta.shape[1]):
            tp, tm = theta[0].copy(), theta[0].copy()
            tp[p] += h
            tm[p] -= h
            fd[p] = (loss(tp) - loss(tm)) / (2 * h)
        t_fd = time.perf_counter() - t0
        err = float(np.max(np.abs(g[0] - fd) / np.maximum(1.0, np.abs(fd))))
        out["cpu_baseline"] = {"fd_seconds_per_instance": t_fd, "fd_seconds_batch_extrapolated": t_fd * len(cfgs),
                               "cores": cores(), "kind": "port",
                               "sample": f"oracle central differences (h = {h}) of instance 0: "
                                         f"{2 * theta.shape[1]} ED + G solves (Householder + implicit QR)",
                               "max_rel_diff_adjoint_vs_fd": err}
    log(f"[bench] grad: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--slices", type=int, default=DEFAULT_SLICES, help="slices of the container (a step is one)")
    ap.add_argument("--cpu-part-terms", type=float, default=1.6e6,
                    help="CPU baseline: per computed element, the largest part with at most this many merged terms")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--survey", action="store_true", help="print per-part term counts (no frequencies)")
    ap.add_argument("--batch-inst", type=int, default=512, help="instances of the batched ED+G solve (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--plain-sample", type=int, default=4, help="instances per bath size timed on the plain per-solve path")
    ap.add_argument("--eigh-n", type=int, default=15876, help="dense eigensolve size (0: skip)")
    ap.add_argument("--gemm-n", type=int, default=8192, help="dense GEMM size (0: skip)")
    ap.add_argument("--grad-inst", type=int, default=64, help="bath-fit gradient instances (0: skip)")
    ap.add_argument("--dense-only", action="store_true", help="only the dense eigensolve and GEMM legs")
    ap.add_argument("--batch-only", action="store_true", help="only the batched-solve leg")
    ap.add_argument("--batch-eager", action="store_true", help="batched solves launched eagerly (no graph replay)")
    ap.add_argument("--chi3-n", type=int, default=64, help="3-point susceptibility leg: n x n pairs (0: skip)")
    ap.add_argument("--chi3-only", action="store_true", help="only the 3-point susceptibility leg")
    ap.add_argument("--grad-only", action="store_true", help="only the bath-fit gradient leg")
    ap.add_argument("--cfg3-sites", type=int, default=9, help="config-3 leg: up-spin sites of G_ij (0: skip the leg)")
    ap.add_argument("--cfg3-cpu-n", type=int, default=2000, help="config-3 CPU eigensolve sample size")
    ap.add_argument("--cfg3-only", action="store_true", help="only the config-3 leg")
    ap.add_argument("--g2-only", action="store_true", help="only the 2PGF leg (and its CPU baseline)")
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
    if args.grad_only:
        print(json.dumps({"grad": run_grad(args)}), flush=True)
        return
    if args.cfg3_only:
        print(json.dumps({"cfg3": run_cfg3(args)}), flush=True)
        return

    import numpy as np

    import pomerol_amd as pa
    from pomerol_amd import workloads as W

    t0 = time.perf_counter()
    c, H, e, box = cfg2_container(pa, W)
    zs = W.g2_grid(32, 8, c.beta)
    nslices = max(1, args.slices)
    slices, nparts = slice_plan(box, nslices)
    log(f"[bench] setup {time.perf_counter() - t0:.2f}s: {e.n_blocks()} blocks, eigh {e.t_eigh:.3f}s, "
        f"rot {e.t_rot:.3f}s, elements {box.elements()}, {nparts} parts in {nslices} slices, {len(zs)} triplets")

    if args.survey:
        for k, el in enumerate(box.elements()):
            for p, desc in enumerate(box.parts(k)):
                t = time.perf_counter()
                r = box.compute_parts(k, [p], zs[:0])
                log(f"element {el} part {p} blocks {desc[:4]} perm {desc[4]} dims {desc[5:]} raw {r['part_raw'][0]} "
                    f"merged {r['part_nonres'][0] + r['part_res'][0]} ({time.perf_counter() - t:.2f}s)")
        return

    nel = len(box.elements())
    counts = {}
    warm = {k: np.zeros(len(zs), complex) for k in range(nel)}
    for i in range(args.warmup):
        run_slice(box, slices[i % nslices], zs, warm, counts)
    acc = {k: np.zeros(len(zs), complex) for k in range(nel)}
    pa.eval_timing(1)
    t0 = time.perf_counter()
    merged = 0
    step_s = []
    for i in range(args.steps):
        ts = time.perf_counter()
        m = run_slice(box, slices[i % nslices], zs, acc, counts)
        step_s.append(time.perf_counter() - ts)
        merged += m
        log(f"[bench] step {i} (slice {i % nslices}): {m} merged terms, {step_s[-1]:.3f}s")
    wall = time.perf_counter() - t0
    stats = pa.kernel_stats()
    pa.eval_timing(0)
    nfreq = len(zs)
    value = merged * nfreq / wall
    roof, classes = roofline(stats, wall)
    # counts of the parts no step reached (generation only), for the container totals
    for k in range(nel):
        rest = [p for p in range(len(box.parts(k))) if (k, p) not in counts]
        if rest:
            r = box.compute_parts(k, rest, zs[:0])
            for p, a, b_, raw in zip(rest, r["part_nonres"], r["part_res"], r["part_raw"]):
                counts[(k, p)] = (int(a) + int(b_), int(raw))
    tot_merged = sum(m for m, raw in counts.values())
    container = {"elements": [list(x) for x in box.elements()],
                 "computed": [list(box.elements()[k]) for k in range(nel) if len(box.parts(k))],
                 "parts": nparts, "merged_terms": tot_merged, "raw_terms": sum(raw for m, raw in counts.values()),
                 "slices": nslices}
    if args.steps >= nslices:
        container["pass_seconds"] = sum(step_s[:nslices])
        container["pass_term_freqs_per_s"] = tot_merged * nfreq / container["pass_seconds"]
    log(f"[bench] container {container}")
    log(f"[bench] kernel classes {classes}")

    cpu = None if args.no_cpu else cpu_baseline(pa, e, H, c.beta, box, zs, counts, args)
    if cpu:
        log(f"[bench] cpu baseline {cpu}")
    g2_only = args.g2_only
    batched = run_batch(args) if args.batch_inst > 0 and not g2_only else None
    dense = run_dense(args) if (args.eigh_n > 0 or args.gemm_n > 0) and not g2_only else None
    chi3 = run_chi3(args) if args.chi3_n > 0 and not g2_only else None
    grad = run_grad(args) if args.grad_inst > 0 and not g2_only else None
    cfg3 = run_cfg3(args) if args.cfg3_sites > 0 and not g2_only else None

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
        "config": {"workload": "config 2: Anderson 1+5 bath, 4096 states, U=3, beta=20, cutoff 1e-12; full 2PGF "
                               "container over all spin combinations (3 computed elements, 582 parts) on 32x32x8 "
                               f"Matsubara triplets; a step is one of {nslices} equal slices of the container's parts, "
                               f"{nslices} consecutive steps are the whole container",
                   "merged_terms_per_step": merged // max(1, args.steps), "freq_triplets": nfreq,
                   "container": container},
        "roofline": roof,
        "kernel_classes": classes,
        "cpu_baseline": cpu,
        "batch": batched,
        "dense": dense,
        "chi3": chi3,
        "grad": grad,
        "cfg3": cfg3,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
