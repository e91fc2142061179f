/* pomx_c.h -- C ABI of the MI355X (gfx950) full-ED kernels.
 *
 * Plain C: opaque context, integer status codes (POMX_OK == 0), plain device pointers and
 * sizes.  No exceptions cross this boundary; pomx_last_error() returns the message of the last
 * failing call on a context.  One context per host thread; every call is stream-ordered on the
 * context's HIP stream and returns after the work it reports (counts, scalars) is complete.
 *
 * The pomerol C++ API (the headers under include/pomerol, implemented in pomerol_amd/src) is
 * the only intended caller; each entry point names the reference routine whose loop it replaces.
 * Every matrix is FP64, dense, row-major unless a pomx_view says otherwise.
 */
#ifndef POMX_C_H
#define POMX_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pomx_ctx pomx_ctx;

enum {
    POMX_OK = 0,
    POMX_ERR_ARG = 1,   /* invalid argument (sizes, null pointers, unsupported option) */
    POMX_ERR_HIP = 2,   /* HIP runtime error (launch, copy, allocation) */
    POMX_ERR_RANGE = 3, /* value outside the range a kernel supports (documented per call) */
    POMX_ERR_NODEV = 4  /* no usable gfx950 device */
};

/* ---- context and memory ------------------------------------------------------------- */
int pomx_ctx_create(int device, pomx_ctx** ctx);
void pomx_ctx_destroy(pomx_ctx* ctx);
const char* pomx_last_error(const pomx_ctx* ctx);
int pomx_malloc(pomx_ctx* ctx, size_t bytes, void** d_ptr);
int pomx_free(pomx_ctx* ctx, void* d_ptr);
int pomx_h2d(pomx_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int pomx_d2h(pomx_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
int pomx_fill_zero(pomx_ctx* ctx, void* d_dst, size_t bytes);
int pomx_synchronize(pomx_ctx* ctx);

/* ---- K1 block assembly ---------------------------------------------------------------
 * Replaces HamiltonianPart::prepareImpl (src/pomerol/HamiltonianPart.cpp:56-76).
 * d_mat[d_pos[e]] += d_val[e] for e < nnz; d_mat (mat_len doubles) is zeroed first. */
int pomx_assemble(pomx_ctx* ctx, int64_t nnz, const int64_t* d_pos, const double* d_val, double* d_mat,
                  int64_t mat_len);

/* ---- K2 batched symmetric eigensolve -------------------------------------------------
 * Replaces HamiltonianPart::computeImpl (src/pomerol/HamiltonianPart.cpp:90-102), i.e. Eigen's
 * SelfAdjointEigenSolver per block.  Block b (size h_n[b]) is the row-major matrix at
 * d_mat + h_mat_off[b]; on return it holds the eigenvectors as columns and d_eval + h_val_off[b]
 * the eigenvalues in ascending order.  Blocks of size <= 128: one workgroup per block
 * (Householder tridiagonalisation, implicit QL, Householder back-transformation), all such
 * blocks concurrently.  Larger blocks, one at a time on the whole device: blocked Householder
 * tridiagonalisation with MFMA trailing updates, divide and conquer on the tridiagonal matrix,
 * blocked (WY) MFMA back-transformation; workspace about 5 n^2 doubles. */
int pomx_eigh_batched(pomx_ctx* ctx, int nblocks, const int* h_n, const int64_t* h_mat_off,
                      const int64_t* h_val_off, double* d_mat, double* d_eval);
/* Divide-and-conquer eigensolver of the symmetric tridiagonal matrix with diagonal h_d (n) and
 * off-diagonal h_e (n-1): eigenvectors as columns of d_Z (n x n row-major, device), eigenvalues
 * ascending into h_w (host).  The tridiagonal stage of the large-block path of
 * pomx_eigh_batched (no separate counterpart in the reference). */
int pomx_eigh_tridiag(pomx_ctx* ctx, int n, const double* h_d, const double* h_e, double* d_Z, double* h_w);

/* ---- K5 Boltzmann weights ------------------------------------------------------------
 * Replaces DensityMatrixPart::computeUnnormalized/normalize (DensityMatrixPart.cpp:23-31) and
 * Hamiltonian::computeGroundEnergy (Hamiltonian.cpp:125-130): E0 = min E, w = exp(-beta(E-E0))/Z. */
int pomx_boltzmann(pomx_ctx* ctx, int64_t nstates, const double* d_E, double beta, double* d_w, double* h_E0,
                   double* h_Z);

/* ---- K3 operator rotation ------------------------------------------------------------
 * Replaces MonomialOperatorPart::computeImpl (MonomialOperatorPart.cpp:48-92):
 *   out[i, j] = sum_t U_to[t, i] * X[t, j],  X[t, j] = sgn(t) * U_from[src(t), j]
 * where d_src[src_off + t] = +-(1 + source inner index) or 0 (no preimage).  Entries with
 * |out| <= cutoff are stored as exact zeros (Eigen's sparseView(Tol) keeps |x| > Tol*1e-12,
 * so the caller passes cutoff = Tol * 1e-12).  FP64 MFMA (v_mfma_f64_16x16x4f64). */
typedef struct {
    int n_to, n_from;
    int64_t u_to_off, u_from_off, src_off, out_off;
} pomx_rot_job;
int pomx_rotate_batched(pomx_ctx* ctx, int njobs, const pomx_rot_job* h_jobs, const double* d_U, const int* d_src,
                        double* d_out, double cutoff);

/* Strided device matrix view: element (r, c) is base[r * rs + c * cs]. */
typedef struct {
    const double* base;
    int64_t rs, cs;
} pomx_view;

/* ---- K6/K10 + K8 two-operator Lehmann terms (G and chi) ------------------------------
 * Replaces GreensFunctionPart::computeImpl (GreensFunctionPart.cpp:65-110) for kind
 * POMX_KIND_GF and SusceptibilityPart::computeImpl (SusceptibilityPart.cpp:76-127) for kind
 * POMX_KIND_CHI, plus the TermList merge (TermList.hpp:65-76): terms are binned by
 * round(pole/pole_res), grouped by (part, bin) with a stable sort, residues summed in
 * emission order, the first pole of a group kept.  Outputs (device, freed with pomx_free):
 * residues and poles of all parts back to back; h_count[p] terms for part p; for CHI also
 * h_zero_pole[p] (the |pole| < pole_res weight). */
enum { POMX_KIND_GF = 0, POMX_KIND_CHI = 1 };
/* Flag OR'ed into the kind of pomx_pair_terms (GF only; ignored for CHI): every element of A and
 * B has magnitude at most 1 (with a relative rounding margin of 1e-9), as for the rotated blocks
 * of a single creation or annihilation operator, whose operator norm is 1.  A candidate can then
 * pass |R| > coeff_tol only if |w_out| + |w_in| > coeff_tol, and the generation skips the pairs of
 * "cold" states (|w| <= coeff_tol / 2 on both sides) without reading them.  The emitted terms
 * are the same as without the flag.  No counterpart in the reference. */
enum { POMX_PAIR_UNIT_ELEMENTS = 0x100 };
typedef struct {
    int n_out, n_in;
    pomx_view A; /* <out|A|in>  n_out x n_in */
    pomx_view B; /* <in|B|out>  n_in x n_out */
    const double *E_out, *E_in, *w_out, *w_in;
} pomx_pair_part;
int pomx_pair_terms(pomx_ctx* ctx, int kind, int nparts, const pomx_pair_part* h_parts, double pole_res,
                    double coeff_tol, double** d_residue, double** d_pole, int64_t* h_count, double* h_zero_pole);

/* ---- K9 Lehmann sum of pole terms on a frequency list --------------------------------
 * GreensFunction::operator() (GreensFunction.hpp:127-136): acc[w] += sum_t R_t / (z_w - P_t);
 * CHI (SusceptibilityPart.hpp:216-220): acc[w] += -sum_t R_t / (z_w - P_t).  z and acc are
 * interleaved complex (re, im) arrays of nz entries. */
int pomx_pair_eval(pomx_ctx* ctx, int kind, int64_t nterms, const double* d_residue, const double* d_pole, int nz,
                   const double* d_z, double* d_acc);

/* As pomx_pair_eval, allocating nothing: the partial sums live in a 16 MiB device array of the
 * library (one per device, set up by pomx_ctx_create, held until the process ends).  No
 * counterpart in the reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  nz <= 1048576
 * (POMX_ERR_RANGE beyond).  Calls that overlap on one device race on the array. */
int pomx_pair_eval_noalloc(pomx_ctx* ctx, int kind, int64_t nterms, const double* d_residue, const double* d_pole,
                           int nz, const double* d_z, double* d_acc);

/* Imaginary-time Lehmann sums on tau in [0, beta] (d_tau: ntau doubles; d_acc: ntau interleaved
 * complex, accumulated, imaginary parts untouched):
 * GF  GreensFunctionPart::Term::operator()(tau, beta) (GreensFunctionPart.cpp:28-31),
 *     acc += sum_t -R_t e^{-tau P_t} / (1 + e^{-beta P_t});
 * CHI SusceptibilityPart::Term::operator()(tau, beta) (SusceptibilityPart.cpp:31-35),
 *     acc += sum_t -R_t e^{-tau P_t} / (e^{-beta P_t} - 1)  (the zero-pole weight is the caller's). */
int pomx_pair_eval_tau(pomx_ctx* ctx, int kind, int64_t nterms, const double* d_residue, const double* d_pole, int ntau,
                       const double* d_tau, double beta, double* d_acc);

/* ---- K7 + K8 + K9 two-particle GF part -----------------------------------------------
 * Replaces TwoParticleGFPart::computeImpl/addMultiterm (TwoParticleGFPart.cpp:189-298), the
 * NonResonant/Resonant TermList merges (TwoParticleGFPart.hpp:83-142,208-268) and the
 * frequency fill (TwoParticleGF.cpp:132-148, TwoParticleGFPart.cpp:309-322).
 * Terms are generated for every (i1,i2,i3,i4) with all four matrix elements non-zero and
 * w1+w2+w3+w4 > coeff_tol, merged on (kind, round(P1/res), round(P2/res), round(P3/res))
 * with count-weighted pole means, and (if nfreq > 0) evaluated at the triplets d_freq
 * (nfreq x 3 complex, the caller's (z1, z2, z3)) into d_acc (nfreq complex, accumulated).
 * h_counts receives the merged term counts (see pomx_g2_terms), h_raw_terms the number of
 * terms emitted before merging.  keep == NULL discards the merged terms after evaluation. */
typedef struct {
    int n1, n2, n3, n4;
    pomx_view O1, O2, O3, CX4; /* <S1|O1|S2> <S2|O2|S3> <S3|O3|S4> <S4|CX4|S1> */
    const double *E1, *E2, *E3, *E4;
    const double *w1, *w2, *w3, *w4;
    int perm;    /* 0..5: permutation of (c_i, c_j, c^dag_k), Misc.cpp:41-42 order */
    double sign; /* its signature */
} pomx_g2_part;

/* Merged terms of one part, packed on the device (freed with pomx_free(ctx, data)).
 * count[0] non-resonant z2 terms, count[1] non-resonant z4 terms (isz4), count[2] resonant
 * z2z3 terms, count[3] resonant z1z2 terms (isz1z2).  Layout: NR terms as SoA
 * (coeff, P1, P2, P3) of n_nr = count[0]+count[1] doubles each, then R terms as SoA
 * (res_coeff, nonres_coeff, P1, P2, P3) of n_r = count[2]+count[3] doubles each. */
typedef struct {
    int64_t count[4];
    double* data;
} pomx_g2_terms;

int pomx_g2_part_compute(pomx_ctx* ctx, const pomx_g2_part* part, double beta, double pole_res, double coeff_tol,
                         int nfreq, const double* d_freq, double* d_acc, pomx_g2_terms* keep, int64_t* h_counts,
                         int64_t* h_raw_terms);
int pomx_g2_terms_eval(pomx_ctx* ctx, const pomx_g2_terms* terms, int perm, double pole_res, int nfreq,
                       const double* d_freq, double* d_acc);

/* ---- K12 + K13 three-point susceptibility part -----------------------------------------
 * Replaces ThreePointSusceptibilityPart::computeImpl/addMultiterm
 * (ThreePointSusceptibilityPart.cpp:243-348), its NonResonantFF/NonResonantFB/Resonant
 * TermLists (ThreePointSusceptibilityPart.hpp:66-370, .cpp:34-149) and the frequency fill
 * (ThreePointSusceptibility.cpp:148-184, ThreePointSusceptibilityPart.cpp:350-359).
 * The part is <S1|Fa|S2><S2|Fb|S3><S3|B1 B2|S1> (Fa = F1, Fb = F2, or swapped Fa = F2, Fb = F1).
 * Terms are generated for every (i1,i2,i3) whose coefficient exceeds coeff_tol, merged on
 * (kind, round(P/res) of each pole) with count-weighted pole means, and (if nfreq > 0)
 * evaluated at the pairs d_freq (nfreq x 2 complex, (z1, z2)) into d_acc (nfreq complex,
 * accumulated).  h_counts receives the merged term counts (see pomx_chi3_terms), h_raw_terms
 * (may be NULL) the number of terms emitted before merging.  keep == NULL discards the terms. */
typedef struct {
    int n1, n2, n3, n4;       /* sizes of S1, S2, S3 and of the block S4 between B1 and B2 */
    pomx_view Fa, Fb, B1, B2; /* <S1|Fa|S2> <S2|Fb|S3> <S3|B1|S4> <S4|B2|S1> */
    const double *E1, *E2, *E3;
    const double *w1, *w2, *w3;
    int channel; /* 0 PP, 1 PH, 2 xPH (Misc.hpp Channel order) */
    int swapped; /* SwappedFermionOps */
} pomx_chi3_part;

/* Merged terms of one part (freed with pomx_free(ctx, data)): count[0] fermion-fermion terms as
 * SoA (coeff, P0, P1), then count[1] fermion-boson terms (coeff, P1, P12), then count[2]
 * resonant terms (coeff, P); every field is count[k] doubles long.  xi = -1 (PP) or 1. */
typedef struct {
    int64_t count[3];
    int xi;
    double* data;
} pomx_chi3_terms;

int pomx_chi3_part_compute(pomx_ctx* ctx, const pomx_chi3_part* part, double beta, double pole_res,
                           double coeff_tol, int nfreq, const double* d_freq, double* d_acc, pomx_chi3_terms* keep,
                           int64_t* h_counts, int64_t* h_raw_terms);
int pomx_chi3_terms_eval(pomx_ctx* ctx, const pomx_chi3_terms* terms, double pole_res, int nfreq,
                         const double* d_freq, double* d_acc);

/* ---- batched FP64 GEMM (matrix cores) --------------------------------------------------
 * C = alpha * op(A) op(B) + beta * C for independent jobs; op(X) = X or X^T (trans_* != 0).
 * Operands are row-major at the given offsets (doubles) into d_A, d_B, d_C, with row strides
 * lda/ldb/ldc.  op(A) is m x k, op(B) is k x n.  Jobs must not write overlapping parts of C.
 * No counterpart in the reference: the dense products of Eigen that this library needs
 * (back-transformations, the adjoint's U^T Ubar and U K) go through this entry. */
typedef struct {
    int m, n, k;
    int trans_a, trans_b;
    int lda, ldb, ldc;
    int64_t a_off, b_off, c_off;
    double alpha, beta;
} pomx_gemm_job;
int pomx_gemm_batched(pomx_ctx* ctx, int njobs, const pomx_gemm_job* h_jobs, const double* d_A, const double* d_B,
                      double* d_C);

/* ---- batched full-ED + G(i w_n) solves (no counterpart in the reference) ---------------
 * Many independent Hamiltonians with uneven sector sizes solved together: block assembly,
 * eigensolve (persistent workgroups pulling blocks from a size-sorted queue), Boltzmann weights
 * per instance, rotation of c^+ for one index, and G(i w_n) = sum R/(i w_n - P) on the first
 * nfreq fermionic Matsubara frequencies of each instance, fused (terms are generated and
 * evaluated on the fly, with the reference's per-term thresholds |R| > coeff_tol and no
 * pole merging).  The whole solve is captured once in a hipGraph and replayed.  All
 * descriptor arrays are host arrays copied at creation; offsets count doubles. */
typedef struct {
    int ninst;   /* instances */
    int nfreq;   /* Matsubara frequencies per instance, n = 0..nfreq-1 */
    int nblocks; /* symmetry blocks of all instances */
    const int* block_n;                /* size of each block */
    const int64_t* block_mat_off;      /* block matrix offset in the matrix arena */
    const int64_t* block_state_off;    /* first state of the block in the state arrays */
    int64_t mat_len, nstates;          /* arena sizes */
    int64_t nnz;                       /* Hamiltonian non-zeros (unique positions) */
    const int64_t* h_pos;              /* position in the matrix arena */
    const double* h_val;
    const double* inst_beta;           /* inverse temperature per instance */
    const int64_t* inst_state_off;     /* states of an instance are contiguous */
    const int64_t* inst_nstates;
    int njobs;                         /* connected (from -> to) block pairs of c^+ */
    const pomx_rot_job* jobs;          /* u_*_off index the matrix arena, out_off the operator arena */
    const int* job_inst;
    const int64_t* job_from_state_off; /* outer block (G: <out|c|in> with out = from block of c^+) */
    const int64_t* job_to_state_off;
    const int* src;                    /* gather maps (pomx_rotate_batched convention) */
    int64_t src_len, ops_len;
    double coeff_tol;
} pomx_batch_desc;

typedef struct pomx_batch pomx_batch;
int pomx_batch_create(pomx_ctx* ctx, const pomx_batch_desc* desc, pomx_batch** batch);
/* One batched solve; use_graph != 0 replays the captured hipGraph (captured on first use). */
int pomx_batch_solve(pomx_ctx* ctx, pomx_batch* batch, int use_graph);
/* h_G: ninst * nfreq interleaved complex; h_E0: ninst ground energies (either may be NULL). */
int pomx_batch_read(pomx_ctx* ctx, pomx_batch* batch, double* h_G, double* h_E0);
/* Autotuner: times the candidate workgroup sizes per block-size bucket and the G work-item
 * sizes on this batch, keeps the fastest and recaptures the graph.  A one-line report of the
 * choices is written to h_report (may be NULL). */
int pomx_batch_tune(pomx_ctx* ctx, pomx_batch* batch, char* h_report, size_t report_len);
void pomx_batch_destroy(pomx_ctx* ctx, pomx_batch* batch);
/* Replace the Hamiltonian values (nnz doubles, in the order of desc->h_val) of every instance;
 * the next pomx_batch_solve uses them (the captured graph reads the same device buffer). */
int pomx_batch_set_values(pomx_ctx* ctx, pomx_batch* batch, const double* h_val);
/* Adjoint of the last pomx_batch_solve (no counterpart in the reference).  h_Gbar holds, per
 * instance and frequency, dL/dRe G + i dL/dIm G (ninst * nfreq interleaved complex); on return
 * h_dval[u] = dL/d(h_val[u]) for every Hamiltonian entry u.  Chain: Lehmann sums -> weights and
 * rotated c^+ -> rotation (recomputed from the eigenvectors) -> eigen-decomposition adjoint
 * (pairs inside a degenerate eigenspace, |E_i - E_j| <= 1e-10 max(1,|E_i|,|E_j|), carry no
 * gradient) -> Hamiltonian entries.  Terms dropped by |R| <= coeff_tol carry no gradient. */
int pomx_batch_adjoint(pomx_ctx* ctx, pomx_batch* batch, const double* h_Gbar, double* h_dval);

/* ---- tangents: first-order response along a Hamiltonian direction ------------------------
 * No counterpart in the reference (forward-mode differentiation of the correlators with respect
 * to one Hamiltonian parameter t, H(t) = H + t dH at t = 0).
 *
 * pomx_eigh_tangent: for blocks already solved by pomx_eigh_batched (eigenvectors as columns of
 * d_U, eigenvalues in d_E) and the direction dH assembled in the same arena layout (d_dH),
 *   A = U^T dH U,  d_dE[i] = A_ii,  d_K = K with K_ij = A_ij / (E_j - E_i) for i != j
 * (dU = U K; K is antisymmetric, zero on the diagonal and between degenerate levels,
 * |E_i - E_j| <= 1e-10 max(1,|E_i|,|E_j|)).  d_K uses the arena layout of d_U.  A rotated operator
 * O = U_L^T M U_R then moves as dO = O K_R - K_L O (pomx_gemm_batched). */
int pomx_eigh_tangent(pomx_ctx* ctx, int nblocks, const int* h_n, const int64_t* h_mat_off, const int64_t* h_val_off,
                      const double* d_U, const double* d_E, const double* d_dH, double* d_K, double* d_dE);
/* Boltzmann weights w = exp(-beta E)/Z (pomx_boltzmann) move as dw = -beta w (dE - sum w dE). */
int pomx_boltzmann_tangent(pomx_ctx* ctx, int64_t nstates, const double* d_w, const double* d_dE, double beta,
                           double* d_dw);
/* Tangent of one two-particle GF part: the derivatives of its rotated operators, eigenvalues and
 * weights (same layouts as the part's own). */
typedef struct {
    pomx_view dO1, dO2, dO3, dCX4;
    const double *dE1, *dE2, *dE3, *dE4;
    const double *dw1, *dw2, *dw3, *dw4;
} pomx_g2_tangent;
/* Value and derivative of a two-particle GF part at the triplets d_freq: the terms are generated,
 * thresholded and merged exactly as by pomx_g2_part_compute (the kept set and the groups are
 * held fixed), each merged group carrying the derivative of its summed coefficients and of its
 * mean poles; d_acc += value, d_dacc += derivative (nfreq complex each). */
int pomx_g2_part_tangent(pomx_ctx* ctx, const pomx_g2_part* part, const pomx_g2_tangent* tangent, double beta,
                         double pole_res, double coeff_tol, int nfreq, const double* d_freq, double* d_acc,
                         double* d_dacc, int64_t* h_counts, int64_t* h_raw_terms);

/* ---- timing of the dominant kernel (HIP events on the context stream) ---------------- */
/* Total time and count of the Lehmann-sum evaluations (every class below except
 * POMX_STAT_G2_GEN) since the last call with enable >= 0; enable >= 0 also resets the
 * statistics and switches them on (1) or off (0). */
int pomx_eval_timing(pomx_ctx* ctx, int enable, double* h_total_ms, int64_t* h_launches);

/* Per-class statistics recorded while timing is on (HIP events around each launch group on the
 * context stream, resolved when read; no synchronisation per launch).  flop: the work model of
 * the algorithm the class runs (FMA = 2, add/mul = 1, reciprocal estimate = 1, one f64 MFMA
 * 16x16x4 = 2048); term_freqs: merged terms x frequencies evaluated; bytes: algorithmic HBM bytes
 * (terms read once, outputs written once).  No counterpart in the reference. */
enum {
    POMX_STAT_G2_SEP = 0,     /* 2PGF, separable grids, matrix cores (g2_eval_mfma_kernel + reduce) */
    POMX_STAT_G2_DIAG = 1,    /* 2PGF, diagonal resonant terms, matrix cores (g2_eval_diag_kernel + reduce) */
    POMX_STAT_G2_GROUPED = 2, /* 2PGF, grouped Matsubara grids, vector (g2_eval_s_kernel) */
    POMX_STAT_G2_GENERAL = 3, /* 2PGF, any triplets, vector (g2_eval_kernel) */
    POMX_STAT_G2_REDUCE = 4,  /* 2PGF, per-slice partial sums into the grid (g2_eval_reduce_kernel) */
    POMX_STAT_PAIR = 5,       /* G / chi Lehmann sums (pair_eval_kernel + reduce) */
    POMX_STAT_G2_GEN = 6,     /* 2PGF term generation and merge of a part, up to its evaluation */
    POMX_STAT_G2_SORT = 7,    /* 2PGF: one sorted chunk (terms with shared pole bins) emitted, sorted, grouped */
    POMX_STAT_NCLASS = 8
};
typedef struct {
    int64_t launches;
    double ms, flop, term_freqs, bytes;
} pomx_kernel_stat;
int pomx_kernel_stats(pomx_ctx* ctx, pomx_kernel_stat* h_stats /* POMX_STAT_NCLASS entries */);

#ifdef __cplusplus
}
#endif
#endif /* POMX_C_H */
