// Device-driven dense symmetric eigensolver for large blocks, many blocks at once.
//
// Replaces Eigen::SelfAdjointEigenSolver in HamiltonianPart::computeImpl
// (src/pomerol/HamiltonianPart.cpp:90-102) for the blocks that do not fit one workgroup.  Every
// stage works on all blocks of a plan together (blockIdx.y = block, tree node or job), the launch
// sequence depends only on the block sizes and every data-dependent quantity (deflation counts,
// permutations, GEMM sizes) lives in device memory, so a plan can be replayed inside a hipGraph
// (in a captured stream everything runs on that stream; outside capture, two-stage plans run
// the next panel's LQ on a second stream of their own, joined by events).
//   1. Blocked Householder tridiagonalisation T = Q^T A Q, panels of kNB reflectors.  Inside a
//      panel the trailing matrix is not updated; w = tau (A22 v - V W^T v - W V^T v) comes from
//      one device-wide matrix-vector product over the upper triangle of A22 (the HBM-bound
//      part) with its partials reduced by small follow-up kernels.  After each panel,
//      A22 -= V W^T + W V^T as MFMA GEMMs on the upper tiles.
//   2. Divide and conquer on (d, e): leaves of at most kLeaf rows (workgroup QL eigensolver),
//      merged pairwise bottom-up.  A merge diagonalises D + rho z z^T: one workgroup merges the
//      two sorted child spectra, normalises z and deflates (negligible z_i; close pairs by a
//      Givens rotation), one wavefront per secular root with the origin at the nearer pole, the
//      Gu-Eisenstat recomputed z for orthogonal eigenvectors, the eigenvector update as an MFMA
//      GEMM whose size the deflation kernel writes, and a rank sort of the merged spectrum.
//   3. Back-transformation X = Q Z, blocks of kBT reflectors (four panels) in reverse, each
//      I - V T V^T (forward triangular factor) with MFMA GEMMs; the wide blocks read and write
//      Z a quarter as often as panel-sized ones would.
// Output as the small-block path: eigenvectors as columns of the row-major block, eigenvalues
// ascending.
#include "pomx_internal.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kNB = 32;   // reflectors per panel
constexpr int kLeaf = 64; // largest divide-and-conquer leaf
constexpr int kMaxSymvGrid = 1024;
constexpr int kDotParts = 64; // row chunks of the panel dot products W^T v, V^T v
constexpr int kBT = 128;       // reflectors per block of the back-transformation (4 panels)
constexpr int kSymTile = 64;   // rows per tile of the symmetric matrix-vector product
constexpr int kSymChunk = 256; // columns per work item of the symmetric matrix-vector product
// rows of a wave whose loads are in flight together.  With 256-column items, 8 rows (128
// registers, four waves per SIMD): config-4 batch trd_symv 446 ms against 528 ms with 512-column
// items and 4 rows (same box); 512 columns with 2 / 4 / 8 rows: 1317 / 1274 / 1352 ms per batch
constexpr int kSymRows = 8;

struct BigBlk {
    int n, gsym;              // block size; symv grid width (partial sums per block)
    int64_t mat_off, val_off; // caller's matrix arena / eigenvalue array
    int64_t dbase, ibase, rbase;
    int64_t xbase; // two-stage plans: band and bulge-chasing reflector scratch (doubles)
};

// ---- two-stage reduction parameters ------------------------------------------------------------
constexpr int kBand = kNB;         // semi-bandwidth of the intermediate band matrix
constexpr int kLQT = 256;          // panel columns per workgroup of the panel LQ kernel
constexpr int kGramW = 128;        // columns per workgroup of the 32 x 32 Gram partial sums
static_assert(kLQT >= 7 * 33 && kLQT % 64 == 0, "the LQ kernel finishes 33 partial sums with 7 threads each");
// row ranges of the stage-1 product X^T = V^T A22 (partial sums; the plan takes min(this,
// smallest block / kBand) so that the partials fit U; 16 against 4: 1.887 against 1.917 s at
// n = 15876)
constexpr int kXvSplitMax = 16;
static_assert(kXvSplitMax * kBand * kBand <= kBT * kBT, "T4 fits the S scratch");
constexpr int kQ2G = 32;           // bulge-chasing sweeps per group of the Q2 back-transformation
constexpr int kQ2Waves = 12;       // waves (groups in flight) per workgroup of the Q2 kernel
constexpr int kBtSplitS = 16;      // two-stage back-transformation: row ranges of S = V^T V
constexpr int kBtSplitW = 4;       // and of W1 = V^T X (partial sums in the free D&C workspace U)
constexpr int kBtBatchMin = 2048;  // two-stage plans of blocks this large: S and T of all panels first
constexpr int kTwoStageMin = 2048; // blocks at least this large take the two-stage reduction
// 128 x 128 GEMM tiles (eight waves) from these sizes on: divide-and-conquer merges of blocks
// of at least kBigGemmDc rows, V^T X of the back-transformation in blocks of at least kBigGemmW1
// rows, X -= V W2 when both the reflector rows and the block have at least kBigGemmX
constexpr int kBigGemmDc = 512, kBigGemmW1 = 1024, kBigGemmX = 256;

// Two-stage per-block scratch (doubles from xbase).  AB: the band in lower band storage,
// AB[j * 2 kBand + d] = A[j + d][j] (d < 2 kBand: the band plus room for the bulge).  V2: the
// bulge-chasing reflectors, row (s * kmax + k) = [tau, v_1 .. v_{kBand-1}] of sweep s, step k
// (v_0 = 1).  lq: panel LQ partial sums (two parities).  gp: Gram partial sums.
struct XOff {
    int64_t AB, V2, lq, gp, end;
    int kmax, nchunk, ngram;
};
__host__ __device__ inline XOff xoff(int n) {
    XOff o;
    int64_t at = 0;
    o.kmax = (n - 2) / kBand + 1;
    o.nchunk = (n + kLQT - 1) / kLQT;
    o.ngram = (n + kGramW - 1) / kGramW;
    o.AB = at, at += int64_t(n) * 2 * kBand;
    o.V2 = at, at += int64_t(n > 2 ? n - 2 : 1) * o.kmax * kBand;
    o.lq = at, at += 2 * (int64_t(o.nchunk) * 40 + 40);
    o.gp = at, at += int64_t(o.ngram) * kBand * kBand;
    o.end = at;
    return o;
}
// bulge-chasing steps of sweep s of an n-block (sweeps 0 .. n-3)
__host__ __device__ inline int sweep_steps(int n, int s) { return (n - 2 - s) / kBand + 1; }

// Per-block scratch layout (doubles from dbase).
struct DOff {
    int64_t V, Z, Q2, U, W, Vt, W3, W1, W2, vbuf, wraw, d, e, tau, lam, zs, dsrt, dl, zl, rt, zh, vals, a2, S, T, rowpart, wpart,
        scal, part, rp, cp, end;
};
__host__ __device__ inline DOff doff(int n, int gsym) {
    DOff o;
    int64_t const nn = int64_t(n) * n;
    int64_t at = 0;
    o.V = at, at += nn;
    o.Z = at, at += nn;
    o.Q2 = at, at += nn;
    o.U = at, at += nn;
    o.W = at, at += int64_t(n) * kNB;  // panel w's, reflector-major: W[p][r] at p*n + r
    o.Vt = at, at += int64_t(n) * kNB; // panel reflectors, reflector-major copy of V
    o.W3 = at, at += int64_t(n) * kNB; // copy of W after Vt: [W; Vt] and [Vt; W3] are contiguous
    o.W1 = at, at += int64_t(n) * kBT;
    o.W2 = at, at += int64_t(n) * kBT;
    o.vbuf = at, at += n;
    o.wraw = at, at += n;
    o.d = at, at += n;
    o.e = at, at += n;
    o.tau = at, at += n;
    o.lam = at, at += n;
    o.zs = at, at += n;
    o.dsrt = at, at += n;
    o.dl = at, at += n;
    o.zl = at, at += n;
    o.rt = at, at += n;
    o.zh = at, at += n;
    o.vals = at, at += n;
    o.a2 = at, at += kNB;
    o.S = at, at += kBT * kBT;
    o.T = at, at += kBT * kBT;
    o.rowpart = at, at += 256;
    o.wpart = at, at += 256;
    o.scal = at, at += 64;
    o.part = at, at += int64_t(gsym) * 64;
    int64_t const nt = (n + kSymTile - 1) / kSymTile, nc = (n + kSymChunk - 1) / kSymChunk;
    o.rp = at, at += nt * nc * kSymTile; // symmetric product: row partials per (row tile, chunk)
    o.cp = at, at += nt * int64_t(n);    // column partials per row tile
    o.end = at;
    return o;
}
// Stage-1 panel buffers [W; Vt; W3] (n x kNB each, contiguous) of the panel starting at row k:
// even panels at W, odd panels in Z (free until the divide and conquer; n >= 3 kNB), so that
// the next panel's LQ can fill its own set while this panel's trailing update reads the other.
__host__ __device__ inline int64_t wv_set(DOff const& o, int k) { return ((k / kBand) & 1) ? o.Z : o.W; }

// Per-block integer scratch (from ibase).
struct IOff {
    int64_t perm, order, org, fperm, cnt, end;
};
__host__ __device__ inline IOff ioff(int n) {
    IOff o;
    int64_t at = 0;
    o.perm = at, at += n;
    o.order = at, at += n;
    o.org = at, at += n;
    o.fperm = at, at += n;
    o.cnt = at, at += 4;
    o.end = at;
    return o;
}

struct DcRot {
    int p, q;
    double c, s;
};

struct DcNode { // one merge of the divide-and-conquer tree
    int blk, o, n, n1;
};
struct DcState {
    int k, nrot;
    double rho;
};
struct DcLeaf {
    int blk, a, L, pad;
    int64_t off; // leaf matrix in the leaf arena
};

__device__ inline double wave_sum(double v) {
    for(int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
    for(int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// 256-thread workgroup reductions.
__device__ inline double wg_sum(double v, double* red) {
    v = wave_sum(v);
    __syncthreads();
    if((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    double const s = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return s;
}
__device__ inline double wg_max(double v, double* red) {
    v = wave_max(v);
    __syncthreads();
    if((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    double const s = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    __syncthreads();
    return s;
}

// ---- 1. tridiagonalisation -------------------------------------------------------------------
//
// Panel state per block: V = Vall[:, j0:j0+kNB] (reflectors, unit entry stored), W[:, p] = w_p
// without its last term, a2[p] = -tau/2 (w_p . v_p); the true w_p = W[:, p] + a2[p] V[:, p] is
// folded in by every reader, and into W once per panel before the trailing update.  Each
// device-wide sum is written as per-workgroup partials and finished by a one-workgroup kernel per
// block (the kernel boundary orders the partials; no device-scope fences inside kernels).

// Row j brought up to date with the panel's first i reflectors (columns across the grid), and
// partial sums of the squared row tail beyond j+1.
__global__ void __launch_bounds__(256) trd_row_kernel(const BigBlk* __restrict__ blks, int j, int j0, int i,
                                                      double* __restrict__ Aar, double* __restrict__ D) {
    __shared__ double red[4];
    __shared__ double vj[kNB], wj[kNB], av[kNB];
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n, tid = threadIdx.x;
    DOff const o = doff(n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    const double* __restrict__ Vt = base + o.Vt;
    const double* __restrict__ Wt = base + o.W;
    const double* __restrict__ a2 = base + o.a2;
    if(tid < i) {
        double const v = Vt[int64_t(tid) * n + j];
        vj[tid] = v;
        av[tid] = a2[tid];
        wj[tid] = Wt[int64_t(tid) * n + j] + a2[tid] * v;
    }
    __syncthreads();
    double* __restrict__ row = Aar + B.mat_off + int64_t(j) * n;
    double s = 0.0;
    for(int c = j + blockIdx.x * 256 + tid; c < n; c += gridDim.x * 256) {
        double acc = row[c];
        for(int p = 0; p < i; ++p) {
            double const v = Vt[int64_t(p) * n + c];
            acc -= vj[p] * fma(av[p], v, Wt[int64_t(p) * n + c]) + wj[p] * v;
        }
        row[c] = acc;
        if(c >= j + 2)
            s = fma(acc, acc, s);
    }
    s = wave_sum(s);
    if((tid & 63) == 0)
        red[tid >> 6] = s;
    __syncthreads();
    if(tid == 0)
        base[o.rowpart + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// One workgroup per block: the Householder reflector of the updated row tail; v to vbuf and to
// column j of Vall, d[j], e[j], tau[j].
__global__ void __launch_bounds__(256) trd_refl_kernel(const BigBlk* __restrict__ blks, int j, int i, int nparts,
                                                       const double* __restrict__ Aar, double* __restrict__ D) {
    __shared__ double red[4];
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n, tid = threadIdx.x;
    DOff const o = doff(n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    const double* __restrict__ row = Aar + B.mat_off + int64_t(j) * n;
    double sum = 0.0;
    for(int b = tid; b < nparts; b += 256)
        sum += base[o.rowpart + b];
    double const ss = wg_sum(sum, red);
    int const m = n - j - 1;
    double const alpha = row[j + 1];
    double t = 0.0, beta = alpha, scale = 0.0;
    if(ss > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + ss), alpha);
        t = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
    }
    for(int c = tid; c < m; c += 256) {
        double v = (c == 0) ? 1.0 : (t != 0.0 ? row[j + 1 + c] * scale : 0.0);
        base[o.vbuf + c] = v;
        base[o.V + int64_t(j + 1 + c) * n + j] = v;
        base[o.Vt + int64_t(i) * n + j + 1 + c] = v;
    }
    if(tid == 0)
        base[o.Vt + int64_t(i) * n + j] = 0.0; // v has no component j
    if(tid == 0) {
        base[o.d + j] = row[j];
        base[o.e + j] = beta;
        base[o.tau + j] = t;
    }
}

// w_raw = A22 v from the upper triangle of A22 only (A22 symmetric; its lower triangle is never
// updated): work item (row tile t, column chunk k) covers rows [kSymTile t, +kSymTile) and
// columns [kSymTile t + kSymChunk k, +kSymChunk) on or above the diagonal; it adds A[r][c] v[c]
// to the row partial of r and A[r][c] v[r] (c > r) to the column partial of c.  Only the item's
// kSymChunk entries of v are staged in LDS; each wave has the row segments of kSymRows rows in
// flight at a time, and the lanes' partial row sums are reduced once per item through LDS.
__global__ void __launch_bounds__(256) trd_symv_kernel(const BigBlk* __restrict__ blks, int j, const double* __restrict__ Aar,
                                                       double* __restrict__ D) {
    __shared__ double vc[kSymChunk];
    // the waves' column partials; before them (same storage) every lane's row partials, [row][lane]
    __shared__ double cred[kSymTile * 65 > 4 * kSymChunk ? kSymTile * 65 : 4 * kSymChunk];
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n;
    DOff const o = doff(n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    int const m = n - j - 1, jb = j + 1;
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double* __restrict__ v = base + o.vbuf;
    int const nt = (m + kSymTile - 1) / kSymTile, ncmax = (m + kSymChunk - 1) / kSymChunk;
    int const ncb = (n + kSymChunk - 1) / kSymChunk; // row-partial stride of this block
    constexpr int U = kSymChunk / 64;
    for(int q = blockIdx.x; q < nt * ncmax; q += gridDim.x) {
        int const t = q / ncmax, k = q % ncmax;
        int const R0 = t * kSymTile, C0 = R0 + k * kSymChunk;
        if(C0 >= m)
            continue; // uniform over the workgroup
        __syncthreads(); // the previous item's readers of vc and cred are done
        for(int c = threadIdx.x; c < kSymChunk; c += 256)
            vc[c] = (C0 + c < m) ? v[C0 + c] : 0.0;
        __syncthreads();
        double cacc[U];
#pragma unroll
        for(int u = 0; u < U; ++u)
            cacc[u] = 0.0;
        // rows wave + 4 i of the tile, i = 0..15, kSymRows of them per step (all their loads in
        // flight together)
        constexpr int NRW = kSymRows;
        for(int g = 0; g < kSymTile / (4 * NRW); ++g) {
            int const r0 = R0 + wave + 4 * NRW * g;
            if(r0 >= m)
                break;
            double x[NRW][U];
#pragma unroll
            for(int rr = 0; rr < NRW; ++rr) {
                int const r = r0 + 4 * rr;
                const double* __restrict__ a = Aar + B.mat_off + int64_t(jb + r) * n + jb;
#pragma unroll
                for(int u = 0; u < U; ++u) {
                    int const c = C0 + lane + 64 * u;
                    x[rr][u] = (r < m && c < m && c >= r) ? a[c] : 0.0;
                }
            }
#pragma unroll
            for(int rr = 0; rr < NRW; ++rr) {
                int const r = r0 + 4 * rr;
                double const vr = r < m ? v[r] : 0.0;
                double sr = 0.0;
#pragma unroll
                for(int u = 0; u < U; ++u) {
                    int const c = C0 + lane + 64 * u;
                    sr = fma(x[rr][u], vc[lane + 64 * u], sr);
                    // column parts (c > r); the diagonal element counts once, in the row part
                    cacc[u] = fma(c > r ? x[rr][u] : 0.0, vr, cacc[u]);
                }
                // lane partials of the row sums (reduced across lanes once per tile, below)
                cred[(r - R0) * 65 + lane] = sr;
            }
        }
        __syncthreads();
        { // row sums: 4 threads per row, 16 lanes each (rows past the matrix are not written)
            int const row = threadIdx.x >> 2, part = threadIdx.x & 3;
            double s4 = 0.0;
#pragma unroll
            for(int l = 0; l < 16; ++l)
                s4 += cred[row * 65 + 16 * part + l];
            s4 += __shfl_xor(s4, 1, 64);
            s4 += __shfl_xor(s4, 2, 64);
            if(part == 0 && R0 + row < m)
                base[o.rp + (int64_t(t) * ncb + k) * kSymTile + row] = s4;
        }
        __syncthreads();
#pragma unroll
        for(int u = 0; u < U; ++u)
            cred[wave * kSymChunk + lane + 64 * u] = cacc[u];
        __syncthreads();
        for(int c = threadIdx.x; c < kSymChunk; c += 256) {
            int const cc = C0 + c;
            if(cc < m)
                base[o.cp + int64_t(t) * n + cc] =
                    (cred[c] + cred[kSymChunk + c]) + (cred[2 * kSymChunk + c] + cred[3 * kSymChunk + c]);
        }
    }
}

// w_raw[x] = sum of the row partials of row x and the column partials of x from every row tile
// at or above it; plus partial sums of W^T v and V^T v over a chunk of rows.
__global__ void __launch_bounds__(256) trd_symv_reduce_kernel(const BigBlk* __restrict__ blks, int j, int j0, int i,
                                                              double* __restrict__ D) {
    __shared__ double red[4][64];
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n;
    DOff const o = doff(n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    const double* __restrict__ Vt = base + o.Vt;
    const double* __restrict__ Wt = base + o.W;
    const double* __restrict__ v = base + o.vbuf;
    int const m = n - j - 1, jb = j + 1;
    int const ncmax = (m + kSymChunk - 1) / kSymChunk, ncb = (n + kSymChunk - 1) / kSymChunk;
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // one 64-row slab (= row tile g) per step; the four waves split its row partials (chunks k)
    // and column partials (row tiles tt <= g) and their sums are added in a fixed order
    int const nslab = (m + kSymTile - 1) / kSymTile;
    for(int g = blockIdx.x; g < nslab; g += gridDim.x) {
        int const x = g * kSymTile + lane;
        int const nck = (m - g * kSymTile + kSymChunk - 1) / kSymChunk; // chunks of row tile g
        double s = 0.0;
        if(x < m) {
            for(int k = wave; k < nck && k < ncmax; k += 4)
                s += base[o.rp + (int64_t(g) * ncb + k) * kSymTile + lane];
            for(int tt = wave; tt <= g; tt += 4)
                s += base[o.cp + int64_t(tt) * n + x];
        }
        red[wave][lane] = s;
        __syncthreads();
        if(wave == 0 && x < m)
            base[o.wraw + x] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
        __syncthreads();
    }
    if(i == 0 || int(blockIdx.x) >= kDotParts)
        return;
    // partial dots over a chunk of rows (the first kDotParts workgroups): wavefront per
    // (matrix, p), lanes along the rows
    int const nparts = min(int(gridDim.x), kDotParts);
    int const chunk = (m + nparts - 1) / nparts;
    int const r0 = blockIdx.x * chunk, r1 = min(m, r0 + chunk);
    for(int q = wave; q < 64; q += 4) { // q < 32: (W^T v)_q, 32 <= q < 64: (V^T v)_{q-32}
        int const p = q & 31;
        double acc = 0.0;
        if(p < i) {
            const double* x = (q < 32 ? Wt : Vt) + int64_t(p) * n + jb;
            for(int r = r0 + lane; r < r1; r += 64)
                acc = fma(x[r], v[r], acc);
        }
        acc = wave_sum(acc);
        if(lane == 0)
            red[0][q] = acc;
    }
    __syncthreads();
    if(threadIdx.x < 64)
        base[o.part + int64_t(blockIdx.x) * 64 + threadIdx.x] = red[0][threadIdx.x];
}

// One workgroup per block: t1 = (true W)^T v and t2 = V^T v from the partials.
__global__ void __launch_bounds__(256) trd_t_kernel(const BigBlk* __restrict__ blks, int i, int nparts,
                                                    double* __restrict__ D) {
    __shared__ double red[4][64];
    BigBlk const B = blks[blockIdx.y];
    DOff const o = doff(B.n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double t = 0.0;
    for(int b = wave; b < nparts; b += 4)
        t += base[o.part + int64_t(b) * 64 + lane];
    red[wave][lane] = t;
    __syncthreads();
    if(threadIdx.x < 32) {
        int const q = threadIdx.x;
        double const t1 = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
        double const t2 = (red[0][32 + q] + red[1][32 + q]) + (red[2][32 + q] + red[3][32 + q]);
        base[o.scal + q] = (q < i) ? t1 + base[o.a2 + q] * t2 : 0.0;
        base[o.scal + 32 + q] = (q < i) ? t2 : 0.0;
    }
}

// w = tau (w_raw - V t1 - (true W) t2) stored without its last term as column i of W, plus
// partial sums of w . v.
__global__ void __launch_bounds__(256) trd_w_kernel(const BigBlk* __restrict__ blks, int j, int j0, int i,
                                                    double* __restrict__ D) {
    __shared__ double ca[kNB], cb[kNB];
    __shared__ double red[4];
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n, tid = threadIdx.x;
    DOff const o = doff(n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    const double* __restrict__ Vt = base + o.Vt;
    double* __restrict__ Wt = base + o.W;
    int const m = n - j - 1, jb = j + 1;
    double const tj = base[o.tau + j];
    if(tid < i) {
        double const t1 = base[o.scal + tid], t2 = base[o.scal + 32 + tid];
        ca[tid] = t1 + base[o.a2 + tid] * t2; // coefficient of V[r][p]
        cb[tid] = t2;                         // coefficient of W[r][p]
    }
    __syncthreads();
    double dot = 0.0;
    for(int r = blockIdx.x * 256 + tid; r < m; r += gridDim.x * 256) {
        double w = base[o.wraw + r];
        for(int p = 0; p < i; ++p)
            w -= Vt[int64_t(p) * n + jb + r] * ca[p] + Wt[int64_t(p) * n + jb + r] * cb[p];
        w *= tj;
        Wt[int64_t(i) * n + jb + r] = w;
        dot = fma(w, base[o.vbuf + r], dot);
    }
    dot = wave_sum(dot);
    if((tid & 63) == 0)
        red[tid >> 6] = dot;
    __syncthreads();
    if(tid == 0)
        base[o.wpart + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// One workgroup per block: a2[i] = -tau/2 (w . v) from the partials.
__global__ void __launch_bounds__(256) trd_a2_kernel(const BigBlk* __restrict__ blks, int j, int i, int nparts,
                                                     double* __restrict__ D) {
    __shared__ double red[4];
    BigBlk const B = blks[blockIdx.y];
    DOff const o = doff(B.n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    double sum = 0.0;
    for(int b = threadIdx.x; b < nparts; b += 256)
        sum += base[o.wpart + b];
    double const tot = wg_sum(sum, red);
    if(threadIdx.x == 0)
        base[o.a2 + i] = -0.5 * base[o.tau + j] * tot;
}

// End of a panel (blocks listed in ids, panel ending at reflector j): W[p][r] += a2[p] V[p][r]
// for the rows of the trailing update.
__global__ void trd_fold_kernel(const BigBlk* __restrict__ blks, const int* __restrict__ ids, int j, double* __restrict__ D) {
    // W += a2 V over the trailing rows, copied to W3; rows beyond a short last panel are zeroed
    // (in W, W3 and Vt) so the trailing update can always take kNB reflectors
    BigBlk const B = blks[ids[blockIdx.y]];
    int const n = B.n;
    DOff const o = doff(n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    int const j0 = (j / kNB) * kNB, nb = j - j0 + 1, r0 = j + 1;
    int64_t const cnt = int64_t(n - r0) * kNB;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int const p = int(q / (n - r0)), r = r0 + int(q % (n - r0));
        int64_t const at = int64_t(p) * n + r;
        double w = 0.0;
        if(p < nb)
            w = base[o.W + at] + base[o.a2 + p] * base[o.Vt + at];
        else
            base[o.Vt + at] = 0.0;
        base[o.W + at] = w;
        base[o.W3 + at] = w;
    }
}

// d[n-1] of every block after its last trailing update.
__global__ void trd_last_kernel(const BigBlk* __restrict__ blks, int nblk, const double* __restrict__ Aar,
                                double* __restrict__ D) {
    int const b = blockIdx.x * blockDim.x + threadIdx.x;
    if(b >= nblk)
        return;
    BigBlk const B = blks[b];
    int const n = B.n;
    DOff const o = doff(n, B.gsym);
    D[B.dbase + o.d + n - 1] = Aar[B.mat_off + int64_t(n - 1) * n + (n - 1)];
    D[B.dbase + o.e + n - 1] = 0.0;
}

// ---- 2. divide and conquer ---------------------------------------------------------------------

// which: 0 the eigenvector matrix Z, 1 the reflector matrix V.
__global__ void fill_blocks_kernel(const BigBlk* __restrict__ blks, double* __restrict__ D, int which, double v) {
    BigBlk const B = blks[blockIdx.y];
    DOff const o = doff(B.n, B.gsym);
    int64_t const off = which == 0 ? o.Z : o.V;
    int64_t const cnt = int64_t(B.n) * B.n;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x)
        D[B.dbase + off + q] = v;
}

// Dense leaf matrix: the tridiagonal rows [a, a+L) with |e| removed from the diagonal next to
// every split (T = diag(T1', T2') + |rho| v v^T at a split between two leaves).
__global__ void dc_leaf_build_kernel(const BigBlk* __restrict__ blks, const DcLeaf* __restrict__ leaves,
                                     const double* __restrict__ D, double* __restrict__ arena) {
    DcLeaf const Lf = leaves[blockIdx.x];
    BigBlk const B = blks[Lf.blk];
    DOff const o = doff(B.n, B.gsym);
    const double* d = D + B.dbase + o.d;
    const double* e = D + B.dbase + o.e;
    int const L = Lf.L, a = Lf.a;
    for(int q = threadIdx.x; q < L * L; q += blockDim.x) {
        int const r = q / L, c = q % L;
        double v = 0.0;
        if(r == c) {
            v = d[a + r];
            if(r == 0 && a > 0)
                v -= fabs(e[a - 1]);
            if(r == L - 1 && a + L < B.n)
                v -= fabs(e[a + L - 1]);
        } else if(c == r + 1 || r == c + 1)
            v = e[a + min(r, c)];
        arena[Lf.off + q] = v;
    }
}

__global__ void dc_leaf_place_kernel(const BigBlk* __restrict__ blks, const DcLeaf* __restrict__ leaves,
                                     const double* __restrict__ arena, double* __restrict__ D) {
    DcLeaf const Lf = leaves[blockIdx.x];
    BigBlk const B = blks[Lf.blk];
    DOff const o = doff(B.n, B.gsym);
    int const L = Lf.L, a = Lf.a, N = B.n;
    for(int q = threadIdx.x; q < L * L; q += blockDim.x) {
        int const r = q / L, c = q % L;
        D[B.dbase + o.Z + int64_t(a + r) * N + a + c] = arena[Lf.off + q];
    }
}

// One workgroup per merge: merged order of the two sorted child spectra, z (last row of the left
// child's eigenvectors, sign(rho) times the first row of the right child's), normalisation and
// deflation.  Writes perm (sorted position -> column), order ([non-deflated..., deflated...]),
// the rotations, the secular data dl/zl, the state and the merge's GEMM job.
__global__ void __launch_bounds__(256) dc_deflate_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes,
                                                         DcState* __restrict__ st, pomx_gemm_job* __restrict__ jobs,
                                                         double* __restrict__ D, int* __restrict__ I,
                                                         DcRot* __restrict__ R) {
    __shared__ double red[4];
    __shared__ int s_nk, s_nrot, s_c1[256], s_c2[256];
    DcNode const N = nodes[blockIdx.y];
    BigBlk const B = blks[N.blk];
    int const NB = B.n, o = N.o, n = N.n, n1 = N.n1, tid = threadIdx.x;
    DOff const od = doff(NB, B.gsym);
    IOff const oi = ioff(NB);
    double* base = D + B.dbase;
    const double* lam = base + od.lam + o;
    const double* Z = base + od.Z;
    double* zs = base + od.zs + o;
    double* ds = base + od.dsrt + o;
    double* dl = base + od.dl + o;
    double* zl = base + od.zl + o;
    int* perm = I + B.ibase + oi.perm + o;
    int* order = I + B.ibase + oi.order + o;
    // column types by sorted position (1: rows of the first half only, 3: second half only, 2:
    // both, after a rotation that mixed the halves); the org array is free until the secular
    // equations.  sig: type-sorted position of each kept column (fperm is free until the ranking).
    int* typ = I + B.ibase + oi.org + o;
    int* sig = I + B.ibase + oi.fperm + o;
    DcRot* rot = R + B.rbase + o;
    double const rho_signed = base[od.e + o + n1 - 1];
    double const sgn = rho_signed >= 0.0 ? 1.0 : -1.0;
    double const theta = fabs(rho_signed);
    for(int c = tid; c < n; c += 256) {
        double const x = lam[c];
        int pos, lo, hi;
        double z;
        if(c < n1) { // left element: right elements strictly smaller come first
            lo = n1;
            hi = n;
            while(lo < hi) {
                int mid = (lo + hi) >> 1;
                if(lam[mid] < x)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            pos = c + (lo - n1);
            z = Z[int64_t(o + n1 - 1) * NB + o + c];
        } else { // right element: left elements smaller or equal come first
            lo = 0;
            hi = n1;
            while(lo < hi) {
                int mid = (lo + hi) >> 1;
                if(lam[mid] <= x)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            pos = (c - n1) + lo;
            z = sgn * Z[int64_t(o + n1) * NB + o + c];
        }
        perm[pos] = c;
        typ[pos] = c < n1 ? 1 : 3;
        ds[pos] = x;
        zs[pos] = z;
    }
    __syncthreads();
    double s2 = 0.0;
    for(int c = tid; c < n; c += 256)
        s2 = fma(zs[c], zs[c], s2);
    double const nz2 = wg_sum(s2, red);
    double const nz = sqrt(nz2);
    double const rho = theta * nz2;
    double dm = 0.0, zm = 0.0;
    for(int c = tid; c < n; c += 256) {
        double const zz = nz > 0.0 ? zs[c] / nz : 0.0;
        zs[c] = zz;
        dm = fmax(dm, fabs(ds[c]));
        zm = fmax(zm, fabs(zz));
    }
    double const dmax = wg_max(dm, red), zmax = wg_max(zm, red);
    __syncthreads();
    if(tid == 0) { // the deflation scan is sequential (a close pair rotates into the next candidate)
        double const tol = 8.0 * DBL_EPSILON * fmax(dmax, zmax);
        int nk = 0, ndf = 0, nrot = 0;
        if(!(nz > 0.0) || rho * zmax <= tol) {
            for(int c = 0; c < n; ++c)
                order[n - 1 - ndf++] = c;
        } else {
            int pj = -1;
            for(int c = 0; c < n; ++c) {
                double const zc = zs[c];
                if(rho * fabs(zc) <= tol) {
                    order[n - 1 - ndf++] = c;
                    continue;
                }
                if(pj < 0) {
                    pj = c;
                    continue;
                }
                double const s0 = zs[pj];
                double const t2 = hypot(zc, s0);
                double const t = ds[c] - ds[pj];
                double const cc = zc / t2, ss = -s0 / t2;
                if(fabs(t * cc * ss) <= tol) {
                    zs[c] = t2;
                    zs[pj] = 0.0;
                    rot[nrot++] = DcRot{pj, c, cc, ss};
                    if(typ[c] != typ[pj])
                        typ[c] = 2;
                    double const dp = ds[pj] * cc * cc + ds[c] * ss * ss;
                    double const dq = ds[pj] * ss * ss + ds[c] * cc * cc;
                    ds[pj] = dp;
                    ds[c] = dq;
                    order[n - 1 - ndf++] = pj;
                } else {
                    order[nk] = pj;
                    dl[nk] = ds[pj];
                    zl[nk] = zs[pj];
                    ++nk;
                }
                pj = c;
            }
            if(pj >= 0) {
                order[nk] = pj;
                dl[nk] = ds[pj];
                zl[nk] = zs[pj];
                ++nk;
            }
        }
        s_nk = nk;
        s_nrot = nrot;
    }
    __syncthreads();
    // The kept columns (Z after the gather by order) sorted by type, first half only, mixed,
    // second half only: C = Z U splits into the first half's rows times the first t1 + tm
    // columns and the second half's rows times the last tm + t3 (the other entries of Z are
    // zero), half the product of the dense n x nk x nk.  U's rows follow (dc_u_kernel, sig).
    // sig (a stable counting sort by type) with every thread: counts over contiguous ranges of
    // the kept columns, then each thread's offsets from the counts of the ranges before it.
    int const nk = s_nk;
    int const per = (nk + 255) / 256, i0 = min(nk, tid * per), i1 = min(nk, i0 + per);
    int c1 = 0, c2 = 0;
    for(int i = i0; i < i1; ++i) {
        int const ty = typ[order[i]];
        c1 += ty == 1;
        c2 += ty == 2;
    }
    s_c1[tid] = c1;
    s_c2[tid] = c2;
    __syncthreads();
    int p1 = 0, p2 = 0, t1 = 0, tm = 0;
    for(int q = 0; q < 256; ++q) {
        p1 += q < tid ? s_c1[q] : 0;
        p2 += q < tid ? s_c2[q] : 0;
        t1 += s_c1[q];
        tm += s_c2[q];
    }
    int const t3 = nk - t1 - tm;
    for(int i = i0, a = p1, bm = t1 + p2, c3 = t1 + tm + (i0 - p1 - p2); i < i1; ++i) {
        int const ty = typ[order[i]];
        sig[i] = ty == 1 ? a++ : ty == 2 ? bm++ : c3++;
    }
    __syncthreads(); // every thread has read typ
    for(int i = tid; i < nk; i += 256) // (typ is no longer needed: the order's copy)
        typ[i] = order[i];
    __syncthreads();
    for(int i = tid; i < nk; i += 256)
        order[sig[i]] = typ[i];
    if(tid != 0)
        return;
    st[blockIdx.y] = DcState{nk, s_nrot, rho};
    int64_t const sq = int64_t(o) * NB + o;
    jobs[2 * blockIdx.y] = pomx_gemm_job{n1, nk, t1 + tm, 0, 0, NB, NB, NB, B.dbase + od.Z + sq, B.dbase + od.U + sq,
                                         B.dbase + od.Q2 + sq, 1.0, 0.0};
    jobs[2 * blockIdx.y + 1] = pomx_gemm_job{n - n1, nk, tm + t3, 0, 0, NB, NB, NB, B.dbase + od.Z + sq + int64_t(n1) * NB + t1,
                                             B.dbase + od.U + sq + int64_t(t1) * NB, B.dbase + od.Q2 + sq + int64_t(n1) * NB,
                                             1.0, 0.0};
}

// dst[o+r][o+c] = src[o+r][o+idx[c]] over each merge's square (which: 0 perm Z->Q2, 1 order
// Q2->Z, 2 fperm Q2->Z).
__global__ void dc_gather_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes, int which,
                                 double* __restrict__ D, const int* __restrict__ I) {
    DcNode const N = nodes[blockIdx.y];
    BigBlk const B = blks[N.blk];
    int const NB = B.n, o = N.o, n = N.n;
    DOff const od = doff(NB, B.gsym);
    IOff const oi = ioff(NB);
    const int* idx = I + B.ibase + (which == 0 ? oi.perm : which == 1 ? oi.order : oi.fperm) + o;
    const double* src = D + B.dbase + (which == 0 ? od.Z : od.Q2);
    double* dst = D + B.dbase + (which == 0 ? od.Q2 : od.Z);
    int64_t const cnt = int64_t(n) * n;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int const r = int(q / n), c = int(q % n);
        int const id = idx[c];
        dst[int64_t(o + r) * NB + o + c] = src[int64_t(o + r) * NB + o + ((id >= 0 && id < n) ? id : c)];
    }
}

// The deflation rotations, in order, on every row of the merge's square of Q2.
__global__ void dc_rot_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes,
                              const DcState* __restrict__ st, double* __restrict__ D, const DcRot* __restrict__ R) {
    DcNode const N = nodes[blockIdx.y];
    int const nrot = st[blockIdx.y].nrot;
    if(nrot == 0)
        return;
    BigBlk const B = blks[N.blk];
    int const NB = B.n, o = N.o;
    DOff const od = doff(NB, B.gsym);
    const DcRot* rot = R + B.rbase + o;
    for(int r = blockIdx.x * blockDim.x + threadIdx.x; r < N.n; r += gridDim.x * blockDim.x) {
        double* row = D + B.dbase + od.Q2 + int64_t(o + r) * NB + o;
        for(int t = 0; t < nrot; ++t) {
            DcRot const Rt = rot[t];
            double const a = row[Rt.p], b = row[Rt.q];
            row[Rt.p] = Rt.c * a + Rt.s * b;
            row[Rt.q] = -Rt.s * a + Rt.c * b;
        }
    }
}

// Root j of 1 + rho sum_i z_i^2/(d_i - lambda) = 0 (d strictly increasing, k >= 1), as
// lambda = d[org] + tau with org the nearer pole, by safeguarded Newton inside a bracket.
// One wavefront per root: the lanes split the sums and run the same iteration.
__device__ void secular_root_wave(int k, const double* __restrict__ d, const double* __restrict__ z, double rho, int j,
                                  int* org, double* tau_out) {
    int const lane = threadIdx.x & 63;
    if(k == 1) {
        *org = 0;
        *tau_out = rho * z[0] * z[0];
        return;
    }
    int o;
    double lo, hi;
    if(j < k - 1) {
        double const mid = 0.5 * (d[j + 1] - d[j]);
        double f = 0.0;
        for(int i = lane; i < k; i += 64)
            f += z[i] * z[i] / ((d[i] - d[j]) - mid);
        f = 1.0 + rho * wave_sum(f);
        if(f >= 0.0) {
            o = j;
            lo = 0.0;
            hi = mid;
        } else {
            o = j + 1;
            lo = -mid;
            hi = 0.0;
        }
    } else {
        double zz = 0.0;
        for(int i = lane; i < k; i += 64)
            zz += z[i] * z[i];
        o = k - 1;
        lo = 0.0;
        hi = rho * wave_sum(zz);
    }
    double const dorg = d[o];
    double tau = 0.5 * (lo + hi);
    for(int it = 0; it < 400; ++it) {
        double g = 0.0, gp = 0.0;
        for(int i = lane; i < k; i += 64) {
            double const r = (d[i] - dorg) - tau;
            double const q = z[i] / r;
            g = fma(z[i], q, g); // z_i^2 / r
            gp = fma(q, q, gp);  // z_i^2 / r^2
        }
        g = 1.0 + rho * wave_sum(g);
        gp = rho * wave_sum(gp);
        if(g == 0.0)
            break;
        if(g < 0.0)
            lo = tau;
        else
            hi = tau;
        double tn = tau - g / gp;
        if(!(tn > lo && tn < hi))
            tn = 0.5 * (lo + hi);
        bool const done =
            fabs(tn - tau) <= 2.0 * DBL_EPSILON * fabs(tn) || (hi - lo) <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi));
        tau = tn;
        if(done)
            break;
    }
    *org = o;
    *tau_out = tau;
}

__global__ void __launch_bounds__(256) dc_secular_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes,
                                                         const DcState* __restrict__ st, double* __restrict__ D,
                                                         int* __restrict__ I) {
    DcNode const N = nodes[blockIdx.y];
    DcState const S = st[blockIdx.y];
    BigBlk const B = blks[N.blk];
    DOff const od = doff(B.n, B.gsym);
    IOff const oi = ioff(B.n);
    double* base = D + B.dbase;
    int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for(int jj = blockIdx.x * 4 + wave; jj < S.k; jj += gridDim.x * 4) {
        int og;
        double t;
        secular_root_wave(S.k, base + od.dl + N.o, base + od.zl + N.o, S.rho, jj, &og, &t);
        if(lane == 0) {
            I[B.ibase + oi.org + N.o + jj] = og;
            base[od.rt + N.o + jj] = t;
        }
    }
}

// Gu-Eisenstat: z_i^2 = prod_j (lambda_j - d_i) / (rho prod_{l != i} (d_l - d_i)) as a product of
// positive ratios, with the sign of the original z_i.
__global__ void dc_zhat_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes,
                               const DcState* __restrict__ st, double* __restrict__ D, const int* __restrict__ I) {
    DcNode const N = nodes[blockIdx.y];
    DcState const S = st[blockIdx.y];
    BigBlk const B = blks[N.blk];
    DOff const od = doff(B.n, B.gsym);
    double* base = D + B.dbase;
    const double* d = base + od.dl + N.o;
    const double* tu = base + od.rt + N.o;
    const int* og = I + B.ibase + ioff(B.n).org + N.o;
    int const k = S.k;
    for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        double const di = d[i];
        double p = ((d[og[k - 1]] - di) + tu[k - 1]) / S.rho;
        for(int jj = 0; jj < i; ++jj)
            p *= ((d[og[jj]] - di) + tu[jj]) / (d[jj] - di);
        for(int jj = i; jj < k - 1; ++jj)
            p *= ((d[og[jj]] - di) + tu[jj]) / (d[jj + 1] - di);
        base[od.zh + N.o + i] = copysign(sqrt(fmax(p, 0.0)), base[od.zl + N.o + i]);
    }
}

// 1 / |column jj| of the merge's eigenvector matrix, |column jj|^2 = sum_i zhat_i^2 / ((d_i -
// d_org) - tau_jj)^2, from zhat, d and the roots alone (before U is written), to wraw (free after
// the reduction).  One thread per column; the (d_i, zhat_i) come through LDS 256 at a time.
__global__ void __launch_bounds__(256) dc_unorm_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes,
                                                       const DcState* __restrict__ st, double* __restrict__ D,
                                                       const int* __restrict__ I) {
    __shared__ double dsh[256], zsh[256];
    DcNode const N = nodes[blockIdx.y];
    int const k = st[blockIdx.y].k;
    BigBlk const B = blks[N.blk];
    int const NB = B.n;
    DOff const od = doff(NB, B.gsym);
    double* base = D + B.dbase;
    const double* d = base + od.dl + N.o;
    const int* og = I + B.ibase + ioff(NB).org + N.o;
    for(int j0 = blockIdx.x * 256; j0 < k; j0 += gridDim.x * 256) { // uniform over the workgroup
        int const jj = j0 + threadIdx.x;
        bool const live = jj < k;
        double const dj = live ? d[og[jj]] : 0.0, rj = live ? base[od.rt + N.o + jj] : 0.0;
        double s = 0.0;
        for(int i0 = 0; i0 < k; i0 += 256) {
            __syncthreads(); // the previous chunk's readers are done
            int const i = i0 + threadIdx.x;
            dsh[threadIdx.x] = i < k ? d[i] : 0.0;
            zsh[threadIdx.x] = i < k ? base[od.zh + N.o + i] : 0.0;
            __syncthreads();
            int const cnt = min(256, k - i0);
            if(live)
                for(int q = 0; q < cnt; ++q) {
                    double const u = zsh[q] / ((dsh[q] - dj) - rj);
                    s = fma(u, u, s);
                }
        }
        if(live)
            base[od.wraw + N.o + jj] = 1.0 / sqrt(s);
    }
}

// U[i][j] = zhat_i / (d_i - lambda_j) / |column j| in the merge's square of U.
__global__ void dc_u_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes, const DcState* __restrict__ st,
                            double* __restrict__ D, const int* __restrict__ I) {
    DcNode const N = nodes[blockIdx.y];
    int const k = st[blockIdx.y].k;
    BigBlk const B = blks[N.blk];
    int const NB = B.n;
    DOff const od = doff(NB, B.gsym);
    double* base = D + B.dbase;
    const double* d = base + od.dl + N.o;
    const int* og = I + B.ibase + ioff(NB).org + N.o;
    const int* sg = I + B.ibase + ioff(NB).fperm + N.o; // row of U for kept column i (type-sorted)
    int64_t const cnt = int64_t(k) * k;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int const i = int(q / k), jj = int(q % k);
        base[od.U + int64_t(N.o + sg[i]) * NB + N.o + jj] =
            base[od.zh + N.o + i] / ((d[i] - d[og[jj]]) - base[od.rt + N.o + jj]) * base[od.wraw + N.o + jj];
    }
}

// Deflated columns (k..n-1) pass through: Q2[o+r][o+c] = Z[o+r][o+c].
__global__ void dc_copycols_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes,
                                   const DcState* __restrict__ st, double* __restrict__ D) {
    DcNode const N = nodes[blockIdx.y];
    int const k = st[blockIdx.y].k, w = N.n - k;
    if(w <= 0)
        return;
    BigBlk const B = blks[N.blk];
    int const NB = B.n, o = N.o;
    DOff const od = doff(NB, B.gsym);
    int64_t const cnt = int64_t(N.n) * w;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int const r = int(q / w), c = k + int(q % w);
        int64_t const at = int64_t(o + r) * NB + o + c;
        D[B.dbase + od.Q2 + at] = D[B.dbase + od.Z + at];
    }
}

// Values of the merged spectrum in column order: roots first, then the deflated values.
__global__ void dc_vals_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes,
                               const DcState* __restrict__ st, double* __restrict__ D, const int* __restrict__ I) {
    DcNode const N = nodes[blockIdx.y];
    int const k = st[blockIdx.y].k;
    BigBlk const B = blks[N.blk];
    DOff const od = doff(B.n, B.gsym);
    IOff const oi = ioff(B.n);
    double* base = D + B.dbase;
    for(int c = blockIdx.x * blockDim.x + threadIdx.x; c < N.n; c += gridDim.x * blockDim.x) {
        double v;
        if(c < k)
            v = base[od.dl + N.o + I[B.ibase + oi.org + N.o + c]] + base[od.rt + N.o + c];
        else
            v = base[od.dsrt + N.o + I[B.ibase + oi.order + N.o + c]];
        base[od.vals + N.o + c] = v;
    }
}

// Ascending order of the merged spectrum (rank by counting, ties by column): fperm and lam.
__global__ void dc_rank_kernel(const BigBlk* __restrict__ blks, const DcNode* __restrict__ nodes, double* __restrict__ D,
                               int* __restrict__ I) {
    DcNode const N = nodes[blockIdx.y];
    BigBlk const B = blks[N.blk];
    DOff const od = doff(B.n, B.gsym);
    IOff const oi = ioff(B.n);
    double* base = D + B.dbase;
    const double* vals = base + od.vals + N.o;
    for(int c = blockIdx.x * blockDim.x + threadIdx.x; c < N.n; c += gridDim.x * blockDim.x) {
        double const v = vals[c];
        int r = 0;
        for(int q = 0; q < N.n; ++q) {
            double const u = vals[q];
            r += (u < v) || (u == v && q < c);
        }
        I[B.ibase + oi.fperm + N.o + r] = c;
        base[od.lam + N.o + r] = v;
    }
}

// ---- 3. back-transformation ----------------------------------------------------------------------

__global__ void copy_z_kernel(const BigBlk* __restrict__ blks, const double* __restrict__ D, double* __restrict__ Aar) {
    BigBlk const B = blks[blockIdx.y];
    DOff const od = doff(B.n, B.gsym);
    int64_t const cnt = int64_t(B.n) * B.n;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x)
        Aar[B.mat_off + q] = D[B.dbase + od.Z + q];
}

// Partial products of the split back-transformation summed in a fixed order, one thread per entry,
// blocks listed in ids: which 0, S = sum of the ks products V_z^T V_z at U + z kBT^2 (row ranges z
// of V); which 1, W1 = sum of the ks products V_z^T X_z at U + ks_S kBT^2 + z nb n.
// batched (plans with bt_batch): the slots of panel j0 (S partials at U + (j0/kBT) ks kBT^2, S to
// Z + (j0/kBT) kBT^2; W1 partials at U).
__global__ void bt_sum_kernel(const BigBlk* __restrict__ blks, const int* __restrict__ ids, int j0, int shift, int which,
                              int ks, int batched, double* __restrict__ D) {
    BigBlk const B = blks[ids[blockIdx.y]];
    DOff const od = doff(B.n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    int const n = B.n, nb = min(kBT, n - shift - j0);
    int64_t const slot = j0 / kBT;
    int64_t const cnt = which == 0 ? int64_t(nb) * nb : int64_t(nb) * n;
    int64_t const src = which == 0 ? od.U + (batched ? slot * ks * kBT * kBT : 0)
                                   : od.U + (batched ? 0 : int64_t(kBtSplitS) * kBT * kBT);
    int64_t const stride = which == 0 ? int64_t(kBT) * kBT : int64_t(nb) * n;
    double* __restrict__ dst = base + (which == 0 ? (batched ? od.Z + slot * kBT * kBT : od.S) : od.W1);
    for(int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < cnt; e += int64_t(gridDim.x) * blockDim.x) {
        double a = 0.0;
        for(int z = 0; z < ks; ++z)
            a += base[src + z * stride + e];
        dst[e] = a;
    }
}

// T (nb x nb, upper, nb <= kBT) of I - V T V^T = H_j0 ... H_{j0+nb-1} from S = V^T V and tau, one
// workgroup of kBT threads per block: column i of T is -tau_i T[:i, :i] S[:i, i], T_ii = tau_i.
// Thread r owns row r of T and builds it in place in the block's scratch (it only reads back
// its own writes); the column factor is shared through LDS.
// BATCH: one workgroup per (block, j0) of pairs, S and T in the panel's slots of Z (S at
// Z + (j0/kBT) kBT^2, T at Z + (np + j0/kBT) kBT^2, np = panels of the block).
template <bool BATCH>
__global__ void __launch_bounds__(kBT) wy_t_kernel(const BigBlk* __restrict__ blks, const int* __restrict__ ids, int j0_,
                                                   const int2* __restrict__ pairs, int shift, double* __restrict__ D) {
    __shared__ double col[kBT];
    int const j0 = BATCH ? pairs[blockIdx.x].y : j0_;
    BigBlk const B = blks[BATCH ? pairs[blockIdx.x].x : ids[blockIdx.x]];
    DOff const od = doff(B.n, B.gsym);
    double* base = D + B.dbase;
    int const nb = min(kBT, B.n - shift - j0); // reflector j acts on rows j+shift..
    int const r = threadIdx.x;
    int64_t const slot = j0 / kBT, np = (B.n - shift + kBT - 1) / kBT;
    int64_t const oS = BATCH ? od.Z + slot * kBT * kBT : od.S, oT = BATCH ? od.Z + (np + slot) * kBT * kBT : od.T;
    double* __restrict__ Tr = base + oT + int64_t(r) * nb; // row r (valid for r < nb)
    if(r < nb)
        for(int c = 0; c < nb; ++c)
            Tr[c] = 0.0;
    for(int i = 0; i < nb; ++i) {
        double const ti = base[od.tau + j0 + i];
        __syncthreads(); // readers of the previous column factor are done
        if(r < i)
            col[r] = -ti * base[oS + r * nb + i];
        __syncthreads();
        if(r < i) {
            // four independent chains (loads in flight), added in a fixed order
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
            int q = r;
            for(; q + 3 < i; q += 4) {
                a0 = fma(Tr[q], col[q], a0);
                a1 = fma(Tr[q + 1], col[q + 1], a1);
                a2 = fma(Tr[q + 2], col[q + 2], a2);
                a3 = fma(Tr[q + 3], col[q + 3], a3);
            }
            for(; q < i; ++q)
                a0 = fma(Tr[q], col[q], a0);
            Tr[i] = (a0 + a1) + (a2 + a3);
        }
        if(r == i)
            Tr[i] = ti;
    }
}

// ---- two-stage reduction: dense -> band (kBand) -> tridiagonal ---------------------------------
//
// Stage 1 brings A to a band of semi-bandwidth kBand panel by panel.  The rows k..k+kBand-1 of the
// upper triangle right of the band, P = A[k:k+kBand, k+kBand:n], are factored P = [L 0] Q^T by
// Householder reflectors (an LQ factorisation, one launch per reflector), Q = I - V T V^T, and the
// trailing matrix becomes Q^T A22 Q = A22 - V Z^T - Z V^T with W = A22 V T and
// Z = W - V (T^T V^T W) / 2: products of kBand-wide panels on the matrix cores instead of one
// matrix-vector product per reflector.  Only the upper triangle of A22 is read and updated.
// Stage 2 chases the band down to tridiagonal form: sweep s annihilates column s below the
// subdiagonal with a reflector of length kBand and chases the bulge this creates down the band,
// one step per kBand rows.  A persistent kernel runs many sweeps at once.
// Eigenvectors: X = Q1 Q2 Z; Q2 (the bulge-chasing reflectors) by groups of sweeps on the vector
// units with the eigenvector columns across the lanes, Q1 by the blocked WY back-transformation
// of the one-stage path with the reflectors shifted by kBand.

// Stage 1, panel k (P[r][c] = A[k + r][k + kBand + c], c < m = n - k - kBand), launch i
// (0 <= i <= nref, nref = min(kBand, m - 1): the reflectors with at least two entries).  Launch i
// finishes reflector q = i-1 from the partial sums of launch i-1 (sigma = squared tail of row q,
// d_r = tail of row r . tail of row q, and the column-q values P[r][q] from the workgroup holding
// column q), applies it to the rows q..kBand-1 of this workgroup's columns (one column per
// thread), stores v_q (reflector-major in Vt, column q of V) and tau; then, from the updated rows,
// it writes the partial sums of reflector i.  The panel ends as [L 0] (L lower triangular).
__global__ void __launch_bounds__(kLQT) lq_kernel(const BigBlk* __restrict__ blks, int k, int i, double* __restrict__ Aar,
                                                  double* __restrict__ D) {
    __shared__ double red[4][33];
    __shared__ double fin[33], fin7[7][33];
    __shared__ double colv[kBand];
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n, m = n - k - kBand;
    if(m < 2)
        return;
    int const nref = min(kBand, m - 1);
    int const nch = (m + kLQT - 1) / kLQT;
    if(i > nref || int(blockIdx.x) >= nch)
        return; // uniform over the workgroup
    XOff const x = xoff(n);
    DOff const o = doff(n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    double* __restrict__ lq = D + B.xbase + x.lq;
    int64_t const par = int64_t(x.nchunk) * 40 + 40; // one parity's partial sums
    int const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int const c = blockIdx.x * kLQT + tid;
    bool const live = c < m;
    int const gc = k + kBand + c; // column in the block
    double* __restrict__ P = Aar + B.mat_off + int64_t(k) * n + (k + kBand);
    double p[kBand]; // all rows of the column (finished rows are stored back unchanged)
#pragma unroll
    for(int r = 0; r < kBand; ++r)
        p[r] = 0.0;
    if(live) {
#pragma unroll
        for(int r = 0; r < kBand; ++r)
            p[r] = P[int64_t(r) * n + c];
    }
    if(i == 0 && live)
        for(int q = nref; q < kBand; ++q) // rows of Vt beyond the panel's reflectors: zero
            base[wv_set(o, k) + int64_t(n) * kNB + int64_t(q) * n + gc] = 0.0;
    if(i == 0 && blockIdx.x == 0 && tid >= nref && tid < kBand)
        base[o.tau + k + tid] = 0.0;
    if(i >= 1) {
        int const q = i - 1;
        const double* __restrict__ src = lq + (q & 1) * par;
        // the partial sums of every chunk, finished in a fixed order: 7 threads per value take
        // every 7th chunk, eight chunks' loads in flight at a time, then one thread per value
        // adds their sums
        if(tid < 7 * 33) {
            int const e = tid % 33, g = tid / 33;
            double s0 = 0.0;
            for(int b0 = g; b0 < nch; b0 += 7 * 8) {
                double t[8];
#pragma unroll
                for(int u = 0; u < 8; ++u) {
                    int const b = b0 + 7 * u;
                    t[u] = b < nch ? src[int64_t(b) * 40 + e] : 0.0;
                }
                s0 += ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
            }
            fin7[g][e] = s0;
        }
        if(tid < kBand)
            colv[tid] = src[int64_t(x.nchunk) * 40 + tid];
        __syncthreads();
        if(tid < 33)
            fin[tid] = ((fin7[0][tid] + fin7[1][tid]) + (fin7[2][tid] + fin7[3][tid])) + ((fin7[4][tid] + fin7[5][tid]) + fin7[6][tid]);
        __syncthreads();
        double const alpha = colv[q], sigma = fin[32];
        double tau = 0.0, beta = alpha, scale = 0.0;
        if(sigma > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
            tau = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        // (rows are selected with compile-time indices: p stays in registers)
        double pq = 0.0;
#pragma unroll
        for(int r = 0; r < kBand; ++r)
            pq = (r == q) ? p[r] : pq;
        double v = 0.0; // v_q[c]: 1 at c = q, the scaled tail beyond, 0 before
        if(live)
            v = (c == q) ? 1.0 : (c > q ? pq * scale : 0.0);
        double const newq = (c == q) ? beta : (c > q ? 0.0 : pq);
#pragma unroll
        for(int r = 0; r < kBand; ++r) {
            if(r > q)
                p[r] = fma(-tau * fma(scale, fin[r], colv[r]), v, p[r]);
            else if(r == q)
                p[r] = newq;
        }
        if(live) {
            base[wv_set(o, k) + int64_t(n) * kNB + int64_t(q) * n + gc] = v;
            if(c >= q)
                base[o.V + int64_t(gc) * n + (k + q)] = v;
        }
        if(blockIdx.x == 0 && tid == 0)
            base[o.tau + k + q] = tau;
    }
    if(i < nref) {
        double* __restrict__ dst = lq + (i & 1) * par;
        bool const tail = live && c > i;
        double pi = 0.0;
#pragma unroll
        for(int r = 0; r < kBand; ++r)
            pi = (r == i && tail) ? p[r] : pi;
#pragma unroll
        for(int r = 0; r < kBand; ++r) {
            double const s = wave_sum(r > i ? p[r] * pi : 0.0);
            if(lane == 0)
                red[wave][r] = s;
        }
        double const sg = wave_sum(pi * pi);
        if(lane == 0)
            red[wave][32] = sg;
        __syncthreads();
        if(tid < 33)
            dst[int64_t(blockIdx.x) * 40 + tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
        if(live && c == i) {
#pragma unroll
            for(int r = 0; r < kBand; ++r)
                if(r >= i)
                    dst[int64_t(x.nchunk) * 40 + r] = p[r];
        }
    }
    if(i >= 1 && live) {
#pragma unroll
        for(int r = 0; r < kBand; ++r)
            P[int64_t(r) * n + c] = p[r];
    }
}

// Partial sums over chunks of kGramW columns of C = X Y^T (kBand x kBand): X = Vt, Y = Vt
// (which 0) or W (which 1), rows of length m from column k + kBand.
__global__ void __launch_bounds__(256) gram_kernel(const BigBlk* __restrict__ blks, int k, int which, double* __restrict__ D) {
    __shared__ double xs[kBand][kGramW + 1], ys[kBand][kGramW + 1];
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n, m = n - k - kBand;
    int const c0 = blockIdx.x * kGramW;
    if(m < 2 || c0 >= m)
        return;
    XOff const x = xoff(n);
    DOff const o = doff(n, B.gsym);
    int64_t const wv = wv_set(o, k), vt = wv + int64_t(n) * kNB;
    const double* __restrict__ X = D + B.dbase + vt + (k + kBand);
    const double* __restrict__ Y = D + B.dbase + (which ? wv : vt) + (k + kBand);
    for(int e = threadIdx.x; e < kBand * kGramW; e += 256) {
        int const r = e / kGramW, cc = e % kGramW, c = c0 + cc;
        xs[r][cc] = c < m ? X[int64_t(r) * n + c] : 0.0;
        ys[r][cc] = c < m ? Y[int64_t(r) * n + c] : 0.0;
    }
    __syncthreads();
    double* __restrict__ out = D + B.xbase + x.gp + int64_t(blockIdx.x) * kBand * kBand;
    for(int e = threadIdx.x; e < kBand * kBand; e += 256) {
        int const p = e / kBand, q = e % kBand;
        double s0 = 0.0, s1 = 0.0;
        for(int cc = 0; cc < kGramW; cc += 2) {
            s0 = fma(xs[p][cc], ys[q][cc], s0);
            s1 = fma(xs[p][cc + 1], ys[q][cc + 1], s1);
        }
        out[e] = s0 + s1;
    }
}

// Stage 1: X^T = Vt A22 (kBand x m) from the upper triangle of A22 only, A(r, c) =
// A22[min(r, c)][max(r, c)].  Workgroup (column tile c0 of 64, block, split z) sums the rows r of
// one of xs contiguous ranges, KC at a time: rows above the tile read A22[r][c0..c0+63]
// (row segments), rows below it A22[c0+j][r..r+KC) (the transposed element, also contiguous), and
// the chunks crossing the diagonal element by element.  The product runs on the matrix cores
// (v_mfma_f64_16x16x4f64, each wave 16 x 32 of the 32 x 64 tile), the next chunk prefetched
// into registers.  Split z writes its partial sum to rows [32 z, 32 z + 32) of U (free until the
// divide and conquer); the next product sums the splits (its K runs over all xs x 32 rows).
__global__ void __launch_bounds__(256) xv_kernel(const BigBlk* __restrict__ blks, int k, int xs, const double* __restrict__ Aar,
                                                 double* __restrict__ D) {
    constexpr int KC = 32;
    __shared__ double Ts[KC][64 + 1];
    __shared__ double Vs[kBand][KC + 1];
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n, m = n - k - kBand, k1 = k + kBand;
    int const c0 = blockIdx.x * 64;
    if(m < 2 || c0 >= m)
        return; // uniform over the workgroup
    DOff const o = doff(n, B.gsym);
    const double* __restrict__ A = Aar + B.mat_off + int64_t(k1) * n + k1;
    const double* __restrict__ Vt = D + B.dbase + wv_set(o, k) + int64_t(n) * kNB + k1;
    double* __restrict__ out = D + B.dbase + o.U + int64_t(blockIdx.z) * kBand * n + k1;
    int const nchunk = (m + KC - 1) / KC;
    int const per = (nchunk + xs - 1) / xs;
    int const ch0 = blockIdx.z * per, ch1 = min(nchunk, ch0 + per);
    int const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int const wm = (wave >> 1) * 16, wn = (wave & 1) * 32;
    dbl4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    double ra[8], rv[4];
    // region of chunk ch: 0 rows above the tile, 1 rows below it, 2 across the diagonal
    auto region = [&](int ch) {
        int const r0 = ch * KC;
        return (r0 + KC <= c0) ? 0 : (r0 >= c0 + 64 ? 1 : 2);
    };
    auto load = [&](int ch) {
        int const r0 = ch * KC, reg = region(ch);
#pragma unroll
        for(int u = 0; u < 8; ++u) {
            if(reg == 1) { // (kk fastest: row c0+j of A22, columns r0..)
                int const kk = tid & 31, j = (tid >> 5) + 8 * u, r = r0 + kk, c = c0 + j;
                ra[u] = (r < m && c < m) ? A[int64_t(c) * n + r] : 0.0;
            } else { // (j fastest)
                int const kk = (tid >> 6) + 4 * u, j = tid & 63, r = r0 + kk, c = c0 + j;
                ra[u] = (r < m && c < m) ? A[int64_t(min(r, c)) * n + max(r, c)] : 0.0;
            }
        }
#pragma unroll
        for(int u = 0; u < 4; ++u) {
            int const kk = tid & 31, p = (tid >> 5) + 8 * u, r = r0 + kk;
            rv[u] = (r < m) ? Vt[int64_t(p) * n + r] : 0.0;
        }
    };
    auto store = [&](int ch) {
        int const reg = region(ch);
#pragma unroll
        for(int u = 0; u < 8; ++u) {
            if(reg == 1)
                Ts[tid & 31][(tid >> 5) + 8 * u] = ra[u];
            else
                Ts[(tid >> 6) + 4 * u][tid & 63] = ra[u];
        }
#pragma unroll
        for(int u = 0; u < 4; ++u)
            Vs[(tid >> 5) + 8 * u][tid & 31] = rv[u];
    };
    if(ch0 < ch1)
        load(ch0);
    for(int ch = ch0; ch < ch1; ++ch) {
        __syncthreads(); // the previous chunk's readers are done
        store(ch);
        __syncthreads();
        if(ch + 1 < ch1)
            load(ch + 1);
#pragma unroll
        for(int kq = 0; kq < KC / 4; ++kq) {
            int const kr = kq * 4 + (lane >> 4);
            double const a = Vs[wm + (lane & 15)][kr];
            double const b0 = Ts[kr][wn + (lane & 15)], b1 = Ts[kr][wn + 16 + (lane & 15)];
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc1, 0, 0, 0);
        }
    }
    // D[i][j] of v_mfma_f64_16x16x4f64 in lane j + 16 (i % 4), register i / 4
#pragma unroll
    for(int r = 0; r < 4; ++r) {
        int const p = wm + (lane >> 4) + 4 * r, c = c0 + wn + (lane & 15);
        if(c < m)
            out[int64_t(p) * n + c] = acc0[r];
        if(c + 16 < m)
            out[int64_t(p) * n + c + 16] = acc1[r];
    }
}

// One workgroup per block: the Gram partial sums finished (fixed order), then which 0: T of
// Q = I - V T V^T = H_0 ... H_{kBand-1} (forward, upper; T[:i, i] = -tau_i T[:i, :i] S[:i, i],
// T_ii = tau_i) to o.T; which 1: N = T^T M (M = V^T W) to o.S.  Both kBand x kBand, leading
// dimension kBand.
__global__ void __launch_bounds__(256) panel_t_kernel(const BigBlk* __restrict__ blks, int k, int which, int xs,
                                                      double* __restrict__ D) {
    __shared__ double S[kBand][kBand + 1], T[kBand][kBand + 1];
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n, m = n - k - kBand;
    if(m < 2)
        return;
    XOff const x = xoff(n);
    DOff const o = doff(n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    int const nchg = (m + kGramW - 1) / kGramW;
    const double* __restrict__ gp = D + B.xbase + x.gp;
    int const tid = threadIdx.x;
    // the Gram partial sums of every column chunk: each thread's four entries of S over eight
    // chunks at a time (32 loads in flight), chunks in a fixed order
    static_assert(kBand * kBand == 4 * 256, "four entries of S per thread");
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for(int b0 = 0; b0 < nchg; b0 += 8) {
        double t[4][8];
#pragma unroll
        for(int u = 0; u < 4; ++u)
#pragma unroll
            for(int w = 0; w < 8; ++w)
                t[u][w] = (b0 + w < nchg) ? gp[int64_t(b0 + w) * kBand * kBand + tid + 256 * u] : 0.0;
#pragma unroll
        for(int u = 0; u < 4; ++u)
            acc[u] += ((t[u][0] + t[u][1]) + (t[u][2] + t[u][3])) + ((t[u][4] + t[u][5]) + (t[u][6] + t[u][7]));
    }
#pragma unroll
    for(int u = 0; u < 4; ++u) {
        int const e = tid + 256 * u;
        S[e / kBand][e % kBand] = acc[u];
        T[e / kBand][e % kBand] = which ? base[o.T + e] : 0.0;
    }
    __syncthreads();
    if(which == 0) {
        if(tid < kBand) { // thread r builds row r of T (reads only its own row and S)
            int const r = tid;
            for(int i = 0; i < kBand; ++i) {
                double const ti = base[o.tau + k + i];
                if(r < i) {
                    double a = 0.0;
                    for(int q = r; q < i; ++q)
                        a = fma(T[r][q], S[q][i], a);
                    T[r][i] = -ti * a;
                } else if(r == i) {
                    T[r][i] = ti;
                }
            }
        }
        __syncthreads();
        for(int e = tid; e < kBand * kBand; e += 256)
            base[o.T + e] = T[e / kBand][e % kBand];
        // T4 = [T^T ... T^T] (kBand x xs kBand) to o.S: W^T = T4 [X_0; X_1; ..] sums the split
        // partial products of X^T
        for(int e = tid; e < kBand * xs * kBand; e += 256) {
            int const p = e / (xs * kBand), q = e % kBand;
            base[o.S + e] = T[q][p];
        }
    } else {
        for(int e = tid; e < kBand * kBand; e += 256) {
            int const a = e / kBand, c = e % kBand;
            double s = 0.0;
            for(int r = 0; r <= a; ++r) // T upper triangular: T[r][a] = 0 for r > a
                s = fma(T[r][a], S[r][c], s);
            base[o.S + e] = s;
        }
    }
}

// W3 = W over the panel's trailing columns (the trailing product reads [W; Vt] and [Vt; W3]).
__global__ void copy_w3_kernel(const BigBlk* __restrict__ blks, int k, double* __restrict__ D) {
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n, m = n - k - kBand;
    if(m < 2)
        return;
    DOff const o = doff(n, B.gsym);
    double* __restrict__ base = D + B.dbase;
    int64_t const cnt = int64_t(kBand) * m;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int64_t const at = (q / m) * n + (k + kBand) + (q % m);
        base[wv_set(o, k) + 2 * int64_t(n) * kNB + at] = base[wv_set(o, k) + at];
    }
}

// The band of the reduced matrix (upper rows, A[j][j+d] for d <= kBand) in lower band storage;
// the bulge room (d > kBand) zeroed.
__global__ void band_kernel(const BigBlk* __restrict__ blks, const double* __restrict__ Aar, double* __restrict__ D) {
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n;
    double* __restrict__ AB = D + B.xbase + xoff(n).AB;
    int64_t const cnt = int64_t(n) * 2 * kBand;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int64_t const j = q / (2 * kBand), d = q % (2 * kBand);
        AB[q] = (d <= kBand && j + d < n) ? Aar[B.mat_off + j * n + j + d] : 0.0;
    }
}

// Stage 2, the bulge chase, one launch per wavefront step t: the launch runs every task (s, k) of
// every block with 3 s + k = t (one workgroup each).  Task (s, k) is step k of sweep s: with W_k =
// rows [s+1+k kBand, s+(k+1) kBand] (clipped to the block), the reflector of the column being
// annihilated (column s at k = 0, else the first column of W_{k-1}, the bulge left by step k-1)
// over W_k is formed and applied from the left to the rest of that bulge block A[W_k, W_{k-1}],
// from both sides to the diagonal block A[W_k, W_k], and from the right to A[W_{k+1}, W_k], which
// creates the next bulge.  It needs step k-1 of its sweep and step k+2 of sweep s-1 (both at
// t-1), and it reads and writes only the band columns [annihilated column, end of W_k), a
// contiguous range of AB staged in LDS; the tasks of one step lie at least three steps apart
// along their sweeps, so their column ranges are disjoint.  The kernel boundaries order the
// steps.
__host__ __device__ inline void sweep_range(int n, int t, int& lo, int& hi) {
    // sweeps s of an n-block with a step k = t - 3 s, 0 <= k < sweep_steps(n, s):
    // k <= (n-2-s)/kBand  <=>  kBand t - (n-2) <= (3 kBand - 1) s
    hi = min(n - 3, t / 3);
    int64_t const num = int64_t(kBand) * t - (n - 2);
    lo = num <= 0 ? 0 : int((num + 3 * kBand - 2) / (3 * kBand - 1));
}

__global__ void __launch_bounds__(256) bulge_step_kernel(const BigBlk* __restrict__ blks, int nblk, int t, double* __restrict__ D) {
    constexpr int b = kBand, LD = 2 * kBand;
    __shared__ double raw[2 * kBand][2 * kBand]; // footprint: band column (from c) x offset d
    __shared__ double vsw[4][kBand], wsw[4][kBand]; // each wave's copies of v and w
    int const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // this workgroup's task: blocks in order, sweeps ascending
    int idx = blockIdx.x, blk = -1, s = 0;
    for(int q = 0; q < nblk; ++q) {
        int lo, hi;
        sweep_range(blks[q].n, t, lo, hi);
        int const cnt = max(0, hi - lo + 1);
        if(idx < cnt) {
            blk = q;
            s = lo + idx;
            break;
        }
        idx -= cnt;
    }
    if(blk < 0)
        return;
    BigBlk const B = blks[blk];
    int const n = B.n, k = t - 3 * s;
    XOff const x = xoff(n);
    double* __restrict__ AB = D + B.xbase + x.AB;
    double* __restrict__ V2 = D + B.xbase + x.V2;
    int const r0 = s + 1 + k * b;       // first row of W_k
    int const L = min(b, n - r0);       // rows of W_k (>= 1)
    int const L2 = min(b, n - r0 - b);  // rows of W_{k+1}
    int const c = (k == 0) ? s : r0 - b; // annihilated column = first footprint column
    int const ncol = r0 + L - c;
    int const nel = ncol * LD;
#pragma unroll
    for(int u = 0; u < (2 * kBand * 2 * kBand) / 256; ++u) { // fixed trip count: loads in flight together
        int const e = tid + 256 * u;
        double const a = AB[int64_t(c) * LD + min(e, nel - 1)];
        if(e < nel)
            raw[e / LD][e % LD] = a;
    }
    __syncthreads();
    // A[i][j] (i >= j) = raw[j - c][i - j].  Phase 1, every wave on its own (the reflector and
    // w = tau D v are formed redundantly in each wave, from its own LDS copies): the reflector,
    // the left application to the bulge block, w and its correction.  Phase 2, after a barrier
    // (every wave has read D and x): the updates of D and of the block below.
    double* const vw = vsw[wave];
    double* const ww = wsw[wave];
    double const xa = (lane < L) ? raw[0][r0 + lane - c] : 0.0;
    double const alpha = __shfl(xa, 0, 64);
    double const sig = wave_sum(lane >= 1 ? xa * xa : 0.0);
    double tau = 0.0, beta = alpha, scale = 0.0;
    if(sig > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + sig), alpha);
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
    }
    if(lane < b) {
        double const v = (lane == 0) ? 1.0 : (lane < L ? xa * scale : 0.0);
        vw[lane] = v;
        if(wave == 0)
            V2[(int64_t(s) * x.kmax + k) * b + lane] = (lane == 0) ? tau : v;
    }
    // the wave's own LDS copies are read by other lanes of the same wave
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int const c0 = r0 - c; // raw column of row/column r0
    if(tau != 0.0) {
        if(k >= 1) { // left: columns c+1 .. c+b-1 of the bulge block, rows W_k (half-wave each)
            int const hw = tid >> 5, a = tid & 31;
            for(int qq = 1 + hw; qq < b; qq += 8) {
                double const e = (a < L) ? raw[qq][b + a - qq] : 0.0;
                double dot = (a < L) ? e * vw[a] : 0.0;
                for(int o = 16; o > 0; o >>= 1)
                    dot += __shfl_xor(dot, o, 32);
                if(a < L)
                    raw[qq][b + a - qq] = fma(-tau * dot, vw[a], e);
            }
        }
        // w = tau D v: lane pair (2a, 2a+1) takes row a, half of the columns each
        int const a = lane >> 1, h = lane & 1;
        double acc = 0.0;
        if(a < L)
            for(int u = 0; u < 16; ++u) {
                int const qv = h * 16 + u;
                if(qv < L) {
                    int const lo = min(a, qv), hi = max(a, qv);
                    acc = fma(raw[c0 + lo][hi - lo], vw[qv], acc);
                }
            }
        acc += __shfl_xor(acc, 1, 64);
        double const wa = (a < L) ? tau * acc : 0.0;
        double const al = -0.5 * tau * wave_sum((h == 0 && a < L) ? wa * vw[a] : 0.0);
        if(h == 0)
            ww[a] = fma(al, vw[a], wa);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads(); // every wave has read x and D
    if(tau != 0.0) {
        for(int e = tid; e < b * b; e += 256) {
            int const a = e / b, qv = e % b;
            if(a < L && qv <= a)
                raw[c0 + qv][a - qv] -= vw[a] * ww[qv] + ww[a] * vw[qv];
        }
        { // right on A[W_{k+1}, W_k] (rows below W_k; disjoint from the entries above)
            int const a2 = tid >> 3, part = tid & 7;
            double acc = 0.0;
            if(a2 < L2)
                for(int u = 0; u < 4; ++u) {
                    int const qv = part * 4 + u;
                    if(qv < L)
                        acc = fma(raw[c0 + qv][b + a2 - qv], vw[qv], acc);
                }
            acc += __shfl_xor(acc, 1, 64);
            acc += __shfl_xor(acc, 2, 64);
            acc += __shfl_xor(acc, 4, 64);
            if(a2 < L2)
                for(int u = 0; u < 4; ++u) {
                    int const qv = part * 4 + u;
                    if(qv < L)
                        raw[c0 + qv][b + a2 - qv] = fma(-tau * acc, vw[qv], raw[c0 + qv][b + a2 - qv]);
                }
        }
        if(wave == 0 && lane < L) // the annihilated column
            raw[0][r0 + lane - c] = (lane == 0) ? beta : 0.0;
    }
    __syncthreads();
#pragma unroll
    for(int u = 0; u < (2 * kBand * 2 * kBand) / 256; ++u) {
        int const e = tid + 256 * u;
        if(e < nel)
            AB[int64_t(c) * LD + e] = raw[e / LD][e % LD];
    }
}

// d and e of the tridiagonal matrix left by the bulge chase.
__global__ void band_de_kernel(const BigBlk* __restrict__ blks, double* __restrict__ D) {
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n;
    DOff const o = doff(n, B.gsym);
    const double* __restrict__ AB = D + B.xbase + xoff(n).AB;
    double* __restrict__ base = D + B.dbase;
    for(int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        base[o.d + j] = AB[int64_t(j) * 2 * kBand];
        base[o.e + j] = (j + 1 < n) ? AB[int64_t(j) * 2 * kBand + 1] : 0.0;
    }
}

// X = Q2 Z in place on the eigenvector matrix (row-major; Q2 acts on its rows).  Q2 = product of
// the bulge-chasing reflectors in generation order, so they are applied last first.  Sweeps go in
// groups of kQ2G, groups last to first; inside a group, steps k ascending, and a step's
// reflectors (rows [S+1+k kBand, S+k kBand+kBand+kQ2G-2] together) later sweep first.  Reflectors
// of later steps of earlier sweeps lie below those of earlier steps of later sweeps, so they
// commute and this order equals the reverse generation order.  One lane per column of Z: a step's
// rows stay in registers and slide down by kBand rows per step, and the reflector entries are
// wave-uniform (scalar loads).  The waves of a workgroup take consecutive groups, each a step
// behind the wave before it (step k of group G-1 needs group G's steps up to k; group G's step
// k+1 and group G-1's step k touch disjoint rows), one barrier per step.
template <int NW>
__global__ void __launch_bounds__(64 * NW) q2_kernel(const BigBlk* __restrict__ blks, const double* __restrict__ D,
                                                     double* __restrict__ Aar) {
    constexpr int b = kBand, G = kQ2G, WIN = kBand + kQ2G - 1;
    BigBlk const B = blks[blockIdx.y];
    int const n = B.n, nsw = n - 2;
    if(int(blockIdx.x) * 64 >= n || nsw <= 0)
        return; // uniform over the workgroup
    int const lane = threadIdx.x & 63;
    int const wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform: scalar branches and loads
    int const col = blockIdx.x * 64 + lane;
    bool const live = col < n;
    XOff const x = xoff(n);
    const double* __restrict__ V2 = D + B.xbase + x.V2;
    double* __restrict__ Z = Aar + B.mat_off;
    int const ngroups = (nsw + G - 1) / G;
    double win[WIN];
    for(int rd = 0; rd * NW < ngroups; ++rd) {
        int tsteps = 0; // the round's steps: max over the waves of (steps of its group + its lag)
        for(int w = 0; w < NW; ++w) {
            int const g = ngroups - 1 - (rd * NW + w);
            if(g >= 0)
                tsteps = max(tsteps, sweep_steps(n, g * G) + w);
        }
        int const gid = ngroups - 1 - (rd * NW + wave);
        int const S = gid * G;
        int const kc = gid >= 0 ? sweep_steps(n, S) : 0;
        for(int t = 0; t < tsteps; ++t) {
            int const k = t - wave;
            if(k >= 0 && k < kc) {
                int const rb = S + 1 + k * b; // first row of the window
                if(k == 0) {
#pragma unroll
                    for(int i = 0; i < WIN; ++i) {
                        int const row = rb + i;
                        win[i] = (live && row < n) ? Z[int64_t(row) * n + col] : 0.0;
                    }
                } else {
#pragma unroll
                    for(int i = 0; i < G - 1; ++i)
                        win[i] = win[i + b];
#pragma unroll
                    for(int i = G - 1; i < WIN; ++i) {
                        int const row = rb + i;
                        win[i] = (live && row < n) ? Z[int64_t(row) * n + col] : 0.0;
                    }
                }
#pragma unroll
                for(int j = G - 1; j >= 0; --j) {
                    int const sj = S + j;
                    if(sj < nsw && k < sweep_steps(n, sj)) { // (tau = 0 rows hold v = 0: no branch on tau)
                        const double* __restrict__ vr = V2 + (int64_t(sj) * x.kmax + k) * b;
                        double const tau = vr[0];
                        double d0 = win[j], d1 = 0.0, d2 = 0.0, d3 = 0.0;
#pragma unroll
                        for(int i = 1; i < b; i += 4) {
                            d1 = fma(vr[i], win[j + i], d1);
                            if(i + 1 < b)
                                d2 = fma(vr[i + 1], win[j + i + 1], d2);
                            if(i + 2 < b)
                                d3 = fma(vr[i + 2], win[j + i + 2], d3);
                            if(i + 3 < b)
                                d0 = fma(vr[i + 3], win[j + i + 3], d0);
                        }
                        double const f = tau * ((d0 + d1) + (d2 + d3));
                        win[j] -= f;
#pragma unroll
                        for(int i = 1; i < b; ++i)
                            win[j + i] = fma(-f, vr[i], win[j + i]);
                    }
                    // one reflector's entries live in scalar registers at a time
                    __builtin_amdgcn_sched_barrier(0);
                }
                int const nst = (k == kc - 1) ? WIN : b; // rows leaving the window (all at the last step)
#pragma unroll
                for(int i = 0; i < WIN; ++i) {
                    int const row = rb + i;
                    if(i < nst && live && row < n)
                        Z[int64_t(row) * n + col] = win[i];
                }
            }
            __syncthreads();
        }
    }
}

__global__ void evals_out_kernel(const BigBlk* __restrict__ blks, const double* __restrict__ D, double* __restrict__ eval) {
    BigBlk const B = blks[blockIdx.y];
    DOff const od = doff(B.n, B.gsym);
    for(int c = blockIdx.x * blockDim.x + threadIdx.x; c < B.n; c += gridDim.x * blockDim.x)
        eval[B.val_off + c] = D[B.dbase + od.lam + c];
}

__global__ void zero_ints_kernel(int* __restrict__ p, int64_t n) {
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x)
        p[q] = 0;
}

unsigned gx_for(int64_t work, int per = 256, int cap = 2048) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((work + per - 1) / per, cap)));
}

// Grid width of a launch over ny blocks/nodes: enough workgroups for the work of one row, but
// about 32k workgroups in all (every batched kernel loops over its row's work).
unsigned gx_rows(int64_t work, int per, int ny) {
    int64_t const cap = std::max<int64_t>(1, (32768 + ny - 1) / std::max(ny, 1));
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((work + per - 1) / per, cap)));
}

// Grid width of the per-column reduction kernels over ny blocks: about 2048 workgroups in all
// (these kernels do O(n) work per block and column, so with many blocks one workgroup per block
// avoids paying dispatch for idle workgroups), at least 1 and at most the row's work needs.
unsigned gx_col(int64_t work, int per, int ny) {
    int64_t const cap = std::max<int64_t>(1, (2048 + ny - 1) / std::max(ny, 1));
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((work + per - 1) / per, cap)));
}

template <class T> T* dupload(pomx_ctx* ctx, std::vector<T> const& h, hipStream_t s, bool& ok) {
    void* p = nullptr;
    if(device_malloc(ctx, &p, sizeof(T) * std::max<size_t>(h.size(), 1)) != hipSuccess) {
        ok = false;
        return nullptr;
    }
    if(!h.empty() && hipMemcpyAsync(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s) != hipSuccess)
        ok = false;
    return static_cast<T*>(p);
}

} // namespace

// Static launch plan of one set of blocks.
struct BigPlan {
    int nblk = 0, nmax = 0, gsym = 1;
    bool tri_only = false; // tridiagonal input (d, e already in scratch): no reduction, no back-transformation
    std::vector<BigBlk> h_blk;
    BigBlk* blk = nullptr;
    double* D = nullptr;
    int* I = nullptr;
    DcRot* R = nullptr;
    int64_t D_len = 0, I_len = 0, R_len = 0;
    struct PanelEnd { // reduction: blocks whose panel ends at reflector j
        int j, nblk, max_tiles;
        int64_t ids_off, jobs_off;
    };
    std::vector<PanelEnd> ends;
    int* ids = nullptr;             // block id lists (panel ends, back-transformation panels)
    pomx_gemm_job* gjobs = nullptr; // static GEMM jobs
    int nleaves = 0;
    DcLeaf* leaves = nullptr;
    EighJob* leaf_jobs = nullptr;
    double* leaf_arena = nullptr;
    double* leaf_ws = nullptr;
    int* status = nullptr; // per leaf
    struct Level {
        int nnodes, nmax;
        int64_t node_off;
    };
    std::vector<Level> levels;
    DcNode* nodes = nullptr;
    DcState* states = nullptr;
    pomx_gemm_job* dc_jobs = nullptr;
    struct BtPanel { // back-transformation panel: jobs S (ks per block), W1 (kw per block), W2, X
        int j0, nblk, nb_max, rows_max, ks, kw;
        int64_t ids_off, jobs_off;
    };
    std::vector<BtPanel> bt;
    // two-stage plans whose blocks all have kBtBatchMin rows or more: S = V^T V and T of every
    // panel first (the products in one launch, T of all panels and blocks in one launch), then the
    // updates of X panel by panel.  Slots: the partial products of S of panel j0 at
    // U + (j0 / kBT) kBtSplitS kBT^2, S at Z + (j0 / kBT) kBT^2, T at Z + (np + j0 / kBT) kBT^2
    // (np = panels of the block), the partial products of W1 at U.
    bool bt_batch = false;
    int64_t bta_off = 0; // S jobs of all panels
    int bta_n = 0;
    int2* pairs = nullptr; // (block, j0) of every panel's T
    int npairs = 0;
    size_t bytes = 0;
    // two-stage reduction (dense -> band -> tridiagonal) instead of the one-stage one
    bool two_stage = false;
    int xsplit = 1; // row ranges of the stage-1 product X^T = V^T A22
    struct Panel2 { // stage-1 panel k: blocks [0, nact) active; GEMM jobs X, W, Z, trailing (nact each)
        int k, nact, nref_max, nchunk, ngram, mmax;
        int64_t jobs_off;
    };
    std::vector<Panel2> p2;
    // look-ahead of stage 1: panel p+1's LQ (and its T) runs on the side stream while panel p's
    // trailing update of the other rows runs on the context stream (events: A after the update
    // of the next panel's rows, B after the next panel's LQ; two of each, by panel parity)
    hipStream_t side = nullptr;
    hipEvent_t evA[2] = {nullptr, nullptr}, evB[2] = {nullptr, nullptr};
    std::vector<int> step_tasks; // bulge chase: tasks of wavefront step t (all blocks)
    BigPlan* next = nullptr; // the two-stage plan of the larger blocks, run after this one
};

size_t big_plan_bytes(BigPlan const* p) { return p ? p->bytes + big_plan_bytes(p->next) : 0; }

void big_plan_destroy(BigPlan* p) {
    if(!p)
        return;
    big_plan_destroy(p->next);
    void* ptrs[] = {p->blk,       p->D,          p->I,       p->R,      p->ids,   p->gjobs,  p->leaves,
                    p->leaf_jobs, p->leaf_arena, p->leaf_ws, p->status, p->nodes, p->states, p->dc_jobs, p->pairs};
    for(void* q : ptrs)
        if(q)
            (void)hipFree(q);
    for(int x = 0; x < 2; ++x) {
        if(p->evA[x])
            (void)hipEventDestroy(p->evA[x]);
        if(p->evB[x])
            (void)hipEventDestroy(p->evB[x]);
    }
    if(p->side)
        (void)hipStreamDestroy(p->side);
    delete p;
}

namespace {

int plan_build(pomx_ctx* ctx, int nblk, const int* n_in, const int64_t* mat_off, const int64_t* val_off, bool tri_only,
               bool two_stage, BigPlan** out) {
    *out = nullptr;
    auto* P = new BigPlan();
    P->tri_only = tri_only;
    P->two_stage = two_stage && !tri_only;
    hipStream_t s = ctx->stream;
    // blocks sorted by size, largest first (the active blocks of a reduction step are a prefix)
    std::vector<int> ord(static_cast<size_t>(nblk));
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return n_in[a] > n_in[b]; });
    P->nblk = nblk;
    P->nmax = nblk ? n_in[ord[0]] : 0;
    if(nblk) // partial products of X^T: xsplit kBand rows of n in U (n^2 doubles) per block
        P->xsplit = std::max(1, std::min(kXvSplitMax, n_in[ord[nblk - 1]] / kBand));
    P->gsym = std::max(1, std::min(kMaxSymvGrid, (P->nmax - 1 + 15) / 16));
    if(nblk > 65535) {
        delete P;
        return pomx_fail(ctx, POMX_ERR_RANGE, "large-block eigensolver: more than 65535 blocks in one plan");
    }
    int64_t D = 0, I = 0, R = 0;
    for(int x : ord) {
        int const n = n_in[x];
        if(n < 2) {
            delete P;
            return pomx_fail(ctx, POMX_ERR_ARG, "large-block eigensolver: blocks of size >= 2 expected");
        }
        P->h_blk.push_back(BigBlk{n, P->gsym, mat_off[x], val_off[x], D, I, R, -1});
        D += doff(n, P->gsym).end;
        if(P->two_stage) {
            P->h_blk.back().xbase = D;
            D += xoff(n).end;
        }
        I += ioff(n).end;
        R += n;
    }
    P->D_len = D;
    P->I_len = I;
    P->R_len = R;
    std::vector<int> ids;
    std::vector<pomx_gemm_job> gj;
    if(P->two_stage) {
        // stage-1 panels: per active block (m = n - k - kBand >= 2; a prefix, blocks largest first)
        // the jobs W^T = T4 [X_0; X_1; ..] (the xsplit partial products), W -= N^T Vt / 2, and
        // A22 -= [W Vt]^T [Vt W3] on the upper tiles
        for(int k = 0; P->nmax - k - kBand >= 2; k += kBand) {
            BigPlan::Panel2 Q{k, 0, 0, 0, 0, P->nmax - k - kBand, int64_t(gj.size())};
            while(Q.nact < nblk && P->h_blk[size_t(Q.nact)].n - k - kBand >= 2)
                ++Q.nact;
            Q.nref_max = std::min(kBand, Q.mmax - 1);
            Q.nchunk = (Q.mmax + kLQT - 1) / kLQT;
            Q.ngram = (Q.mmax + kGramW - 1) / kGramW;
            // kinds: 1 W^T, 2 Z^T, 3 the trailing update of the next panel's rows (first kBand rows
            // of A22), 4 the rest of the trailing update (rows and columns from kBand on)
            for(int kind = 1; kind < 5; ++kind)
                for(int b = 0; b < Q.nact; ++b) {
                    BigBlk const& B = P->h_blk[size_t(b)];
                    int const n = B.n, m = n - k - kBand, k1 = k + kBand;
                    DOff const o = doff(n, B.gsym);
                    int64_t const a22 = B.mat_off + int64_t(k1) * n + k1;
                    int64_t const wv = B.dbase + wv_set(o, k), vt = wv + int64_t(n) * kNB; // [W; Vt; W3]
                    if(kind == 1)
                        gj.push_back(pomx_gemm_job{kBand, m, P->xsplit * kBand, 0, 0, P->xsplit * kBand, n, n, B.dbase + o.S,
                                                   B.dbase + o.U + k1, wv + k1, 1.0, 0.0});
                    else if(kind == 2)
                        gj.push_back(pomx_gemm_job{kBand, m, kBand, 1, 0, kBand, n, n, B.dbase + o.S, vt + k1, wv + k1, -0.5, 1.0});
                    else if(kind == 3)
                        gj.push_back(pomx_gemm_job{std::min(m, kBand), m, 2 * kBand, 1, 0, n, n, n, wv + k1, vt + k1, a22, -1.0,
                                                   1.0});
                    else {
                        int const mr = std::max(0, m - kBand);
                        gj.push_back(pomx_gemm_job{mr, mr, 2 * kBand, 1, 0, n, n, n, wv + k1 + kBand, vt + k1 + kBand,
                                                   a22 + int64_t(kBand) * n + kBand, -1.0, 1.0});
                    }
                }
            P->p2.push_back(Q);
        }
        // bulge chase: tasks per wavefront step (step t runs the tasks (s, t - 3 s) of every block)
        if(P->nmax >= 3) {
            int const tend = 3 * (P->nmax - 3) + sweep_steps(P->nmax, 0);
            P->step_tasks.assign(size_t(tend) + 1, 0);
            for(int t = 0; t <= tend; ++t)
                for(auto const& B : P->h_blk) {
                    int lo, hi;
                    sweep_range(B.n, t, lo, hi);
                    P->step_tasks[size_t(t)] += std::max(0, hi - lo + 1);
                }
        }
    }
    if(!tri_only && !P->two_stage) {
        int const jmax = P->nmax - 2;
        std::vector<std::vector<int>> at(size_t(std::max(jmax + 1, 0)));
        for(int b = 0; b < nblk; ++b) {
            int const n = P->h_blk[size_t(b)].n;
            for(int j0 = 0; j0 < n - 1; j0 += kNB)
                at[size_t(std::min(j0 + kNB, n - 1) - 1)].push_back(b);
        }
        for(int j = 0; j <= jmax; ++j) {
            auto const& L = at[size_t(j)];
            if(L.empty())
                continue;
            BigPlan::PanelEnd E{j, int(L.size()), 0, int64_t(ids.size()), int64_t(gj.size())};
            int const j0 = (j / kNB) * kNB, nb = j - j0 + 1, j1 = j + 1;
            int mtmax = 0;
            for(int b : L) { // A22 -= [V2 W2] [W2 V2]^T as one product over 2 kNB reflector rows
                BigBlk const& B = P->h_blk[size_t(b)];
                int const n = B.n, mt = n - j1;
                DOff const o = doff(n, B.gsym);
                ids.push_back(b);
                mtmax = std::max(mtmax, mt);
                gj.push_back(pomx_gemm_job{mt, mt, 2 * kNB, 1, 0, n, n, n, B.dbase + o.Vt + j1, B.dbase + o.W + j1,
                                           B.mat_off + int64_t(j1) * n + j1, -1.0, 1.0});
            }
            int64_t const t = (mtmax + 63) / 64;
            E.max_tiles = int(t * t);
            P->ends.push_back(E);
        }
    }
    // divide-and-conquer tree per block: leaves of kLeaf rows merged pairwise bottom-up
    std::vector<DcLeaf> leaves;
    std::vector<EighJob> ljobs;
    int64_t larena = 0, lws = 0;
    std::vector<std::vector<DcNode>> lv;
    for(int b = 0; b < nblk; ++b) {
        BigBlk const& B = P->h_blk[size_t(b)];
        int const n = B.n, nl = (n + kLeaf - 1) / kLeaf;
        DOff const o = doff(n, B.gsym);
        for(int q = 0; q < nl; ++q) {
            int const a = q * kLeaf, L = std::min(kLeaf, n - a);
            leaves.push_back(DcLeaf{b, a, L, 0, larena});
            ljobs.push_back(EighJob{L, 0, larena, B.dbase + o.lam + a, lws});
            larena += int64_t(L) * L;
            lws += eigh_ws_doubles(L);
        }
        for(int h = 1; (1 << (h - 1)) < nl; ++h) {
            if(int(lv.size()) < h)
                lv.resize(size_t(h));
            int const span = 1 << h, half = 1 << (h - 1);
            for(int q = 0; q * span < nl; ++q) {
                int const l0 = q * span, lm = l0 + half, l1 = std::min(nl, l0 + span);
                if(lm >= l1)
                    continue; // nothing to merge at this level
                int const o0 = l0 * kLeaf, om = lm * kLeaf, o1 = std::min(n, l1 * kLeaf);
                lv[size_t(h - 1)].push_back(DcNode{b, o0, o1 - o0, om - o0});
            }
        }
    }
    std::vector<DcNode> nodes;
    for(auto const& L : lv) {
        if(L.empty())
            continue;
        if(L.size() > 65535) {
            delete P;
            return pomx_fail(ctx, POMX_ERR_RANGE, "large-block eigensolver: more than 65535 merges in one tree level");
        }
        int nm = 0;
        for(auto const& x : L)
            nm = std::max(nm, x.n);
        P->levels.push_back(BigPlan::Level{int(L.size()), nm, int64_t(nodes.size())});
        nodes.insert(nodes.end(), L.begin(), L.end());
    }
    P->nleaves = int(leaves.size());
    // back-transformation panels, last to first; reflector j acts on rows j+sh.. (sh = 1, or kBand
    // after the two-stage reduction, whose reflectors are j = 0 .. n-kBand-2)
    int const sh = P->two_stage ? kBand : 1;
    std::vector<pomx_gemm_job> gja; // batched S jobs, appended after the panels' jobs
    std::vector<int2> pairs;
    P->bt_batch = P->two_stage && nblk > 0 && P->h_blk[size_t(nblk - 1)].n >= kBtBatchMin;
    if(!tri_only && P->nmax - sh >= 1) {
        int const last = ((P->nmax - sh - 1) / kBT) * kBT;
        for(int j0 = last; j0 >= 0; j0 -= kBT) {
            std::vector<int> L;
            for(int b = 0; b < nblk; ++b)
                if(j0 < P->h_blk[size_t(b)].n - sh)
                    L.push_back(b);
            if(L.empty())
                continue;
            // two-stage plans (blocks of 2048 rows and more, few of them) split the long-K products
            // S = V^T V and W1 = V^T X over row ranges into partial sums in U (free after the
            // divide and conquer), so that they run on many workgroups
            // (when every block of the panel has room for the partial sums in U: n^2 doubles)
            int nmin = P->nmax;
            for(int b : L)
                nmin = std::min(nmin, P->h_blk[size_t(b)].n);
            bool const split = P->two_stage && int64_t(nmin) * nmin >= int64_t(kBtSplitS) * kBT * kBT +
                                                                          int64_t(kBtSplitW) * kBT * nmin;
            int const ks = split ? kBtSplitS : 1, kw = split ? kBtSplitW : 1;
            if(P->bt_batch && !split) {
                delete P;
                return pomx_fail(ctx, POMX_ERR_ARG, "large-block eigensolver: batched back-transformation without split");
            }
            BigPlan::BtPanel T{j0, int(L.size()), 0, 0, ks, kw, int64_t(ids.size()), int64_t(gj.size())};
            for(int kind = 0; kind < 4; ++kind) // jobs grouped by kind: S, W1, W2, X
                for(int z = 0; z < (kind == 0 ? ks : kind == 1 ? kw : 1); ++z)
                    for(int b : L) {
                        BigBlk const& B = P->h_blk[size_t(b)];
                        int const n = B.n, nb = std::min(kBT, n - sh - j0), rows = n - j0 - sh;
                        DOff const o = doff(n, B.gsym);
                        int64_t const voff = B.dbase + o.V + int64_t(j0 + sh) * n + j0;
                        int64_t const slot = j0 / kBT, np = (n - sh + kBT - 1) / kBT;
                        int64_t const w1 = P->bt_batch ? B.dbase + o.U
                                                       : (kw > 1 ? B.dbase + o.U + int64_t(ks) * kBT * kBT : B.dbase + o.W1);
                        if(kind == 0) {
                            int const rc = (rows + ks - 1) / ks, r0 = std::min(rows, z * rc), r1 = std::min(rows, r0 + rc);
                            int64_t const out = P->bt_batch ? B.dbase + o.U + (slot * ks + z) * kBT * kBT
                                                : ks > 1     ? B.dbase + o.U + int64_t(z) * kBT * kBT
                                                             : B.dbase + o.S;
                            (P->bt_batch ? gja : gj)
                                .push_back(pomx_gemm_job{nb, nb, r1 - r0, 1, 0, n, n, nb, voff + int64_t(r0) * n,
                                                         voff + int64_t(r0) * n, out, 1.0, 0.0});
                            if(z == 0 && P->bt_batch)
                                pairs.push_back(int2{b, j0});
                            if(z == 0) {
                                ids.push_back(b);
                                T.nb_max = std::max(T.nb_max, nb);
                                T.rows_max = std::max(T.rows_max, rows);
                            }
                        } else if(kind == 1) {
                            int const rc = (rows + kw - 1) / kw, r0 = std::min(rows, z * rc), r1 = std::min(rows, r0 + rc);
                            gj.push_back(pomx_gemm_job{nb, n, r1 - r0, 1, 0, n, n, n, voff + int64_t(r0) * n,
                                                       B.mat_off + int64_t(j0 + sh + r0) * n, w1 + int64_t(z) * nb * n,
                                                       1.0, 0.0});
                        } else if(kind == 2) { // (split: W1 summed from its partial products first)
                            int64_t const tt = P->bt_batch ? B.dbase + o.Z + (np + slot) * kBT * kBT : B.dbase + o.T;
                            gj.push_back(pomx_gemm_job{nb, n, nb, 0, 0, nb, n, n, tt, B.dbase + o.W1, B.dbase + o.W2, 1.0, 0.0});
                        } else
                            gj.push_back(pomx_gemm_job{rows, n, nb, 0, 0, n, n, n, voff, B.dbase + o.W2,
                                                       B.mat_off + int64_t(j0 + sh) * n, -1.0, 1.0});
                    }
            P->bt.push_back(T);
        }
    }
    P->bta_off = int64_t(gj.size());
    P->bta_n = int(gja.size());
    gj.insert(gj.end(), gja.begin(), gja.end());
    P->npairs = int(pairs.size());
    bool ok = true;
    P->pairs = pairs.empty() ? nullptr : dupload(ctx, pairs, s, ok);
    P->blk = dupload(ctx, P->h_blk, s, ok);
    P->ids = dupload(ctx, ids, s, ok);
    P->gjobs = dupload(ctx, gj, s, ok);
    P->leaves = dupload(ctx, leaves, s, ok);
    P->leaf_jobs = dupload(ctx, ljobs, s, ok);
    P->nodes = dupload(ctx, nodes, s, ok);
    std::vector<DcState> st0(nodes.size(), DcState{0, 0, 0.0}); // host sources live until the sync below
    std::vector<pomx_gemm_job> jb0(2 * nodes.size()); // two products per merge (the halves' rows)
    P->states = dupload(ctx, st0, s, ok);
    P->dc_jobs = dupload(ctx, jb0, s, ok);
    auto alloc = [&](void** q, size_t bytes) {
        if(ok && device_malloc(ctx, q, std::max<size_t>(bytes, 8)) != hipSuccess)
            ok = false;
        P->bytes += bytes;
    };
    alloc(reinterpret_cast<void**>(&P->D), sizeof(double) * size_t(D));
    alloc(reinterpret_cast<void**>(&P->I), sizeof(int) * size_t(I));
    alloc(reinterpret_cast<void**>(&P->R), sizeof(DcRot) * size_t(R));
    alloc(reinterpret_cast<void**>(&P->leaf_arena), sizeof(double) * size_t(larena));
    alloc(reinterpret_cast<void**>(&P->leaf_ws), sizeof(double) * size_t(std::max<int64_t>(lws, 1)));
    alloc(reinterpret_cast<void**>(&P->status), sizeof(int) * size_t(std::max(P->nleaves, 1)));
    if(ok && P->p2.size() > 1) { // stage-1 look-ahead (a non-blocking stream of default priority)
        ok = hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking) == hipSuccess;
        for(int x = 0; ok && x < 2; ++x)
            ok = hipEventCreateWithFlags(&P->evA[x], hipEventDisableTiming) == hipSuccess &&
                 hipEventCreateWithFlags(&P->evB[x], hipEventDisableTiming) == hipSuccess;
    }
    if(hipStreamSynchronize(s) != hipSuccess || !ok) { // (the uploads from host vectors have finished)
        big_plan_destroy(P);
        return pomx_fail(ctx, POMX_ERR_HIP, "large-block eigensolver: out of device memory for the plan");
    }
    *out = P;
    return POMX_OK;
}

// Stage 1 (dense -> band) and stage 2 (band -> tridiagonal, d and e to the block scratch).
int run_two_stage(pomx_ctx* ctx, BigPlan* P, double* Aar) {
    hipStream_t s = ctx->stream;
    int rc;
    int const nb = P->nblk;
    fill_blocks_kernel<<<dim3(gx_rows(int64_t(P->nmax) * P->nmax, 256, nb), unsigned(nb)), 256, 0, s>>>(P->blk, P->D, 1, 0.0);
    POMX_CHECK_LAUNCH(ctx);
    // the panel LQ with its T (on stream q)
    auto panel_lq = [&](BigPlan::Panel2 const& Q, hipStream_t q) {
        unsigned const na = unsigned(Q.nact);
        for(int i = 0; i <= Q.nref_max; ++i)
            lq_kernel<<<dim3(unsigned(Q.nchunk), na), kLQT, 0, q>>>(P->blk, Q.k, i, Aar, P->D);
        gram_kernel<<<dim3(unsigned(Q.ngram), na), 256, 0, q>>>(P->blk, Q.k, 0, P->D);
        panel_t_kernel<<<dim3(1, na), 256, 0, q>>>(P->blk, Q.k, 0, P->xsplit, P->D);
    };
    // look-ahead on the side stream, unless the context stream is being captured into a graph
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    POMX_HIP(ctx, hipStreamIsCapturing(s, &cap));
    bool const ahead = P->side && cap == hipStreamCaptureStatusNone;
    size_t const np = P->p2.size();
    for(size_t pi = 0; pi < np; ++pi) {
        auto const& Q = P->p2[pi];
        unsigned const na = unsigned(Q.nact);
        if(pi == 0 || !ahead)
            panel_lq(Q, s);
        else
            POMX_HIP(ctx, hipStreamWaitEvent(s, P->evB[pi & 1], 0)); // this panel's LQ, run ahead
        POMX_CHECK_LAUNCH(ctx);
        int64_t const tc = (Q.mmax + 63) / 64;
        pomx_gemm_job const* J = P->gjobs + Q.jobs_off;
        xv_kernel<<<dim3(unsigned(tc), na, unsigned(P->xsplit)), 256, 0, s>>>(P->blk, Q.k, P->xsplit, Aar, P->D);
        POMX_CHECK_LAUNCH(ctx);
        if((rc = launch_gemm_jobs(ctx, 0, 0, J, Q.nact, tc, P->D, P->D, P->D)))
            return rc;
        gram_kernel<<<dim3(unsigned(Q.ngram), na), 256, 0, s>>>(P->blk, Q.k, 1, P->D);
        panel_t_kernel<<<dim3(1, na), 256, 0, s>>>(P->blk, Q.k, 1, P->xsplit, P->D);
        POMX_CHECK_LAUNCH(ctx);
        if((rc = launch_gemm_jobs(ctx, 1, 0, J + Q.nact, Q.nact, tc, P->D, P->D, P->D)))
            return rc;
        copy_w3_kernel<<<dim3(gx_rows(int64_t(kBand) * Q.mmax, 256, Q.nact), na), 256, 0, s>>>(P->blk, Q.k, P->D);
        POMX_CHECK_LAUNCH(ctx);
        // trailing update: the next panel's rows first, then (beside the next panel's LQ) the rest
        if((rc = launch_gemm_jobs(ctx, 1, 0, J + 2 * Q.nact, Q.nact, tc, P->D, P->D, Aar, true)))
            return rc;
        if(ahead && pi + 1 < np) {
            POMX_HIP(ctx, hipEventRecord(P->evA[pi & 1], s));
            POMX_HIP(ctx, hipStreamWaitEvent(P->side, P->evA[pi & 1], 0));
            panel_lq(P->p2[pi + 1], P->side);
            POMX_CHECK_LAUNCH(ctx);
            POMX_HIP(ctx, hipEventRecord(P->evB[(pi + 1) & 1], P->side));
        }
        if((rc = launch_gemm_jobs(ctx, 1, 0, J + 3 * Q.nact, Q.nact, tc * tc, P->D, P->D, Aar, true)))
            return rc;
    }
    band_kernel<<<dim3(gx_rows(int64_t(P->nmax) * 2 * kBand, 256, nb), unsigned(nb)), 256, 0, s>>>(P->blk, Aar, P->D);
    for(size_t t = 0; t < P->step_tasks.size(); ++t)
        if(P->step_tasks[t] > 0)
            bulge_step_kernel<<<unsigned(P->step_tasks[t]), 256, 0, s>>>(P->blk, nb, int(t), P->D);
    band_de_kernel<<<dim3(gx_rows(P->nmax, 256, nb), unsigned(nb)), 256, 0, s>>>(P->blk, P->D);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

int plan_run(pomx_ctx* ctx, BigPlan* P, double* Aar, double* eval) {
    hipStream_t s = ctx->stream;
    int rc;
    int const nb = P->nblk;
    if(nb == 0)
        return POMX_OK;
    zero_ints_kernel<<<gx_for(P->I_len), 256, 0, s>>>(P->I, P->I_len);
    zero_ints_kernel<<<gx_for(std::max(P->nleaves, 1)), 256, 0, s>>>(P->status, std::max(P->nleaves, 1));
    // 1. reduction to tridiagonal form
    if(P->two_stage) {
        if((rc = run_two_stage(ctx, P, Aar)))
            return rc;
    } else if(!P->tri_only) {
        fill_blocks_kernel<<<dim3(gx_rows(int64_t(P->nmax) * P->nmax, 256, nb), unsigned(nb)), 256, 0, s>>>(P->blk, P->D, 1,
                                                                                                        0.0);
        size_t e = 0;
        int nact = nb;
        int const n0 = P->h_blk[0].n;
        for(int j = 0; j <= P->nmax - 2; ++j) {
            while(nact > 0 && P->h_blk[size_t(nact - 1)].n - 2 < j)
                --nact;
            int const j0 = (j / kNB) * kNB, i = j - j0, m = n0 - j - 1;
            unsigned const g1 = std::min(256u, gx_col(n0 - j, 256, nact));
            unsigned const g2 = std::min(unsigned(P->gsym), gx_col(m, kSymTile, nact));
            unsigned const g3 = std::min(256u, gx_col(m, 256, nact));
            dim3 const one(1, unsigned(nact));
            trd_row_kernel<<<dim3(g1, unsigned(nact)), 256, 0, s>>>(P->blk, j, j0, i, Aar, P->D);
            trd_refl_kernel<<<one, 256, 0, s>>>(P->blk, j, i, int(g1), Aar, P->D);
            int64_t const items = int64_t((m + kSymTile - 1) / kSymTile) * ((m + kSymChunk - 1) / kSymChunk);
            unsigned const gs = gx_rows(items, 1, nact);
            trd_symv_kernel<<<dim3(gs, unsigned(nact)), 256, 0, s>>>(P->blk, j, Aar, P->D);
            trd_symv_reduce_kernel<<<dim3(g2, unsigned(nact)), 256, 0, s>>>(P->blk, j, j0, i, P->D);
            if(i > 0)
                trd_t_kernel<<<one, 256, 0, s>>>(P->blk, i, int(std::min(g2, unsigned(kDotParts))), P->D);
            trd_w_kernel<<<dim3(g3, unsigned(nact)), 256, 0, s>>>(P->blk, j, j0, i, P->D);
            trd_a2_kernel<<<one, 256, 0, s>>>(P->blk, j, i, int(g3), P->D);
            if(e < P->ends.size() && P->ends[e].j == j) {
                BigPlan::PanelEnd const& E = P->ends[e++];
                trd_fold_kernel<<<dim3(gx_rows(int64_t(std::max(m, 1)) * kNB, 256, E.nblk), unsigned(E.nblk)), 256, 0, s>>>(
                    P->blk, P->ids + E.ids_off, j, P->D);
                if((rc = launch_gemm_jobs(ctx, 1, 0, P->gjobs + E.jobs_off, E.nblk, E.max_tiles, P->D, P->D, Aar, true)))
                    return rc;
            }
        }
        trd_last_kernel<<<gx_for(nb), 256, 0, s>>>(P->blk, nb, Aar, P->D);
        POMX_CHECK_LAUNCH(ctx);
    }
    // 2. divide and conquer: leaves
    dc_leaf_build_kernel<<<unsigned(P->nleaves), 256, 0, s>>>(P->blk, P->leaves, P->D, P->leaf_arena);
    if((rc = launch_eigh(ctx, P->leaf_jobs, P->nleaves, kLeaf, 64, 0, nullptr, P->leaf_arena, P->D, P->leaf_ws, P->status)))
        return rc;
    fill_blocks_kernel<<<dim3(gx_rows(int64_t(P->nmax) * P->nmax, 256, nb), unsigned(nb)), 256, 0, s>>>(P->blk, P->D, 0, 0.0);
    dc_leaf_place_kernel<<<unsigned(P->nleaves), 256, 0, s>>>(P->blk, P->leaves, P->leaf_arena, P->D);
    POMX_CHECK_LAUNCH(ctx);
    // merges, level by level
    for(auto const& L : P->levels) {
        DcNode const* nd = P->nodes + L.node_off;
        DcState* st = P->states + L.node_off;
        pomx_gemm_job* jb = P->dc_jobs + 2 * L.node_off; // two products per merge
        dim3 const gsq(gx_rows(int64_t(L.nmax) * L.nmax, 256, L.nnodes), unsigned(L.nnodes));
        dim3 const grow(gx_rows(L.nmax, 256, L.nnodes), unsigned(L.nnodes));
        dim3 const gwave(gx_rows(L.nmax, 4, L.nnodes), unsigned(L.nnodes));
        dc_deflate_kernel<<<dim3(1, unsigned(L.nnodes)), 256, 0, s>>>(P->blk, nd, st, jb, P->D, P->I, P->R);
        dc_gather_kernel<<<gsq, 256, 0, s>>>(P->blk, nd, 0, P->D, P->I);
        dc_rot_kernel<<<grow, 256, 0, s>>>(P->blk, nd, st, P->D, P->R);
        dc_gather_kernel<<<gsq, 256, 0, s>>>(P->blk, nd, 1, P->D, P->I);
        dc_secular_kernel<<<gwave, 256, 0, s>>>(P->blk, nd, st, P->D, P->I);
        dc_zhat_kernel<<<grow, 256, 0, s>>>(P->blk, nd, st, P->D, P->I);
        dc_unorm_kernel<<<grow, 256, 0, s>>>(P->blk, nd, st, P->D, P->I);
        dc_u_kernel<<<gsq, 256, 0, s>>>(P->blk, nd, st, P->D, P->I);
        POMX_CHECK_LAUNCH(ctx);
        int64_t const t = (L.nmax + 63) / 64;
        for(int j0 = 0; j0 < 2 * L.nnodes; j0 += 65534) // (at most 65535 jobs per launch)
            if((rc = launch_gemm_jobs(ctx, 0, 0, jb + j0, std::min(65534, 2 * L.nnodes - j0), t * t, P->D, P->D, P->D, false,
                                      L.nmax >= kBigGemmDc)))
                return rc;
        dc_copycols_kernel<<<gsq, 256, 0, s>>>(P->blk, nd, st, P->D);
        dc_vals_kernel<<<grow, 256, 0, s>>>(P->blk, nd, st, P->D, P->I);
        dc_rank_kernel<<<grow, 256, 0, s>>>(P->blk, nd, P->D, P->I);
        dc_gather_kernel<<<gsq, 256, 0, s>>>(P->blk, nd, 2, P->D, P->I);
        POMX_CHECK_LAUNCH(ctx);
    }
    // 3. back-transformation, eigenvalues
    if(!P->tri_only) {
        copy_z_kernel<<<dim3(gx_rows(int64_t(P->nmax) * P->nmax, 256, nb), unsigned(nb)), 256, 0, s>>>(P->blk, P->D, Aar);
        if(P->two_stage) { // Q2 first: X = Q1 (Q2 Z)
            q2_kernel<kQ2Waves><<<dim3(unsigned((P->nmax + 63) / 64), unsigned(nb)), 64 * kQ2Waves, 0, s>>>(P->blk, P->D, Aar);
            POMX_CHECK_LAUNCH(ctx);
        }
        int const sh = P->two_stage ? kBand : 1;
        int64_t const tn = (P->nmax + 63) / 64;
        int const bb = P->bt_batch ? 1 : 0;
        if(P->bt_batch) { // S and T of every panel first: they do not depend on X
            for(int64_t j0 = 0; j0 < P->bta_n; j0 += 65535)
                if((rc = launch_gemm_jobs(ctx, 1, 0, P->gjobs + P->bta_off + j0, int(std::min<int64_t>(65535, P->bta_n - j0)),
                                          4, P->D, P->D, P->D)))
                    return rc;
            for(auto const& T : P->bt)
                bt_sum_kernel<<<dim3(unsigned((T.nb_max * T.nb_max + 255) / 256), unsigned(T.nblk)), 256, 0, s>>>(
                    P->blk, P->ids + T.ids_off, T.j0, sh, 0, T.ks, 1, P->D);
            wy_t_kernel<true><<<unsigned(P->npairs), kBT, 0, s>>>(P->blk, nullptr, 0, P->pairs, sh, P->D);
            POMX_CHECK_LAUNCH(ctx);
        }
        for(auto const& T : P->bt) {
            pomx_gemm_job const* J = P->gjobs + T.jobs_off;
            int const ns = P->bt_batch ? 0 : T.ks * T.nblk, nw = T.kw * T.nblk;
            int64_t const tb = (T.nb_max + 63) / 64;
            int64_t const tr = (T.rows_max + 63) / 64;
            if(!P->bt_batch) {
                if((rc = launch_gemm_jobs(ctx, 1, 0, J, ns, tb * tb, P->D, P->D, P->D)))
                    return rc;
                if(T.ks > 1)
                    bt_sum_kernel<<<dim3(unsigned((T.nb_max * T.nb_max + 255) / 256), unsigned(T.nblk)), 256, 0, s>>>(
                        P->blk, P->ids + T.ids_off, T.j0, sh, 0, T.ks, 0, P->D);
                wy_t_kernel<false><<<unsigned(T.nblk), kBT, 0, s>>>(P->blk, P->ids + T.ids_off, T.j0, nullptr, sh, P->D);
            }
            if((rc = launch_gemm_jobs(ctx, 1, 0, J + ns, nw, tb * tn, P->D, Aar, P->D, false, P->nmax >= kBigGemmW1)))
                return rc;
            if(T.kw > 1)
                bt_sum_kernel<<<dim3(gx_rows(int64_t(T.nb_max) * P->nmax, 256, T.nblk), unsigned(T.nblk)), 256, 0, s>>>(
                    P->blk, P->ids + T.ids_off, T.j0, sh, 1, T.kw, bb, P->D);
            if((rc = launch_gemm_jobs(ctx, 0, 0, J + ns + nw, T.nblk, tb * tn, P->D, P->D, P->D)) ||
               (rc = launch_gemm_jobs(ctx, 0, 0, J + ns + nw + T.nblk, T.nblk, tr * tn, P->D, P->D, Aar, false,
                                      T.rows_max >= kBigGemmX && P->nmax >= kBigGemmX)))
                return rc;
        }
        evals_out_kernel<<<dim3(gx_rows(P->nmax, 256, nb), unsigned(nb)), 256, 0, s>>>(P->blk, P->D, eval);
        POMX_CHECK_LAUNCH(ctx);
    }
    return POMX_OK;
}

} // namespace

// Blocks of at least this size take the two-stage reduction: kTwoStageMin, or a positive integer
// in the environment variable POMX_EIGH_TWO_STAGE_MIN (a test hook, so that the two-stage path
// is tested on small blocks).
int two_stage_min() {
    if(const char* e = std::getenv("POMX_EIGH_TWO_STAGE_MIN")) {
        long const v = std::atol(e);
        if(v > 0) // at least 3 kNB: the odd panels' stage-1 buffers live in Z
            return int(std::min<long>(std::max<long>(v, 3 * kNB), 1L << 30));
    }
    return kTwoStageMin;
}

size_t big_block_bytes(int n) {
    int const gsym = std::max(1, std::min(kMaxSymvGrid, (n - 1 + 15) / 16));
    int64_t d = doff(n, gsym).end + (n >= two_stage_min() ? xoff(n).end : 0);
    return sizeof(double) * size_t(d) + sizeof(int) * size_t(ioff(n).end) + sizeof(DcRot) * size_t(n);
}

int big_plan_create(pomx_ctx* ctx, int nblk, const int* n, const int64_t* mat_off, const int64_t* val_off, BigPlan** plan) {
    *plan = nullptr;
    int const tmin = two_stage_min();
    std::vector<int> n1, n2;
    std::vector<int64_t> m1, m2, v1, v2;
    for(int b = 0; b < nblk; ++b) {
        bool const big = n[b] >= tmin;
        (big ? n2 : n1).push_back(n[b]);
        (big ? m2 : m1).push_back(mat_off[b]);
        (big ? v2 : v1).push_back(val_off[b]);
    }
    BigPlan* P = nullptr;
    int rc = plan_build(ctx, int(n1.size()), n1.data(), m1.data(), v1.data(), false, false, &P);
    if(rc)
        return rc;
    if(!n2.empty() && (rc = plan_build(ctx, int(n2.size()), n2.data(), m2.data(), v2.data(), false, true, &P->next))) {
        big_plan_destroy(P);
        return rc;
    }
    *plan = P;
    return POMX_OK;
}

int big_plan_enqueue(pomx_ctx* ctx, BigPlan* plan, double* d_mat, double* d_eval) {
    for(BigPlan* p = plan; p; p = p->next) {
        int rc = plan_run(ctx, p, d_mat, d_eval);
        if(rc)
            return rc;
    }
    return POMX_OK;
}

int big_plan_status(pomx_ctx* ctx, BigPlan* P) {
    if(!P)
        return POMX_OK;
    if(P->nleaves > 0) {
        std::vector<int> st(size_t(P->nleaves));
        POMX_HIP(ctx, hipMemcpy(st.data(), P->status, sizeof(int) * st.size(), hipMemcpyDeviceToHost));
        for(int v : st)
            if(v)
                return pomx_fail(ctx, POMX_ERR_RANGE, "large-block eigensolver: a leaf eigensolve did not converge");
    }
    return big_plan_status(ctx, P->next);
}

extern "C" int pomx_eigh_tridiag(pomx_ctx* ctx, int n, const double* h_d, const double* h_e, double* d_Z, double* h_w) {
    POMX_REQUIRE(ctx, n >= 1 && h_d && (h_e || n == 1) && d_Z && h_w, "pomx_eigh_tridiag: bad arguments");
    if(n == 1) {
        double one = 1.0;
        h_w[0] = h_d[0];
        POMX_HIP(ctx, hipMemcpy(d_Z, &one, sizeof(double), hipMemcpyHostToDevice));
        return POMX_OK;
    }
    int64_t zero = 0;
    BigPlan* P = nullptr;
    int rc = plan_build(ctx, 1, &n, &zero, &zero, true, false, &P);
    if(rc)
        return rc;
    BigBlk const& B = P->h_blk[0];
    DOff const o = doff(n, B.gsym);
    std::vector<double> e(size_t(n), 0.0);
    for(int q = 0; q + 1 < n; ++q)
        e[size_t(q)] = h_e[q];
    hipStream_t s = ctx->stream;
    if(hipMemcpyAsync(P->D + o.d, h_d, sizeof(double) * size_t(n), hipMemcpyHostToDevice, s) != hipSuccess ||
       h2d_sync(ctx, P->D + o.e, e.data(), sizeof(double) * size_t(n)) != hipSuccess) {
        big_plan_destroy(P);
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_eigh_tridiag: copy failed");
    }
    rc = plan_run(ctx, P, nullptr, nullptr);
    if(!rc && (hipMemcpyAsync(d_Z, P->D + o.Z, sizeof(double) * size_t(n) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
               hipMemcpyAsync(h_w, P->D + o.lam, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, s) != hipSuccess ||
               hipStreamSynchronize(s) != hipSuccess))
        rc = pomx_fail(ctx, POMX_ERR_HIP, "pomx_eigh_tridiag: copy failed");
    if(!rc)
        rc = big_plan_status(ctx, P);
    big_plan_destroy(P);
    return rc;
}
