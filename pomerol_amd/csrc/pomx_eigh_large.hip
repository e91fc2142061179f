// Dense symmetric eigensolver for large blocks (n > 512): the whole device works on one block.
//
// Replaces Eigen::SelfAdjointEigenSolver in HamiltonianPart::computeImpl
// (src/pomerol/HamiltonianPart.cpp:90-102) for the blocks that do not fit one workgroup.
//   1. Blocked Householder tridiagonalisation T = Q^T A Q, panels of kNB reflectors.  Inside a
//      panel the trailing matrix is not updated; each reflector's w = tau (A22 v - V W^T v -
//      W V^T v) comes from one device-wide matrix-vector product over A22 (the HBM-bound part).
//      After the panel, A22 -= V W^T + W V^T as two FP64 MFMA GEMMs.  The reflectors are kept
//      column-wise in a separate matrix for step 3.
//   2. Divide and conquer on (d, e): the tridiagonal matrix is split recursively into leaves of
//      at most kLeaf rows, solved by the per-block workgroup eigensolver.  Each merge of two
//      solved halves diagonalises D + rho z z^T: deflation of negligible z_i and of close
//      eigenvalue pairs (Givens rotation of the two columns) on the host, the secular equation
//      1 + rho sum z_i^2/(d_i - lambda) = 0 one root per thread with the origin at the nearer
//      pole, the Gu-Eisenstat recomputed z for orthogonal eigenvectors, and the eigenvector
//      update as an FP64 MFMA GEMM.  All merges of one tree level run together.
//   3. Back-transformation X = Q Z, panels in reverse order, each as I - V T V^T (forward
//      accumulated triangular factor T) with three MFMA GEMMs.
// Output as the small-block path: eigenvectors as columns of the row-major block, eigenvalues
// ascending.
#include "pomx_internal.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <functional>
#include <numeric>

namespace {

constexpr int kNB = 32;     // reflectors per panel
constexpr int kLeaf = 128;  // largest divide-and-conquer leaf
constexpr int kSymvGrid = 1024;

__device__ inline double wave_sum(double v) {
    for(int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup (blockDim.x a multiple of 64, at most 1024).
__device__ inline double block_sum(double v, double* red) {
    v = wave_sum(v);
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if(lane == 0)
        red[wave] = v;
    __syncthreads();
    double s = 0.0;
    for(int i = 0; i < int(blockDim.x >> 6); ++i)
        s += red[i];
    __syncthreads();
    return s;
}

// ---- 1. tridiagonalisation ---------------------------------------------------------------
//
// Panel state: V = Vall[:, j0:j0+kNB] (reflectors, unit entry stored), Wp[:, p] = w_p without
// its last term, a2[p] = -tau/2 (w_p . v_p); the true w_p is Wp[:, p] + a2[p] V[:, p] (folded
// in by every reader, and into Wp once per panel before the trailing update).  Device-wide
// reductions finish in the last workgroup to arrive (counter ticket, device-scope fences), so
// a reflector costs three launches and no single-workgroup pass over the matrix.

struct TrdScal {
    double scale;   // v = x * scale (x = updated row tail), v[0] = 1
    double t12[64]; // t1 = W~^T v (p < 32), t2 = V^T v (32 + p)
};

// Arrive at the ticket counter; true in the last workgroup (which then sees every other
// workgroup's stores) and the counter is rearmed for the next launch.
__device__ inline bool last_block(unsigned* counter) {
    __shared__ bool last;
    __threadfence();
    __syncthreads();
    if(threadIdx.x == 0) {
        unsigned t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
        if(last)
            atomicExch(counter, 0u);
    }
    __syncthreads();
    if(last)
        __threadfence();
    return last;
}

// Row j brought up to date with the panel's first i reflectors (columns across the grid);
// the last workgroup forms the reflector of the row tail and writes v (vbuf and Vall column j).
__global__ void __launch_bounds__(256) trd_row_kernel(int n, int j, int j0, int i, double* __restrict__ A,
                                                      double* __restrict__ Vall, const double* __restrict__ Wp,
                                                      const double* __restrict__ a2, double* __restrict__ vbuf,
                                                      double* __restrict__ d, double* __restrict__ e,
                                                      double* __restrict__ tau, double* __restrict__ rowpart,
                                                      unsigned* __restrict__ counter) {
    __shared__ double red[4];
    __shared__ double vj[kNB], wj[kNB], av[kNB];
    int const tid = threadIdx.x;
    if(tid < i) {
        double const v = Vall[int64_t(j) * n + j0 + tid];
        vj[tid] = v;
        av[tid] = a2[tid];
        wj[tid] = Wp[int64_t(j) * kNB + tid] + a2[tid] * v;
    }
    __syncthreads();
    double* __restrict__ row = A + int64_t(j) * n;
    double s = 0.0;
    for(int c = j + blockIdx.x * 256 + tid; c < n; c += gridDim.x * 256) {
        double acc = row[c];
        const double* wc = Wp + int64_t(c) * kNB;
        const double* vc = Vall + int64_t(c) * n + j0;
        for(int p = 0; p < i; ++p) {
            double const v = vc[p];
            acc -= vj[p] * fma(av[p], v, wc[p]) + wj[p] * v;
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
        rowpart[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    if(!last_block(counter))
        return;
    double sum = 0.0;
    for(int b = tid; b < int(gridDim.x); b += 256)
        sum += rowpart[b];
    sum = wave_sum(sum);
    __syncthreads();
    if((tid & 63) == 0)
        red[tid >> 6] = sum;
    __syncthreads();
    double const ss = red[0] + red[1] + red[2] + red[3];
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
        vbuf[c] = v;
        Vall[int64_t(j + 1 + c) * n + j] = v;
    }
    if(tid == 0) {
        d[j] = row[j];
        e[j] = beta;
        tau[j] = t;
    }
}

// w_raw = A22 v (A22 = rows/columns j+1.. at the start of the panel), one wavefront per row,
// and partial sums of Wp^T v, V^T v over a chunk of rows; the last workgroup reduces them to
// t1 = W~^T v = Wp^T v + a2 o (V^T v) and t2 = V^T v.
__global__ void __launch_bounds__(256) trd_symv_kernel(int n, int j, int j0, int i, const double* __restrict__ A,
                                                       const double* __restrict__ vbuf, const double* __restrict__ Vall,
                                                       const double* __restrict__ Wp, const double* __restrict__ a2,
                                                       double* __restrict__ wraw, double* __restrict__ part,
                                                       TrdScal* __restrict__ sc, unsigned* __restrict__ counter) {
    __shared__ double red[4][64];
    int const m = n - j - 1, base = j + 1;
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for(int r = blockIdx.x * 4 + wave; r < m; r += gridDim.x * 4) {
        const double* __restrict__ a = A + int64_t(base + r) * n + base;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int c = lane;
        for(; c + 192 < m; c += 256) {
            s0 = fma(a[c], vbuf[c], s0);
            s1 = fma(a[c + 64], vbuf[c + 64], s1);
            s2 = fma(a[c + 128], vbuf[c + 128], s2);
            s3 = fma(a[c + 192], vbuf[c + 192], s3);
        }
        for(; c < m; c += 64)
            s0 = fma(a[c], vbuf[c], s0);
        double s = wave_sum((s0 + s1) + (s2 + s3));
        if(lane == 0)
            wraw[r] = s;
    }
    if(i == 0)
        return; // uniform over the grid: no reduction needed
    int const chunk = (m + gridDim.x - 1) / gridDim.x;
    int const r0 = blockIdx.x * chunk, r1 = min(m, r0 + chunk);
    int const p = lane; // p < 32: (Wp^T v)_p, 32 <= p < 64: (V^T v)_{p-32}
    double acc = 0.0;
    if((p & 31) < i)
        for(int r = r0 + wave; r < r1; r += 4) {
            double x = (p < 32) ? Wp[int64_t(base + r) * kNB + p] : Vall[int64_t(base + r) * n + j0 + (p - 32)];
            acc = fma(x, vbuf[r], acc);
        }
    red[wave][lane] = acc;
    __syncthreads();
    if(wave == 0)
        part[int64_t(blockIdx.x) * 64 + lane] = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
    if(!last_block(counter))
        return;
    // 256 threads: 4 groups of 64 sum strided workgroups, then combine
    double t = 0.0;
    for(int b = wave; b < int(gridDim.x); b += 4)
        t += part[int64_t(b) * 64 + lane];
    __syncthreads();
    red[wave][lane] = t;
    __syncthreads();
    if(wave == 0) {
        double v = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
        red[0][lane] = v;
    }
    __syncthreads();
    if(threadIdx.x < 32) {
        int const q = threadIdx.x;
        double const t2 = red[0][32 + q];
        sc->t12[q] = (q < i) ? red[0][q] + a2[q] * t2 : 0.0;
        sc->t12[32 + q] = (q < i) ? t2 : 0.0;
    }
}

// w = tau (w_raw - V t1 - W~ t2) stored (without its last term) as column i of Wp; the last
// workgroup finishes a2[i] = -tau/2 (w . v).
__global__ void __launch_bounds__(256) trd_w_kernel(int n, int j, int j0, int i, const double* __restrict__ wraw,
                                                    const double* __restrict__ vbuf, const double* __restrict__ Vall,
                                                    double* __restrict__ Wp, double* __restrict__ a2,
                                                    const double* __restrict__ tau, const TrdScal* __restrict__ sc,
                                                    double* __restrict__ wpart, unsigned* __restrict__ counter) {
    __shared__ double ca[kNB], cb[kNB];
    __shared__ double red[4];
    int const tid = threadIdx.x;
    int const m = n - j - 1, base = j + 1;
    double const tj = tau[j];
    if(tid < i) {
        double const t1 = sc->t12[tid], t2 = sc->t12[32 + tid];
        ca[tid] = t1 + a2[tid] * t2; // coefficient of V[r][p] (t1 plus the deferred a2 V t2)
        cb[tid] = t2;                // coefficient of Wp[r][p]
    }
    __syncthreads();
    double dot = 0.0;
    for(int r = blockIdx.x * 256 + tid; r < m; r += gridDim.x * 256) {
        double w = wraw[r];
        const double* vr = Vall + int64_t(base + r) * n + j0;
        const double* wr = Wp + int64_t(base + r) * kNB;
        for(int p = 0; p < i; ++p)
            w -= vr[p] * ca[p] + wr[p] * cb[p];
        w *= tj;
        Wp[int64_t(base + r) * kNB + i] = w;
        dot = fma(w, vbuf[r], dot);
    }
    dot = wave_sum(dot);
    if((tid & 63) == 0)
        red[tid >> 6] = dot;
    __syncthreads();
    if(tid == 0)
        wpart[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    if(!last_block(counter))
        return;
    double sum = 0.0;
    for(int b = tid; b < int(gridDim.x); b += 256)
        sum += wpart[b];
    sum = wave_sum(sum);
    __syncthreads();
    if((tid & 63) == 0)
        red[tid >> 6] = sum;
    __syncthreads();
    if(tid == 0)
        a2[i] = -0.5 * tj * (red[0] + red[1] + red[2] + red[3]);
}

// End of a panel: Wp[r][p] += a2[p] V[r][p] for the rows of the trailing update.
__global__ void trd_fold_kernel(int n, int j0, int nb, int r0, const double* __restrict__ Vall, double* __restrict__ Wp,
                                const double* __restrict__ a2) {
    int64_t const cnt = int64_t(n - r0) * nb;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int const r = r0 + int(q / nb), p = int(q % nb);
        Wp[int64_t(r) * kNB + p] += a2[p] * Vall[int64_t(r) * n + j0 + p];
    }
}

__global__ void copy_diag_kernel(int n, const double* __restrict__ A, double* __restrict__ d) {
    if(threadIdx.x == 0 && blockIdx.x == 0)
        d[n - 1] = A[int64_t(n - 1) * n + (n - 1)];
}

__global__ void fill_kernel(double* __restrict__ p, int64_t len, double v) {
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < len; q += int64_t(gridDim.x) * blockDim.x)
        p[q] = v;
}

// ---- 2. divide and conquer -----------------------------------------------------------------

struct DcMerge {
    int o, n, n1, k;             // node offset/size, left child size, non-deflated count
    int64_t perm_off, rot_off;   // into the level's permutation and rotation lists
    int nrot, pad;
    int64_t sec_off, u_off;      // secular data (k entries), rank-one eigenvector matrix (k x k)
    double rho;                  // rho * |z|^2 (z normalised)
};

struct DcRot {
    int p, q;
    double c, s;
};

// z of every merge: last row of the left child's eigenvectors, first row of the right child's.
__global__ void dc_z_kernel(int nm, const DcMerge* __restrict__ mg, const double* __restrict__ Q, int N,
                            double* __restrict__ z) {
    DcMerge const M = mg[blockIdx.y];
    for(int c = blockIdx.x * blockDim.x + threadIdx.x; c < M.n; c += gridDim.x * blockDim.x) {
        int const r = (c < M.n1) ? (M.o + M.n1 - 1) : (M.o + M.n1);
        z[M.o + c] = Q[int64_t(r) * N + M.o + c];
    }
}

// dst[o+r][o+c] = src[o+r][o+perm[c]] inside each merge's square (c < ncols, all n rows).
__global__ void dc_colgather_kernel(const DcMerge* __restrict__ mg, const int* __restrict__ perm, int which,
                                    const double* __restrict__ src, double* __restrict__ dst, int N) {
    DcMerge const M = mg[blockIdx.y];
    const int* pm = perm + (which == 0 ? M.perm_off : M.perm_off + M.n);
    int64_t const cnt = int64_t(M.n) * M.n;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int const r = int(q / M.n), c = int(q % M.n);
        dst[int64_t(M.o + r) * N + M.o + c] = src[int64_t(M.o + r) * N + M.o + pm[c]];
    }
}

// The deflation rotations of each merge, in order, on every row of its square.
__global__ void dc_rot_kernel(const DcMerge* __restrict__ mg, const DcRot* __restrict__ rots, double* __restrict__ Q, int N) {
    DcMerge const M = mg[blockIdx.y];
    if(M.nrot == 0)
        return;
    for(int r = blockIdx.x * blockDim.x + threadIdx.x; r < M.n; r += gridDim.x * blockDim.x) {
        double* row = Q + int64_t(M.o + r) * N + M.o;
        for(int t = 0; t < M.nrot; ++t) {
            DcRot const R = rots[M.rot_off + t];
            double const a = row[R.p], b = row[R.q];
            row[R.p] = R.c * a + R.s * b;
            row[R.q] = -R.s * a + R.c * b;
        }
    }
}

// Root j of 1 + rho sum_i z_i^2/(d_i - lambda) = 0 (d strictly increasing, k >= 1), as
// lambda = d[org] + tau with org the nearer pole, by safeguarded Newton inside a bracket.
// One wavefront per root: the lanes split the sums, every lane runs the same iteration.
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
        bool const done = fabs(tn - tau) <= 2.0 * DBL_EPSILON * fabs(tn) ||
                          (hi - lo) <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi));
        tau = tn;
        if(done)
            break;
    }
    *org = o;
    *tau_out = tau;
}

__global__ void __launch_bounds__(256) dc_secular_kernel(int nm, const DcMerge* __restrict__ mg, const double* __restrict__ dl,
                                                         const double* __restrict__ zl, int* __restrict__ org,
                                                         double* __restrict__ tau) {
    DcMerge const M = mg[blockIdx.y];
    int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for(int jj = blockIdx.x * 4 + wave; jj < M.k; jj += gridDim.x * 4) {
        int o;
        double t;
        secular_root_wave(M.k, dl + M.sec_off, zl + M.sec_off, M.rho, jj, &o, &t);
        if(lane == 0) {
            org[M.sec_off + jj] = o;
            tau[M.sec_off + jj] = t;
        }
    }
}

// Gu-Eisenstat: z_i^2 = prod_j (lambda_j - d_i) / (rho prod_{l != i} (d_l - d_i)), as a product of
// positive ratios; the sign is that of the original z_i.
__global__ void dc_zhat_kernel(const DcMerge* __restrict__ mg, const double* __restrict__ dl, const double* __restrict__ zl,
                               const int* __restrict__ org, const double* __restrict__ tau, double* __restrict__ zh) {
    DcMerge const M = mg[blockIdx.y];
    const double* d = dl + M.sec_off;
    const int* og = org + M.sec_off;
    const double* tu = tau + M.sec_off;
    int const k = M.k;
    for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        double const di = d[i];
        double p = ((d[og[k - 1]] - di) + tu[k - 1]) / M.rho;
        for(int jj = 0; jj < i; ++jj)
            p *= ((d[og[jj]] - di) + tu[jj]) / (d[jj] - di);
        for(int jj = i; jj < k - 1; ++jj)
            p *= ((d[og[jj]] - di) + tu[jj]) / (d[jj + 1] - di);
        zh[M.sec_off + i] = copysign(sqrt(fmax(p, 0.0)), zl[M.sec_off + i]);
    }
}

// U[i][j] = zhat_i / (d_i - lambda_j), columns normalised afterwards.
__global__ void dc_u_kernel(const DcMerge* __restrict__ mg, const double* __restrict__ dl, const double* __restrict__ zh,
                            const int* __restrict__ org, const double* __restrict__ tau, double* __restrict__ U) {
    DcMerge const M = mg[blockIdx.y];
    int const k = M.k;
    const double* d = dl + M.sec_off;
    int64_t const cnt = int64_t(k) * k;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int const i = int(q / k), jj = int(q % k);
        int const o = org[M.sec_off + jj];
        U[M.u_off + q] = zh[M.sec_off + i] / ((d[i] - d[o]) - tau[M.sec_off + jj]);
    }
}

__global__ void __launch_bounds__(256) dc_colnorm_kernel(const DcMerge* __restrict__ mg, double* __restrict__ U) {
    DcMerge const M = mg[blockIdx.y];
    int const k = M.k;
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for(int jj = blockIdx.x * 4 + wave; jj < k; jj += gridDim.x * 4) {
        double* u = U + M.u_off + jj;
        double s = 0.0;
        for(int i = lane; i < k; i += 64)
            s = fma(u[int64_t(i) * k], u[int64_t(i) * k], s);
        s = wave_sum(s);
        double const inv = 1.0 / sqrt(s);
        for(int i = lane; i < k; i += 64)
            u[int64_t(i) * k] *= inv;
    }
}

// Deflated columns (positions k..n-1) pass through: dst[o+r][o+c] = src[o+r][o+c].
__global__ void dc_copycols_kernel(const DcMerge* __restrict__ mg, const double* __restrict__ src, double* __restrict__ dst,
                                   int N) {
    DcMerge const M = mg[blockIdx.y];
    int const w = M.n - M.k;
    if(w <= 0)
        return;
    int64_t const cnt = int64_t(M.n) * w;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int const r = int(q / w), c = M.k + int(q % w);
        dst[int64_t(M.o + r) * N + M.o + c] = src[int64_t(M.o + r) * N + M.o + c];
    }
}

// Leaf eigenvectors (per-leaf arena, n_l x n_l) into the diagonal blocks of Q.
struct DcLeaf {
    int o, n;
    int64_t off;
};
__global__ void dc_place_kernel(const DcLeaf* __restrict__ lv, const double* __restrict__ arena, double* __restrict__ Q, int N) {
    DcLeaf const L = lv[blockIdx.y];
    int64_t const cnt = int64_t(L.n) * L.n;
    for(int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < cnt; q += int64_t(gridDim.x) * blockDim.x) {
        int const r = int(q / L.n), c = int(q % L.n);
        Q[int64_t(L.o + r) * N + L.o + c] = arena[L.off + q];
    }
}

// ---- 3. back-transformation ---------------------------------------------------------------

// T (nb x nb, upper) of I - V T V^T = H_0 ... H_{nb-1} from S = V^T V and tau (one workgroup).
__global__ void wy_t_kernel(int nb, const double* __restrict__ S, const double* __restrict__ tau, double* __restrict__ T) {
    __shared__ double sT[kNB][kNB];
    __shared__ double col[kNB];
    int const r = threadIdx.x;
    if(r < kNB)
        for(int c = 0; c < kNB; ++c)
            sT[r][c] = 0.0;
    __syncthreads();
    for(int i = 0; i < nb; ++i) {
        double const ti = tau[i];
        if(r < i)
            col[r] = -ti * S[r * nb + i];
        __syncthreads();
        if(r < i) {
            double acc = 0.0;
            for(int q = r; q < i; ++q)
                acc += sT[r][q] * col[q];
            sT[r][i] = acc;
        }
        if(r == i)
            sT[i][i] = ti;
        __syncthreads();
    }
    if(r < nb)
        for(int c = 0; c < nb; ++c)
            T[r * nb + c] = sT[r][c];
}

unsigned blocks_for(int64_t n, int per = 256, int cap = 4096) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + per - 1) / per, cap)));
}

template <class T> int put(pomx_ctx* ctx, WsBuf<T>& buf, std::vector<T> const& h) {
    buf = WsBuf<T>(ctx, std::max<size_t>(h.size(), 1));
    if(!buf.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "eigh_large: out of device memory");
    if(!h.empty())
        POMX_HIP(ctx, hipMemcpyAsync(buf.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    return POMX_OK;
}

struct Node {
    int o, n, n1, left, right, height;
};

// Host part of one merge: sort, normalise, deflate.  Fills the permutation (sorted position ->
// local column), the rotation list, the order [non-deflated..., deflated...] and the secular data.
struct MergeHost {
    std::vector<int> perm, order;
    std::vector<DcRot> rots;
    std::vector<double> dl, zl, dval; // non-deflated d, z; values of all positions after deflation
    int k = 0;
    double rho = 0.0;
};

void deflate(int n, int n1, double rho_signed, const double* lam, const double* zraw, MergeHost& H) {
    double const sgn = rho_signed >= 0.0 ? 1.0 : -1.0;
    double const theta = std::fabs(rho_signed);
    std::vector<double> z(zraw, zraw + n);
    for(int c = n1; c < n; ++c)
        z[size_t(c)] *= sgn;
    H.perm.resize(size_t(n));
    std::iota(H.perm.begin(), H.perm.end(), 0);
    std::stable_sort(H.perm.begin(), H.perm.end(), [&](int a, int b) { return lam[a] < lam[b]; });
    std::vector<double> d(static_cast<size_t>(n)), zs(static_cast<size_t>(n));
    for(int c = 0; c < n; ++c) {
        d[size_t(c)] = lam[H.perm[size_t(c)]];
        zs[size_t(c)] = z[size_t(H.perm[size_t(c)])];
    }
    double nz = 0.0;
    for(double x : zs)
        nz += x * x;
    nz = std::sqrt(nz);
    double const rho = theta * nz * nz;
    double dmax = 0.0, zmax = 0.0;
    for(int c = 0; c < n; ++c) {
        if(nz > 0.0)
            zs[size_t(c)] /= nz;
        dmax = std::max(dmax, std::fabs(d[size_t(c)]));
        zmax = std::max(zmax, std::fabs(zs[size_t(c)]));
    }
    double const tol = 8.0 * DBL_EPSILON * std::max(dmax, zmax);
    std::vector<int> nd, df;
    H.rots.clear();
    if(!(nz > 0.0) || rho * zmax <= tol) {
        for(int c = 0; c < n; ++c)
            df.push_back(c);
    } else {
        int pj = -1;
        for(int c = 0; c < n; ++c) {
            if(rho * std::fabs(zs[size_t(c)]) <= tol) {
                df.push_back(c);
                continue;
            }
            if(pj < 0) {
                pj = c;
                continue;
            }
            double const s0 = zs[size_t(pj)], c0 = zs[size_t(c)];
            double const t2 = std::hypot(c0, s0);
            double const t = d[size_t(c)] - d[size_t(pj)];
            double const cc = c0 / t2, ss = -s0 / t2;
            if(std::fabs(t * cc * ss) <= tol) {
                zs[size_t(c)] = t2;
                zs[size_t(pj)] = 0.0;
                H.rots.push_back(DcRot{pj, c, cc, ss});
                double const dp = d[size_t(pj)] * cc * cc + d[size_t(c)] * ss * ss;
                double const dq = d[size_t(pj)] * ss * ss + d[size_t(c)] * cc * cc;
                d[size_t(pj)] = dp;
                d[size_t(c)] = dq;
                df.push_back(pj);
            } else
                nd.push_back(pj);
            pj = c;
        }
        if(pj >= 0)
            nd.push_back(pj);
    }
    H.k = int(nd.size());
    H.rho = rho;
    H.order = nd;
    H.order.insert(H.order.end(), df.begin(), df.end());
    H.dl.clear();
    H.zl.clear();
    for(int c : nd) {
        H.dl.push_back(d[size_t(c)]);
        H.zl.push_back(zs[size_t(c)]);
    }
    H.dval.resize(size_t(n));
    for(int c = 0; c < n; ++c)
        H.dval[size_t(c)] = d[size_t(c)];
}

// Eigen-decomposition of the symmetric tridiagonal (d, e): eigenvalues ascending into lam (host),
// eigenvectors as columns of d_Z (N x N row-major).  Workspace: d_Q2, d_U (N x N each).
int dc_tridiag(pomx_ctx* ctx, int N, std::vector<double> d, std::vector<double> const& e, double* d_Z, double* d_Q2,
               double* d_U, std::vector<double>& lam) {
    hipStream_t s = ctx->stream;
    lam.assign(size_t(N), 0.0);
    if(N == 1) {
        lam[0] = d[0];
        double one = 1.0;
        POMX_HIP(ctx, hipMemcpyAsync(d_Z, &one, sizeof(double), hipMemcpyHostToDevice, s));
        POMX_HIP(ctx, hipStreamSynchronize(s));
        return POMX_OK;
    }
    // tree
    std::vector<Node> nodes;
    std::vector<int> leaves;
    std::function<int(int, int)> build = [&](int o, int n) -> int {
        int id = int(nodes.size());
        nodes.push_back(Node{o, n, 0, -1, -1, 0});
        if(n <= kLeaf) {
            leaves.push_back(id);
            return id;
        }
        int n1 = n / 2;
        int sp = o + n1 - 1;
        double r = std::fabs(e[size_t(sp)]);
        d[size_t(sp)] -= r;
        d[size_t(sp + 1)] -= r;
        int l = build(o, n1);
        int rr = build(o + n1, n - n1);
        nodes[size_t(id)].n1 = n1;
        nodes[size_t(id)].left = l;
        nodes[size_t(id)].right = rr;
        nodes[size_t(id)].height = 1 + std::max(nodes[size_t(l)].height, nodes[size_t(rr)].height);
        return id;
    };
    build(0, N);
    // leaves: dense tridiagonal blocks through the workgroup eigensolver
    std::vector<DcLeaf> lv;
    std::vector<EighJob> jobs;
    std::vector<double> arena;
    int64_t ws = 0;
    int nmax = 1;
    for(int id : leaves) {
        Node const& L = nodes[size_t(id)];
        int64_t off = int64_t(arena.size());
        arena.resize(arena.size() + size_t(L.n) * L.n, 0.0);
        for(int r = 0; r < L.n; ++r) {
            arena[size_t(off + int64_t(r) * L.n + r)] = d[size_t(L.o + r)];
            if(r + 1 < L.n) {
                arena[size_t(off + int64_t(r) * L.n + r + 1)] = e[size_t(L.o + r)];
                arena[size_t(off + int64_t(r + 1) * L.n + r)] = e[size_t(L.o + r)];
            }
        }
        lv.push_back(DcLeaf{L.o, L.n, off});
        jobs.push_back(EighJob{L.n, 0, off, int64_t(L.o), ws});
        ws += eigh_ws_doubles(L.n);
        nmax = std::max(nmax, L.n);
    }
    WsBuf<double> d_arena, d_ws, d_lam;
    WsBuf<EighJob> d_jobs;
    WsBuf<DcLeaf> d_lv;
    WsBuf<int> d_status(ctx, jobs.size());
    int rc;
    if((rc = put(ctx, d_arena, arena)) || (rc = put(ctx, d_jobs, jobs)) || (rc = put(ctx, d_lv, lv)))
        return rc;
    d_ws = WsBuf<double>(ctx, size_t(std::max<int64_t>(ws, 1)));
    d_lam = WsBuf<double>(ctx, size_t(N));
    if(!d_ws.p || !d_lam.p || !d_status.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "eigh_large: out of device memory");
    if((rc = pomx_fill_zero(ctx, d_status.p, sizeof(int) * jobs.size())))
        return rc;
    if((rc = launch_eigh(ctx, d_jobs, int(jobs.size()), nmax, 128, 0, nullptr, d_arena, d_lam, d_ws, d_status)))
        return rc;
    fill_kernel<<<blocks_for(int64_t(N) * N), 256, 0, s>>>(d_Z, int64_t(N) * N, 0.0);
    dc_place_kernel<<<dim3(blocks_for(int64_t(kLeaf) * kLeaf, 256, 64), unsigned(lv.size())), 256, 0, s>>>(d_lv, d_arena,
                                                                                                         d_Z, N);
    POMX_CHECK_LAUNCH(ctx);
    std::vector<int> st(jobs.size());
    POMX_HIP(ctx, hipMemcpyAsync(lam.data(), d_lam.p, sizeof(double) * size_t(N), hipMemcpyDeviceToHost, s));
    POMX_HIP(ctx, hipMemcpyAsync(st.data(), d_status.p, sizeof(int) * st.size(), hipMemcpyDeviceToHost, s));
    POMX_HIP(ctx, hipStreamSynchronize(s));
    for(int v : st)
        if(v)
            return pomx_fail(ctx, POMX_ERR_RANGE, "eigh_large: leaf eigensolver did not converge");
    // merges, level by level
    int const top = nodes[0].height;
    WsBuf<double> d_z(ctx, size_t(N));
    std::vector<double> hz(static_cast<size_t>(N));
    for(int h = 1; h <= top; ++h) {
        std::vector<int> ids;
        for(size_t q = 0; q < nodes.size(); ++q)
            if(nodes[q].height == h)
                ids.push_back(int(q));
        if(ids.empty())
            continue;
        std::vector<DcMerge> mg(ids.size());
        for(size_t q = 0; q < ids.size(); ++q) {
            Node const& nd = nodes[size_t(ids[q])];
            mg[q] = DcMerge{nd.o, nd.n, nd.n1, 0, 0, 0, 0, 0, 0, 0, 0.0};
        }
        unsigned const nm = unsigned(mg.size());
        WsBuf<DcMerge> d_mg;
        if((rc = put(ctx, d_mg, mg)))
            return rc;
        dc_z_kernel<<<dim3(blocks_for(N, 256, 64), nm), 256, 0, s>>>(int(nm), d_mg, d_Z, N, d_z);
        POMX_CHECK_LAUNCH(ctx);
        POMX_HIP(ctx, hipMemcpyAsync(hz.data(), d_z.p, sizeof(double) * size_t(N), hipMemcpyDeviceToHost, s));
        POMX_HIP(ctx, hipStreamSynchronize(s));
        // host: deflation
        std::vector<MergeHost> H(mg.size());
        std::vector<int> perms;
        std::vector<DcRot> rots;
        std::vector<double> dl, zl;
        std::vector<pomx_gemm_job> gj;
        int64_t u_len = 0;
        for(size_t q = 0; q < mg.size(); ++q) {
            Node const& nd = nodes[size_t(ids[q])];
            double const rho_signed = e[size_t(nd.o + nd.n1 - 1)];
            deflate(nd.n, nd.n1, rho_signed, lam.data() + nd.o, hz.data() + nd.o, H[q]);
            mg[q].k = H[q].k;
            mg[q].perm_off = int64_t(perms.size());
            perms.insert(perms.end(), H[q].perm.begin(), H[q].perm.end());
            perms.insert(perms.end(), H[q].order.begin(), H[q].order.end());
            mg[q].rot_off = int64_t(rots.size());
            mg[q].nrot = int(H[q].rots.size());
            rots.insert(rots.end(), H[q].rots.begin(), H[q].rots.end());
            mg[q].sec_off = int64_t(dl.size());
            dl.insert(dl.end(), H[q].dl.begin(), H[q].dl.end());
            zl.insert(zl.end(), H[q].zl.begin(), H[q].zl.end());
            mg[q].u_off = u_len;
            mg[q].rho = H[q].rho;
            int const k = H[q].k;
            if(k > 0)
                gj.push_back(pomx_gemm_job{nd.n, k, k, 0, 0, N, k, N, int64_t(nd.o) * N + nd.o, u_len,
                                           int64_t(nd.o) * N + nd.o, 1.0, 0.0});
            u_len += int64_t(k) * k;
        }
        if(u_len > int64_t(N) * N)
            return pomx_fail(ctx, POMX_ERR_RANGE, "eigh_large: rank-one eigenvector workspace exceeded");
        WsBuf<int> d_perm, d_org;
        WsBuf<DcRot> d_rot;
        WsBuf<double> d_dl, d_zl, d_tau, d_zh;
        if((rc = put(ctx, d_mg, mg)) || (rc = put(ctx, d_perm, perms)) || (rc = put(ctx, d_rot, rots)) ||
           (rc = put(ctx, d_dl, dl)) || (rc = put(ctx, d_zl, zl)))
            return rc;
        size_t const nsec = std::max<size_t>(dl.size(), 1);
        d_org = WsBuf<int>(ctx, nsec);
        d_tau = WsBuf<double>(ctx, nsec);
        d_zh = WsBuf<double>(ctx, nsec);
        if(!d_org.p || !d_tau.p || !d_zh.p)
            return pomx_fail(ctx, POMX_ERR_HIP, "eigh_large: out of device memory");
        int maxn = 0, maxk = 0;
        for(auto const& m : mg) {
            maxn = std::max(maxn, m.n);
            maxk = std::max(maxk, m.k);
        }
        unsigned const gsq = blocks_for(int64_t(maxn) * maxn, 256, 2048);
        // permute columns into sorted order, apply deflation rotations, then order as
        // [non-deflated..., deflated...]
        dc_colgather_kernel<<<dim3(gsq, nm), 256, 0, s>>>(d_mg, d_perm, 0, d_Z, d_Q2, N);
        dc_rot_kernel<<<dim3(blocks_for(maxn, 256, 64), nm), 256, 0, s>>>(d_mg, d_rot, d_Q2, N);
        dc_colgather_kernel<<<dim3(gsq, nm), 256, 0, s>>>(d_mg, d_perm, 1, d_Q2, d_Z, N);
        if(maxk > 0) {
            dc_secular_kernel<<<dim3(blocks_for(maxk, 4, 4096), nm), 256, 0, s>>>(int(nm), d_mg, d_dl, d_zl, d_org, d_tau);
            dc_zhat_kernel<<<dim3(blocks_for(maxk, 64, 1024), nm), 64, 0, s>>>(d_mg, d_dl, d_zl, d_org, d_tau, d_zh);
            dc_u_kernel<<<dim3(blocks_for(int64_t(maxk) * maxk, 256, 2048), nm), 256, 0, s>>>(d_mg, d_dl, d_zh, d_org,
                                                                                             d_tau, d_U);
            dc_colnorm_kernel<<<dim3(blocks_for(maxk, 4, 1024), nm), 256, 0, s>>>(d_mg, d_U);
            POMX_CHECK_LAUNCH(ctx);
            // Q2[:, 0:k] = Z[:, 0:k] U  (all merges of the level in one launch)
            std::vector<GemmTile> tiles;
            gemm_tiles(gj.data(), int(gj.size()), tiles);
            WsBuf<GemmTile> d_tiles;
            WsBuf<pomx_gemm_job> d_gj;
            if((rc = put(ctx, d_tiles, tiles)) || (rc = put(ctx, d_gj, gj)))
                return rc;
            if((rc = launch_gemm(ctx, 0, 0, d_tiles, int(tiles.size()), d_gj, d_Z, d_U, d_Q2)))
                return rc;
            POMX_HIP(ctx, hipStreamSynchronize(s)); // tiles/jobs are pooled temporaries
        }
        dc_copycols_kernel<<<dim3(gsq, nm), 256, 0, s>>>(d_mg, d_Z, d_Q2, N);
        POMX_CHECK_LAUNCH(ctx);
        // eigenvalues of the level, final ascending order
        std::vector<int> horg(dl.size());
        std::vector<double> htau(dl.size());
        if(!dl.empty()) {
            POMX_HIP(ctx, hipMemcpyAsync(horg.data(), d_org.p, sizeof(int) * dl.size(), hipMemcpyDeviceToHost, s));
            POMX_HIP(ctx, hipMemcpyAsync(htau.data(), d_tau.p, sizeof(double) * dl.size(), hipMemcpyDeviceToHost, s));
        }
        POMX_HIP(ctx, hipStreamSynchronize(s));
        std::vector<int> fperm;
        for(size_t q = 0; q < mg.size(); ++q) {
            MergeHost const& M = H[q];
            int const n = mg[q].n, k = M.k;
            std::vector<double> vals(static_cast<size_t>(n));
            for(int jj = 0; jj < k; ++jj)
                vals[size_t(jj)] = M.dl[size_t(horg[size_t(mg[q].sec_off + jj)])] + htau[size_t(mg[q].sec_off + jj)];
            for(int c = k; c < n; ++c)
                vals[size_t(c)] = M.dval[size_t(M.order[size_t(c)])];
            std::vector<int> f(static_cast<size_t>(n));
            std::iota(f.begin(), f.end(), 0);
            std::stable_sort(f.begin(), f.end(), [&](int a, int b) { return vals[size_t(a)] < vals[size_t(b)]; });
            for(int c = 0; c < n; ++c)
                lam[size_t(mg[q].o + c)] = vals[size_t(f[size_t(c)])];
            mg[q].perm_off = int64_t(fperm.size());
            fperm.insert(fperm.end(), f.begin(), f.end());
        }
        WsBuf<int> d_fperm;
        if((rc = put(ctx, d_mg, mg)) || (rc = put(ctx, d_fperm, fperm)))
            return rc;
        dc_colgather_kernel<<<dim3(gsq, nm), 256, 0, s>>>(d_mg, d_fperm, 0, d_Q2, d_Z, N);
        POMX_CHECK_LAUNCH(ctx);
        POMX_HIP(ctx, hipStreamSynchronize(s));
    }
    return POMX_OK;
}

} // namespace

int eigh_large(pomx_ctx* ctx, int n, double* d_A, double* d_w) {
    hipStream_t s = ctx->stream;
    if(n <= 0)
        return POMX_OK;
    int64_t const nn = int64_t(n) * n;
    WsBuf<double> d_V(ctx, size_t(nn)), d_Z(ctx, size_t(nn)), d_Q2(ctx, size_t(nn)), d_U(ctx, size_t(nn));
    WsBuf<double> d_W(ctx, size_t(n) * kNB), d_v(ctx, size_t(n)), d_wraw(ctx, size_t(n)), d_part(ctx, size_t(kSymvGrid) * 64);
    WsBuf<double> d_a2(ctx, kNB), d_rowpart(ctx, 256), d_wpart(ctx, 256);
    WsBuf<TrdScal> d_sc(ctx, 1);
    WsBuf<unsigned> d_cnt(ctx, 4);
    WsBuf<double> d_d(ctx, size_t(n)), d_e(ctx, size_t(n)), d_tau(ctx, size_t(n));
    WsBuf<double> d_S(ctx, size_t(kNB) * kNB), d_T(ctx, size_t(kNB) * kNB), d_W1(ctx, size_t(kNB) * n),
        d_W2(ctx, size_t(kNB) * n);
    for(void* p : {(void*)d_V.p, (void*)d_Z.p, (void*)d_Q2.p, (void*)d_U.p, (void*)d_W.p, (void*)d_v.p, (void*)d_wraw.p,
                   (void*)d_part.p, (void*)d_d.p, (void*)d_e.p, (void*)d_tau.p, (void*)d_S.p, (void*)d_T.p, (void*)d_W1.p,
                   (void*)d_W2.p, (void*)d_a2.p, (void*)d_rowpart.p, (void*)d_wpart.p, (void*)d_sc.p, (void*)d_cnt.p})
        if(!p)
            return pomx_fail(ctx, POMX_ERR_HIP, "eigh_large: out of device memory for a block of size " + std::to_string(n));
    fill_kernel<<<blocks_for(nn), 256, 0, s>>>(d_V, nn, 0.0);
    fill_kernel<<<blocks_for(int64_t(n) * kNB), 256, 0, s>>>(d_W, int64_t(n) * kNB, 0.0);
    fill_kernel<<<blocks_for(n), 256, 0, s>>>(d_e, n, 0.0);
    fill_kernel<<<blocks_for(n), 256, 0, s>>>(d_tau, n, 0.0);
    fill_kernel<<<1, 64, 0, s>>>(d_a2, kNB, 0.0);
    POMX_CHECK_LAUNCH(ctx);
    int rc;
    if((rc = pomx_fill_zero(ctx, d_cnt.p, 4 * sizeof(unsigned))))
        return rc;
    // 1. tridiagonalisation
    for(int j0 = 0; j0 < n - 1; j0 += kNB) {
        int const nb = std::min(kNB, n - 1 - j0);
        for(int i = 0; i < nb; ++i) {
            int const j = j0 + i, m = n - j - 1;
            int const g1 = std::max(1, std::min(256, (n - j + 255) / 256));
            trd_row_kernel<<<g1, 256, 0, s>>>(n, j, j0, i, d_A, d_V, d_W, d_a2, d_v, d_d, d_e, d_tau, d_rowpart, d_cnt.p);
            int const g2 = std::max(1, std::min(kSymvGrid, (m + 15) / 16));
            trd_symv_kernel<<<g2, 256, 0, s>>>(n, j, j0, i, d_A, d_v, d_V, d_W, d_a2, d_wraw, d_part, d_sc.p, d_cnt.p + 1);
            int const g3 = std::max(1, std::min(256, (m + 255) / 256));
            trd_w_kernel<<<g3, 256, 0, s>>>(n, j, j0, i, d_wraw, d_v, d_V, d_W, d_a2, d_tau, d_sc.p, d_wpart, d_cnt.p + 2);
        }
        POMX_CHECK_LAUNCH(ctx);
        int const j1 = j0 + nb, mt = n - j1;
        if(mt > 0) { // A22 -= V2 W2^T + W2 V2^T
            trd_fold_kernel<<<blocks_for(int64_t(mt) * nb), 256, 0, s>>>(n, j0, nb, j1, d_V, d_W, d_a2);
            pomx_gemm_job g1{mt, mt, nb, 0, 1, n, kNB, n, int64_t(j1) * n + j0, int64_t(j1) * kNB, int64_t(j1) * n + j1,
                             -1.0, 1.0};
            pomx_gemm_job g2{mt, mt, nb, 0, 1, kNB, n, n, int64_t(j1) * kNB, int64_t(j1) * n + j0, int64_t(j1) * n + j1,
                             -1.0, 1.0};
            if((rc = launch_gemm1(ctx, g1, d_V, d_W, d_A)) || (rc = launch_gemm1(ctx, g2, d_W, d_V, d_A)))
                return rc;
        }
    }
    copy_diag_kernel<<<1, 64, 0, s>>>(n, d_A, d_d);
    POMX_CHECK_LAUNCH(ctx);
    std::vector<double> hd(static_cast<size_t>(n)), he(static_cast<size_t>(n)), htau(static_cast<size_t>(n));
    POMX_HIP(ctx, hipMemcpyAsync(hd.data(), d_d.p, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, s));
    POMX_HIP(ctx, hipMemcpyAsync(he.data(), d_e.p, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, s));
    POMX_HIP(ctx, hipMemcpyAsync(htau.data(), d_tau.p, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, s));
    POMX_HIP(ctx, hipStreamSynchronize(s));
    for(int q = 0; q < n; ++q)
        if(!std::isfinite(hd[size_t(q)]) || !std::isfinite(he[size_t(q)]))
            return pomx_fail(ctx, POMX_ERR_RANGE, "eigh_large: non-finite value in the tridiagonal reduction");
    // 2. divide and conquer
    std::vector<double> lam;
    if((rc = dc_tridiag(ctx, n, hd, he, d_Z, d_Q2, d_U, lam)))
        return rc;
    // 3. back-transformation X = H_0 ... H_{n-2} Z, panels last to first, X in place of A
    POMX_HIP(ctx, hipMemcpyAsync(d_A, d_Z.p, sizeof(double) * size_t(nn), hipMemcpyDeviceToDevice, s));
    int const last = ((n - 2) / kNB) * kNB;
    for(int j0 = last; j0 >= 0; j0 -= kNB) {
        int const nb = std::min(kNB, n - 1 - j0);
        if(nb <= 0)
            continue;
        int const rows = n - j0 - 1; // V has rows j0+1 .. n-1
        int64_t const voff = int64_t(j0 + 1) * n + j0;
        pomx_gemm_job gs{nb, nb, rows, 1, 0, n, n, nb, voff, voff, 0, 1.0, 0.0};
        if((rc = launch_gemm1(ctx, gs, d_V, d_V, d_S)))
            return rc;
        wy_t_kernel<<<1, 64, 0, s>>>(nb, d_S, d_tau.p + j0, d_T);
        pomx_gemm_job g1{nb, n, rows, 1, 0, n, n, n, voff, int64_t(j0 + 1) * n, 0, 1.0, 0.0};
        pomx_gemm_job g2{nb, n, nb, 0, 0, nb, n, n, 0, 0, 0, 1.0, 0.0};
        pomx_gemm_job g3{rows, n, nb, 0, 0, n, n, n, voff, 0, int64_t(j0 + 1) * n, -1.0, 1.0};
        if((rc = launch_gemm1(ctx, g1, d_V, d_A, d_W1)) || (rc = launch_gemm1(ctx, g2, d_T, d_W1, d_W2)) ||
           (rc = launch_gemm1(ctx, g3, d_V, d_W2, d_A)))
            return rc;
    }
    POMX_HIP(ctx, hipMemcpyAsync(d_w, lam.data(), sizeof(double) * size_t(n), hipMemcpyHostToDevice, s));
    POMX_HIP(ctx, hipStreamSynchronize(s));
    return POMX_OK;
}

extern "C" int pomx_eigh_tridiag(pomx_ctx* ctx, int n, const double* h_d, const double* h_e, double* d_Z, double* h_w) {
    POMX_REQUIRE(ctx, n >= 1 && h_d && (h_e || n == 1) && d_Z && h_w, "pomx_eigh_tridiag: bad arguments");
    int64_t const nn = int64_t(n) * n;
    WsBuf<double> d_Q2(ctx, size_t(nn)), d_U(ctx, size_t(nn));
    if(!d_Q2.p || !d_U.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_eigh_tridiag: out of device memory");
    std::vector<double> d(h_d, h_d + n), e(size_t(n), 0.0), lam;
    for(int q = 0; q + 1 < n; ++q)
        e[size_t(q)] = h_e[q];
    int rc = dc_tridiag(ctx, n, d, e, d_Z, d_Q2, d_U, lam);
    if(rc)
        return rc;
    std::copy(lam.begin(), lam.end(), h_w);
    return POMX_OK;
}
