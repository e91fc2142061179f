// Batched FP64 GEMM on the matrix cores: C = alpha * op(A) op(B) + beta * C for a list of
// independent jobs of arbitrary shapes (row-major operands at offsets into three arenas).
//
// Used by the adjoint (U^T Ubar, U K, X Abar^T, U_t Abar over every block / block pair) and by
// the large-block eigensolver.  Tile: 64 x 64 outputs per workgroup of four waves, each wave
// 32 x 32 as 2 x 2 v_mfma_f64_16x16x4f64 accumulators; K advances 16 at a time through LDS,
// with the next K slab prefetched into registers while the current one is multiplied.
#include "pomx_internal.hpp"

#include <algorithm>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// One 64 x 64 output tile (rows i0.., columns j0..) of job J.
template <int TA, int TB>
__device__ inline void gemm_tile(pomx_gemm_job const& J, int const i0, int const j0, const double* __restrict__ A,
                                 const double* __restrict__ B, double* __restrict__ C) {
    __shared__ double As[16][64 + 4];
    __shared__ double Bs[16][64 + 4];
    int const M = J.m, N = J.n, K = J.k;
    const double* __restrict__ a = A + J.a_off;
    const double* __restrict__ b = B + J.b_off;
    int const tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    int const wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

    // slab coordinates of this thread's 4 A and 4 B elements
    int ak[4], ai[4], bk[4], bj[4];
#pragma unroll
    for(int r = 0; r < 4; ++r) {
        int e = tid + 256 * r;
        if(TA == 0) { // op(A)(i, t) = a[i*lda + t]: consecutive threads along t
            ak[r] = e & 15;
            ai[r] = e >> 4;
        } else { // op(A)(i, t) = a[t*lda + i]: consecutive threads along i
            ai[r] = e & 63;
            ak[r] = e >> 6;
        }
        if(TB == 0) { // op(B)(t, j) = b[t*ldb + j]
            bj[r] = e & 63;
            bk[r] = e >> 6;
        } else { // op(B)(t, j) = b[j*ldb + t]
            bk[r] = e & 15;
            bj[r] = e >> 4;
        }
    }
    auto load = [&](int t0, double* ra, double* rb) {
#pragma unroll
        for(int r = 0; r < 4; ++r) {
            int t = t0 + ak[r], i = i0 + ai[r];
            ra[r] = (t < K && i < M) ? (TA == 0 ? a[int64_t(i) * J.lda + t] : a[int64_t(t) * J.lda + i]) : 0.0;
            int tb = t0 + bk[r], j = j0 + bj[r];
            rb[r] = (tb < K && j < N) ? (TB == 0 ? b[int64_t(tb) * J.ldb + j] : b[int64_t(j) * J.ldb + tb]) : 0.0;
        }
    };

    dbl4 acc[2][2];
    for(int x = 0; x < 2; ++x)
        for(int y = 0; y < 2; ++y)
            acc[x][y] = dbl4{0.0, 0.0, 0.0, 0.0};
    double ra[4], rb[4];
    load(0, ra, rb);
    for(int t0 = 0; t0 < K; t0 += 16) {
        __syncthreads();
#pragma unroll
        for(int r = 0; r < 4; ++r) {
            As[ak[r]][ai[r]] = ra[r];
            Bs[bk[r]][bj[r]] = rb[r];
        }
        __syncthreads();
        if(t0 + 16 < K)
            load(t0 + 16, ra, rb);
#pragma unroll
        for(int kq = 0; kq < 4; ++kq) {
            int kr = kq * 4 + (lane >> 4);
            double a0 = As[kr][wm + (lane & 15)], a1 = As[kr][wm + 16 + (lane & 15)];
            double b0 = Bs[kr][wn + (lane & 15)], b1 = Bs[kr][wn + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
    }
    double* __restrict__ c = C + J.c_off;
    double const alpha = J.alpha, beta = J.beta;
#pragma unroll
    for(int mt = 0; mt < 2; ++mt)
#pragma unroll
        for(int nt = 0; nt < 2; ++nt)
#pragma unroll
            for(int r = 0; r < 4; ++r) {
                // v_mfma_f64_16x16x4f64 result layout: D[i][j] in lane j + 16*(i%4), register i/4
                int i = i0 + wm + mt * 16 + (lane >> 4) + 4 * r;
                int j = j0 + wn + nt * 16 + (lane & 15);
                if(i < M && j < N) {
                    double* p = c + int64_t(i) * J.ldc + j;
                    double v = alpha * acc[mt][nt][r];
                    *p = (beta == 0.0) ? v : fma(beta, *p, v);
                }
            }
}

template <int TA, int TB>
__global__ void __launch_bounds__(256) gemm_kernel(const GemmTile* __restrict__ tiles, const pomx_gemm_job* __restrict__ jobs,
                                                   const double* __restrict__ A, const double* __restrict__ B,
                                                   double* __restrict__ C) {
    GemmTile const T = tiles[blockIdx.x];
    gemm_tile<TA, TB>(jobs[T.job], T.i0, T.j0, A, B, C);
}

// Job blockIdx.y, tile blockIdx.x (row-major over the job's 64 x 64 tiles).  Tiles outside the
// job exit at once, so the launch can be sized for the largest job and the job sizes may be
// written on the device by an earlier kernel.
template <int TA, int TB, bool UPPER>
__global__ void __launch_bounds__(256) gemmj_kernel(const pomx_gemm_job* __restrict__ jobs, const double* __restrict__ A,
                                                    const double* __restrict__ B, double* __restrict__ C) {
    pomx_gemm_job const J = jobs[blockIdx.y];
    int const tn = (J.n + 63) / 64, tm = (J.m + 63) / 64;
    // grid-stride over the job's tiles (uniform per workgroup: the barriers inside stay uniform);
    // UPPER: only tiles that touch the upper triangle (column >= row) of the output
    for(int t = blockIdx.x; t < tm * tn; t += gridDim.x) {
        int const i0 = (t / tn) * 64, j0 = (t % tn) * 64;
        if(UPPER && j0 + 63 < i0)
            continue;
        __syncthreads();
        gemm_tile<TA, TB>(J, i0, j0, A, B, C);
    }
}

// One job; the tile is (blockIdx.y, blockIdx.x), so nothing needs to be uploaded per launch.
template <int TA, int TB>
__global__ void __launch_bounds__(256) gemm1_kernel(pomx_gemm_job const J, const double* __restrict__ A,
                                                    const double* __restrict__ B, double* __restrict__ C) {
    gemm_tile<TA, TB>(J, int(blockIdx.y) * 64, int(blockIdx.x) * 64, A, B, C);
}

} // namespace

void gemm_tiles(const pomx_gemm_job* jobs, int njobs, std::vector<GemmTile>& tiles) {
    for(int q = 0; q < njobs; ++q)
        for(int i0 = 0; i0 < jobs[q].m; i0 += 64)
            for(int j0 = 0; j0 < jobs[q].n; j0 += 64)
                tiles.push_back(GemmTile{q, i0, j0, 0});
}

int launch_gemm(pomx_ctx* ctx, int trans_a, int trans_b, const GemmTile* d_tiles, int ntiles,
                const pomx_gemm_job* d_jobs, const double* d_A, const double* d_B, double* d_C) {
    if(ntiles <= 0)
        return POMX_OK;
    unsigned g = unsigned(ntiles);
    if(!trans_a && !trans_b)
        gemm_kernel<0, 0><<<g, 256, 0, ctx->stream>>>(d_tiles, d_jobs, d_A, d_B, d_C);
    else if(!trans_a && trans_b)
        gemm_kernel<0, 1><<<g, 256, 0, ctx->stream>>>(d_tiles, d_jobs, d_A, d_B, d_C);
    else if(trans_a && !trans_b)
        gemm_kernel<1, 0><<<g, 256, 0, ctx->stream>>>(d_tiles, d_jobs, d_A, d_B, d_C);
    else
        gemm_kernel<1, 1><<<g, 256, 0, ctx->stream>>>(d_tiles, d_jobs, d_A, d_B, d_C);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

int launch_gemm1(pomx_ctx* ctx, pomx_gemm_job const& J, const double* d_A, const double* d_B, double* d_C) {
    if(J.m <= 0 || J.n <= 0)
        return POMX_OK;
    if((J.m + 63) / 64 > 65535)
        return pomx_fail(ctx, POMX_ERR_RANGE, "launch_gemm1: more than 65535 row tiles");
    dim3 g(unsigned((J.n + 63) / 64), unsigned((J.m + 63) / 64));
    if(!J.trans_a && !J.trans_b)
        gemm1_kernel<0, 0><<<g, 256, 0, ctx->stream>>>(J, d_A, d_B, d_C);
    else if(!J.trans_a && J.trans_b)
        gemm1_kernel<0, 1><<<g, 256, 0, ctx->stream>>>(J, d_A, d_B, d_C);
    else if(J.trans_a && !J.trans_b)
        gemm1_kernel<1, 0><<<g, 256, 0, ctx->stream>>>(J, d_A, d_B, d_C);
    else
        gemm1_kernel<1, 1><<<g, 256, 0, ctx->stream>>>(J, d_A, d_B, d_C);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

int launch_gemm_jobs(pomx_ctx* ctx, int trans_a, int trans_b, const pomx_gemm_job* d_jobs, int njobs, int64_t max_tiles,
                     const double* d_A, const double* d_B, double* d_C, bool upper) {
    if(njobs <= 0 || max_tiles <= 0)
        return POMX_OK;
    if(njobs > 65535)
        return pomx_fail(ctx, POMX_ERR_RANGE, "launch_gemm_jobs: more than 65535 jobs in one launch");
    // about 32k workgroups in all: many small jobs loop over their tiles instead of launching
    // one workgroup per tile
    int64_t const gx = std::max<int64_t>(1, std::min<int64_t>(max_tiles, (32768 + njobs - 1) / njobs));
    dim3 const g{unsigned(gx), unsigned(njobs), 1u};
    if(upper && trans_a && !trans_b)
        gemmj_kernel<1, 0, true><<<g, 256, 0, ctx->stream>>>(d_jobs, d_A, d_B, d_C);
    else if(upper)
        return pomx_fail(ctx, POMX_ERR_ARG, "launch_gemm_jobs: upper-triangle output only for op(A)^T op(B)");
    else if(!trans_a && !trans_b)
        gemmj_kernel<0, 0, false><<<g, 256, 0, ctx->stream>>>(d_jobs, d_A, d_B, d_C);
    else if(!trans_a && trans_b)
        gemmj_kernel<0, 1, false><<<g, 256, 0, ctx->stream>>>(d_jobs, d_A, d_B, d_C);
    else if(trans_a && !trans_b)
        gemmj_kernel<1, 0, false><<<g, 256, 0, ctx->stream>>>(d_jobs, d_A, d_B, d_C);
    else
        gemmj_kernel<1, 1, false><<<g, 256, 0, ctx->stream>>>(d_jobs, d_A, d_B, d_C);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

bool gemm_job_ok(pomx_gemm_job const& j) {
    if(j.m < 0 || j.n < 0 || j.k < 0 || j.a_off < 0 || j.b_off < 0 || j.c_off < 0)
        return false;
    int const ra = j.trans_a ? j.k : j.m, ca = j.trans_a ? j.m : j.k; // stored rows x cols of A
    int const rb = j.trans_b ? j.n : j.k, cb = j.trans_b ? j.k : j.n;
    return (ra == 0 || j.lda >= ca) && (rb == 0 || j.ldb >= cb) && (j.m == 0 || j.ldc >= j.n);
}

extern "C" int pomx_gemm_batched(pomx_ctx* ctx, int njobs, const pomx_gemm_job* h_jobs, const double* d_A,
                                 const double* d_B, double* d_C) {
    POMX_REQUIRE(ctx, njobs >= 0 && (njobs == 0 || (h_jobs && d_A && d_B && d_C)), "pomx_gemm_batched: bad arguments");
    if(njobs == 0)
        return POMX_OK;
    for(int q = 0; q < njobs; ++q)
        POMX_REQUIRE(ctx, gemm_job_ok(h_jobs[q]), "pomx_gemm_batched: job " + std::to_string(q) +
                                                     " has a negative size/offset or a leading dimension too small");
    // one launch per transposition pattern
    for(int ta = 0; ta < 2; ++ta)
        for(int tb = 0; tb < 2; ++tb) {
            std::vector<pomx_gemm_job> js;
            for(int q = 0; q < njobs; ++q)
                if((h_jobs[q].trans_a != 0) == (ta != 0) && (h_jobs[q].trans_b != 0) == (tb != 0))
                    js.push_back(h_jobs[q]);
            if(js.empty())
                continue;
            std::vector<GemmTile> tiles;
            gemm_tiles(js.data(), int(js.size()), tiles);
            if(tiles.empty())
                continue;
            WsBuf<GemmTile> d_tiles(ctx, tiles.size());
            WsBuf<pomx_gemm_job> d_jobs(ctx, js.size());
            if(!d_tiles.p || !d_jobs.p)
                return pomx_fail(ctx, POMX_ERR_HIP, "pomx_gemm_batched: out of device memory");
            POMX_HIP(ctx, hipMemcpyAsync(d_tiles.p, tiles.data(), sizeof(GemmTile) * tiles.size(), hipMemcpyHostToDevice,
                                         ctx->stream));
            POMX_HIP(ctx, hipMemcpyAsync(d_jobs.p, js.data(), sizeof(pomx_gemm_job) * js.size(), hipMemcpyHostToDevice,
                                         ctx->stream));
            int rc = launch_gemm(ctx, ta, tb, d_tiles, int(tiles.size()), d_jobs, d_A, d_B, d_C);
            if(rc)
                return rc;
            // the pooled descriptors are released below: wait for the launch that reads them
            POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
    return POMX_OK;
}
