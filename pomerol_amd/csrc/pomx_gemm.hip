// Batched FP64 GEMM on the matrix cores: C = alpha * op(A) op(B) + beta * C for a list of
// independent jobs of arbitrary shapes (row-major operands at offsets into three arenas).
//
// Used by the adjoint (U^T Ubar, U K, X Abar^T, U_t Abar over every block / block pair) and by
// the large-block eigensolver.  Tiles: 64 x 64 outputs per workgroup of four waves, each wave
// 32 x 32 as 2 x 2 v_mfma_f64_16x16x4f64 accumulators, or, for large jobs, 128 x 128 outputs per
// workgroup of eight waves of 32 x 64 (half the operand reads per multiply-add); K advances 16
// at a time through LDS, with the next K slab prefetched into registers while the current one
// is multiplied.
#include "pomx_internal.hpp"

#include <algorithm>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// One TT x TT output tile (rows i0.., columns j0..) of job J, computed by WGM x 2 waves of
// (TT/WGM) x (TT/2); the workgroup has 128 * WGM threads.
template <int TA, int TB, int TT = 64, int WGM = 2>
__device__ inline void gemm_tile(pomx_gemm_job const& J, int const i0, int const j0, const double* __restrict__ A,
                                 const double* __restrict__ B, double* __restrict__ C) {
    constexpr int NT = 128 * WGM;               // threads
    constexpr int WM = TT / WGM, WN = TT / 2;   // wave tile
    constexpr int NBM = WM / 16, NBN = WN / 16; // 16 x 16 blocks per wave
    constexpr int NE = 16 * TT / NT;            // slab elements per thread and operand
    static_assert(NT % 64 == 0 && (16 * TT) % NT == 0 && NT % TT == 0 && WM % 16 == 0, "tile shape");
    __shared__ double As[16][TT + 4];
    __shared__ double Bs[16][TT + 4];
    int const M = J.m, N = J.n, K = J.k;
    const double* __restrict__ a = A + J.a_off;
    const double* __restrict__ b = B + J.b_off;
    int const tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    int const wm = (wave >> 1) * WM, wn = (wave & 1) * WN;

    // slab coordinates of this thread's NE A and NE B elements
    int ak[NE], ai[NE], bk[NE], bj[NE];
#pragma unroll
    for(int r = 0; r < NE; ++r) {
        int e = tid + NT * r;
        if(TA == 0) { // op(A)(i, t) = a[i*lda + t]: consecutive threads along t
            ak[r] = e & 15;
            ai[r] = e >> 4;
        } else { // op(A)(i, t) = a[t*lda + i]: consecutive threads along i
            ai[r] = e % TT;
            ak[r] = e / TT;
        }
        if(TB == 0) { // op(B)(t, j) = b[t*ldb + j]
            bj[r] = e % TT;
            bk[r] = e / TT;
        } else { // op(B)(t, j) = b[j*ldb + t]
            bk[r] = e & 15;
            bj[r] = e >> 4;
        }
    }
    auto load = [&](int t0, double* ra, double* rb) {
#pragma unroll
        for(int r = 0; r < NE; ++r) {
            int t = t0 + ak[r], i = i0 + ai[r];
            ra[r] = (t < K && i < M) ? (TA == 0 ? a[int64_t(i) * J.lda + t] : a[int64_t(t) * J.lda + i]) : 0.0;
            int tb = t0 + bk[r], j = j0 + bj[r];
            rb[r] = (tb < K && j < N) ? (TB == 0 ? b[int64_t(tb) * J.ldb + j] : b[int64_t(j) * J.ldb + tb]) : 0.0;
        }
    };

    dbl4 acc[NBM][NBN];
#pragma unroll
    for(int x = 0; x < NBM; ++x)
#pragma unroll
        for(int y = 0; y < NBN; ++y)
            acc[x][y] = dbl4{0.0, 0.0, 0.0, 0.0};
    double ra[NE], rb[NE];
    load(0, ra, rb);
    double* __restrict__ c = C + J.c_off;
    double const alpha = J.alpha, beta = J.beta;
    // 64 x 64 tiles with beta != 0 (updates, often with a short K): the C tile is loaded before
    // the products, so its latency overlaps them
    constexpr bool PRE = TT == 64;
    double cpre[PRE ? NBM : 1][PRE ? NBN : 1][4];
    if constexpr(PRE) {
        if(beta != 0.0) {
#pragma unroll
            for(int mt = 0; mt < NBM; ++mt)
#pragma unroll
                for(int nt = 0; nt < NBN; ++nt)
#pragma unroll
                    for(int r = 0; r < 4; ++r) {
                        int const i = i0 + wm + mt * 16 + (lane >> 4) + 4 * r;
                        int const j = j0 + wn + nt * 16 + (lane & 15);
                        cpre[mt][nt][r] = (i < M && j < N) ? c[int64_t(i) * J.ldc + j] : 0.0;
                    }
        }
    }
    for(int t0 = 0; t0 < K; t0 += 16) {
        __syncthreads();
#pragma unroll
        for(int r = 0; r < NE; ++r) {
            As[ak[r]][ai[r]] = ra[r];
            Bs[bk[r]][bj[r]] = rb[r];
        }
        __syncthreads();
        if(t0 + 16 < K)
            load(t0 + 16, ra, rb);
#pragma unroll
        for(int kq = 0; kq < 4; ++kq) {
            int kr = kq * 4 + (lane >> 4);
            double av[NBM], bv[NBN];
#pragma unroll
            for(int m = 0; m < NBM; ++m)
                av[m] = As[kr][wm + 16 * m + (lane & 15)];
#pragma unroll
            for(int n = 0; n < NBN; ++n)
                bv[n] = Bs[kr][wn + 16 * n + (lane & 15)];
#pragma unroll
            for(int m = 0; m < NBM; ++m)
#pragma unroll
                for(int n = 0; n < NBN; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bv[n], acc[m][n], 0, 0, 0);
        }
    }
#pragma unroll
    for(int mt = 0; mt < NBM; ++mt)
#pragma unroll
        for(int nt = 0; nt < NBN; ++nt)
#pragma unroll
            for(int r = 0; r < 4; ++r) {
                // v_mfma_f64_16x16x4f64 result layout: D[i][j] in lane j + 16*(i%4), register i/4
                int i = i0 + wm + mt * 16 + (lane >> 4) + 4 * r;
                int j = j0 + wn + nt * 16 + (lane & 15);
                if(i < M && j < N) {
                    double* p = c + int64_t(i) * J.ldc + j;
                    double v = alpha * acc[mt][nt][r];
                    if constexpr(PRE)
                        *p = (beta == 0.0) ? v : fma(beta, cpre[mt][nt][r], v);
                    else
                        *p = (beta == 0.0) ? v : fma(beta, *p, v);
                }
            }
}

// (the eight-wave variant is held to 128 registers: four waves per SIMD, two workgroups per CU)
template <int TA, int TB, int TT = 64, int WGM = 2>
__global__ void __launch_bounds__(128 * WGM, WGM == 4 ? 4 : 1) gemm_kernel(const GemmTile* __restrict__ tiles, const pomx_gemm_job* __restrict__ jobs,
                                                         const double* __restrict__ A, const double* __restrict__ B,
                                                         double* __restrict__ C) {
    GemmTile const T = tiles[blockIdx.x];
    gemm_tile<TA, TB, TT, WGM>(jobs[T.job], T.i0, T.j0, A, B, C);
}

// A job large enough for 128 x 128 tiles: at least 256 x 256 outputs and a K of 64.
bool gemm_job_big(pomx_gemm_job const& j) { return j.m >= 256 && j.n >= 256 && j.k >= 64; }

// Job blockIdx.y, tile blockIdx.x (row-major over the job's TT x TT tiles).  Tiles outside the
// job exit at once, so the launch can be sized for the largest job and the job sizes may be
// written on the device by an earlier kernel.
template <int TA, int TB, bool UPPER, int TT = 64, int WGM = 2>
__global__ void __launch_bounds__(128 * WGM, WGM == 4 ? 4 : 1) gemmj_kernel(const pomx_gemm_job* __restrict__ jobs, const double* __restrict__ A,
                                                          const double* __restrict__ B, double* __restrict__ C) {
    pomx_gemm_job const J = jobs[blockIdx.y];
    int const tn = (J.n + TT - 1) / TT, tm = (J.m + TT - 1) / TT;
    // grid-stride over the job's tiles (uniform per workgroup: the barriers inside stay uniform);
    // UPPER: only tiles that touch the upper triangle (column >= row) of the output
    for(int t = blockIdx.x; t < tm * tn; t += gridDim.x) {
        int const i0 = (t / tn) * TT, j0 = (t % tn) * TT;
        if(UPPER && j0 + TT - 1 < i0)
            continue;
        __syncthreads();
        gemm_tile<TA, TB, TT, WGM>(J, i0, j0, A, B, C);
    }
}

// One job; the tile is (blockIdx.y, blockIdx.x), so nothing needs to be uploaded per launch.
template <int TA, int TB>
__global__ void __launch_bounds__(256) gemm1_kernel(pomx_gemm_job const J, const double* __restrict__ A,
                                                    const double* __restrict__ B, double* __restrict__ C) {
    gemm_tile<TA, TB>(J, int(blockIdx.y) * 64, int(blockIdx.x) * 64, A, B, C);
}

} // namespace

void gemm_tiles(const pomx_gemm_job* jobs, int njobs, std::vector<GemmTile>& tiles, int tt) {
    for(int q = 0; q < njobs; ++q)
        for(int i0 = 0; i0 < jobs[q].m; i0 += tt)
            for(int j0 = 0; j0 < jobs[q].n; j0 += tt)
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

// 128 x 128 tiles of eight waves (tile list built with gemm_tiles(..., 128)).
int launch_gemm_big(pomx_ctx* ctx, int trans_a, int trans_b, const GemmTile* d_tiles, int ntiles,
                    const pomx_gemm_job* d_jobs, const double* d_A, const double* d_B, double* d_C) {
    if(ntiles <= 0)
        return POMX_OK;
    unsigned g = unsigned(ntiles);
    if(!trans_a && !trans_b)
        gemm_kernel<0, 0, 128, 4><<<g, 512, 0, ctx->stream>>>(d_tiles, d_jobs, d_A, d_B, d_C);
    else if(!trans_a && trans_b)
        gemm_kernel<0, 1, 128, 4><<<g, 512, 0, ctx->stream>>>(d_tiles, d_jobs, d_A, d_B, d_C);
    else if(trans_a && !trans_b)
        gemm_kernel<1, 0, 128, 4><<<g, 512, 0, ctx->stream>>>(d_tiles, d_jobs, d_A, d_B, d_C);
    else
        gemm_kernel<1, 1, 128, 4><<<g, 512, 0, ctx->stream>>>(d_tiles, d_jobs, d_A, d_B, d_C);
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
                     const double* d_A, const double* d_B, double* d_C, bool upper, bool big) {
    if(njobs <= 0 || max_tiles <= 0)
        return POMX_OK;
    if(njobs > 65535)
        return pomx_fail(ctx, POMX_ERR_RANGE, "launch_gemm_jobs: more than 65535 jobs in one launch");
    if(big) { // 128 x 128 tiles: a quarter of the 64 x 64 count (rounded up per dimension)
        if(upper || trans_b)
            return pomx_fail(ctx, POMX_ERR_ARG, "launch_gemm_jobs: 128 x 128 tiles only for NN and TN without upper");
        int64_t const gx = std::max<int64_t>(1, std::min<int64_t>((max_tiles + 3) / 4 + 64, (16384 + njobs - 1) / njobs));
        dim3 const g{unsigned(gx), unsigned(njobs), 1u};
        if(trans_a)
            gemmj_kernel<1, 0, false, 128, 4><<<g, 512, 0, ctx->stream>>>(d_jobs, d_A, d_B, d_C);
        else
            gemmj_kernel<0, 0, false, 128, 4><<<g, 512, 0, ctx->stream>>>(d_jobs, d_A, d_B, d_C);
        POMX_CHECK_LAUNCH(ctx);
        return POMX_OK;
    }
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
    // one launch per transposition pattern and tile size (128 x 128 tiles for large jobs)
    for(int ta = 0; ta < 2; ++ta)
        for(int tb = 0; tb < 2; ++tb)
            for(int big = 0; big < 2; ++big) {
                std::vector<pomx_gemm_job> js;
                for(int q = 0; q < njobs; ++q)
                    if((h_jobs[q].trans_a != 0) == (ta != 0) && (h_jobs[q].trans_b != 0) == (tb != 0) &&
                       gemm_job_big(h_jobs[q]) == (big != 0))
                        js.push_back(h_jobs[q]);
                if(js.empty())
                    continue;
                std::vector<GemmTile> tiles;
                gemm_tiles(js.data(), int(js.size()), tiles, big ? 128 : 64);
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
                int rc = big ? launch_gemm_big(ctx, ta, tb, d_tiles, int(tiles.size()), d_jobs, d_A, d_B, d_C)
                             : launch_gemm(ctx, ta, tb, d_tiles, int(tiles.size()), d_jobs, d_A, d_B, d_C);
                if(rc) {
                    (void)hipStreamSynchronize(ctx->stream); // a queued launch may read the pooled descriptors
                    return rc;
                }
                // the pooled descriptors are released below: wait for the launch that reads them
                POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
            }
    return POMX_OK;
}
