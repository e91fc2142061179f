// K3: rotation of a monomial operator into the eigenbasis, batched over connected block pairs.
//
// Replaces MonomialOperatorPart::computeImpl (src/pomerol/MonomialOperatorPart.cpp:48-92):
//   OURight = M * U_R   (M has at most one +-1 per column: a signed row gather)
//   O       = U_L^T * OURight,  then sparseView(Tol) (entries |x| <= Tol*1e-12 dropped).
// The gather is fused into the B-operand load of an FP64 MFMA GEMM
// (v_mfma_f64_16x16x4f64): 64x64 output tile per workgroup, four waves of 32x32, K staged
// through LDS 16 rows at a time.
#include "pomx_internal.hpp"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) rotate_kernel(const RotTile* __restrict__ tiles, const pomx_rot_job* __restrict__ jobs,
                                                     const double* __restrict__ U, const int* __restrict__ src,
                                                     double* __restrict__ out, double cutoff) {
    __shared__ double As[16][64 + 4];
    __shared__ double Bs[16][64 + 4];
    RotTile const T = tiles[blockIdx.x];
    pomx_rot_job const J = jobs[T.job];
    int const M = J.n_to, N = J.n_from, K = J.n_to;
    const double* __restrict__ Ut = U + J.u_to_off;
    const double* __restrict__ Uf = U + J.u_from_off;
    const int* __restrict__ sr = src + J.src_off;
    int const tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    int const wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    dbl4 acc[2][2];
    for(int a = 0; a < 2; ++a)
        for(int b = 0; b < 2; ++b)
            acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
    for(int t0 = 0; t0 < K; t0 += 16) {
        for(int e = tid; e < 16 * 64; e += 256) {
            int kk = e >> 6, cc = e & 63;
            int t = t0 + kk;
            double a = 0.0, b = 0.0;
            if(t < K) {
                int i = T.i0 + cc;
                if(i < M)
                    a = Ut[int64_t(t) * M + i];
                int sv = sr[t];
                int j = T.j0 + cc;
                if(sv != 0 && j < N) {
                    int s = (sv > 0 ? sv : -sv) - 1;
                    double x = Uf[int64_t(s) * N + j];
                    b = sv > 0 ? x : -x;
                }
            }
            As[kk][cc] = a;
            Bs[kk][cc] = b;
        }
        __syncthreads();
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
        __syncthreads();
    }
    double* __restrict__ O = out + J.out_off;
#pragma unroll
    for(int mt = 0; mt < 2; ++mt)
#pragma unroll
        for(int nt2 = 0; nt2 < 2; ++nt2)
#pragma unroll
            for(int r = 0; r < 4; ++r) {
                // v_mfma_f64_16x16x4f64 result layout: D[i][j] in lane j + 16*(i%4), register i/4
                int i = T.i0 + wm + mt * 16 + (lane >> 4) + 4 * r;
                int j = T.j0 + wn + nt2 * 16 + (lane & 15);
                if(i < M && j < N) {
                    double x = acc[mt][nt2][r];
                    O[int64_t(i) * N + j] = fabs(x) <= cutoff ? 0.0 : x;
                }
            }
}

} // namespace

int launch_rotate(pomx_ctx* ctx, const RotTile* d_tiles, int ntiles, const pomx_rot_job* d_jobs, const double* d_U,
                  const int* d_src, double* d_out, double cutoff) {
    if(ntiles <= 0)
        return POMX_OK;
    rotate_kernel<<<unsigned(ntiles), 256, 0, ctx->stream>>>(d_tiles, d_jobs, d_U, d_src, d_out, cutoff);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

extern "C" int pomx_rotate_batched(pomx_ctx* ctx, int njobs, const pomx_rot_job* h_jobs, const double* d_U,
                                   const int* d_src, double* d_out, double cutoff) {
    POMX_REQUIRE(ctx, njobs >= 0 && (njobs == 0 || (h_jobs && d_U && d_src && d_out)), "pomx_rotate_batched: bad arguments");
    if(njobs == 0)
        return POMX_OK;
    std::vector<RotTile> tiles;
    for(int q = 0; q < njobs; ++q) {
        POMX_REQUIRE(ctx, h_jobs[q].n_to >= 1 && h_jobs[q].n_from >= 1, "pomx_rotate_batched: empty block");
        for(int i0 = 0; i0 < h_jobs[q].n_to; i0 += 64)
            for(int j0 = 0; j0 < h_jobs[q].n_from; j0 += 64)
                tiles.push_back(RotTile{q, i0, j0, 0});
    }
    WsBuf<RotTile> d_tiles(ctx, tiles.size());
    WsBuf<pomx_rot_job> d_jobs(ctx, njobs);
    if(!d_tiles.p || !d_jobs.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_rotate_batched: out of device memory");
    POMX_HIP(ctx, hipMemcpyAsync(d_tiles.p, tiles.data(), sizeof(RotTile) * tiles.size(), hipMemcpyHostToDevice, ctx->stream));
    POMX_HIP(ctx, h2d_sync(ctx, d_jobs.p, h_jobs, sizeof(pomx_rot_job) * njobs)); // (both copies done)
    int rc = launch_rotate(ctx, d_tiles, int(tiles.size()), d_jobs, d_U, d_src, d_out, cutoff);
    if(rc)
        return rc;
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}
