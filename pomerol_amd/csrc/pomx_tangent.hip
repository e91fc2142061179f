// First-order response of solved blocks and of the Boltzmann weights along a Hamiltonian
// direction dH (H(t) = H + t dH at t = 0).  No counterpart in the reference: the forward-mode
// half of the differentiation of the correlators.
//
//   H U = U diag(E)  =>  dH U + H dU = dU diag(E) + U diag(dE); with dU = U K:
//   A = U^T dH U,  dE_i = A_ii,  K_ij (E_j - E_i) = A_ij  (i != j),  K_ii = 0,
// so K is antisymmetric and a rotated operator O = U_L^T M U_R moves as dO = O K_R - K_L O.
// Pairs inside a degenerate eigenspace get K_ij = 0: the correlators are sums over complete
// degenerate eigenspaces, so a rotation inside one changes nothing, and a direction that keeps the
// symmetry behind the degeneracy has A_ij = 0 there.
//   w_s = exp(-beta E_s) / Z  =>  dw_s = -beta w_s (dE_s - sum_t w_t dE_t).
#include "pomx_internal.hpp"

#include <algorithm>

namespace {

// K in place of A and dE from its diagonal, one workgroup per block.
__global__ void tangent_k_kernel(const EighJob* __restrict__ blocks, const double* __restrict__ E, double* __restrict__ A,
                                 double* __restrict__ dE) {
    EighJob const B = blocks[blockIdx.x];
    int const n = B.n;
    const double* e = E + B.val_off;
    double* a = A + B.mat_off;
    for(int64_t q = threadIdx.x; q < int64_t(n) * n; q += blockDim.x) {
        int const i = int(q / n), j = int(q % n);
        double v = 0.0;
        if(i == j)
            dE[B.val_off + i] = a[q];
        else {
            double const d = e[j] - e[i];
            double const scale = fmax(1.0, fmax(fabs(e[i]), fabs(e[j])));
            v = (fabs(d) <= 1e-10 * scale) ? 0.0 : a[q] / d;
        }
        a[q] = v;
    }
}

// partial[b] = sum of w dE over the grid-stride share of block b (fixed order)
__global__ void wde_partial_kernel(int64_t n, const double* __restrict__ w, const double* __restrict__ dE,
                                   double* __restrict__ partial) {
    __shared__ double s[kBlock];
    double acc = 0.0;
    for(int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x)
        acc = fma(w[k], dE[k], acc);
    s[threadIdx.x] = acc;
    __syncthreads();
    for(int o = kBlock / 2; o > 0; o >>= 1) {
        if(int(threadIdx.x) < o)
            s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if(threadIdx.x == 0)
        partial[blockIdx.x] = s[0];
}

__global__ void wde_final_kernel(int np, double* __restrict__ partial) {
    __shared__ double s[kBlock];
    double acc = 0.0;
    for(int k = threadIdx.x; k < np; k += blockDim.x)
        acc += partial[k];
    s[threadIdx.x] = acc;
    __syncthreads();
    for(int o = kBlock / 2; o > 0; o >>= 1) {
        if(int(threadIdx.x) < o)
            s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if(threadIdx.x == 0)
        partial[0] = s[0];
}

__global__ void dw_kernel(int64_t n, const double* __restrict__ w, const double* __restrict__ dE,
                          const double* __restrict__ mean, double beta, double* __restrict__ dw) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n)
        dw[k] = -beta * w[k] * (dE[k] - mean[0]);
}

} // namespace

extern "C" int pomx_eigh_tangent(pomx_ctx* ctx, int nb, const int* h_n, const int64_t* h_mat_off,
                                 const int64_t* h_val_off, const double* d_U, const double* d_E, const double* d_dH,
                                 double* d_K, double* d_dE) {
    POMX_REQUIRE(ctx, nb >= 0 && (nb == 0 || (h_n && h_mat_off && h_val_off && d_U && d_E && d_dH && d_K && d_dE)),
                 "pomx_eigh_tangent: bad arguments");
    if(nb == 0)
        return POMX_OK;
    int64_t arena = 0;
    std::vector<EighJob> blocks(static_cast<size_t>(nb));
    std::vector<pomx_gemm_job> mul, rot;
    for(int b = 0; b < nb; ++b) {
        int const n = h_n[b];
        POMX_REQUIRE(ctx, n >= 0 && h_mat_off[b] >= 0 && h_val_off[b] >= 0, "pomx_eigh_tangent: bad block");
        blocks[b] = EighJob{n, 0, h_mat_off[b], h_val_off[b], 0};
        arena = std::max(arena, h_mat_off[b] + int64_t(n) * n);
        if(n == 0)
            continue;
        // T = dH U, then A = U^T T into d_K
        mul.push_back(pomx_gemm_job{n, n, n, 0, 0, n, n, n, h_mat_off[b], h_mat_off[b], h_mat_off[b], 1.0, 0.0});
        rot.push_back(pomx_gemm_job{n, n, n, 1, 0, n, n, n, h_mat_off[b], h_mat_off[b], h_mat_off[b], 1.0, 0.0});
    }
    WsBuf<double> T(ctx, size_t(std::max<int64_t>(arena, 1)));
    WsBuf<EighJob> d_blocks(ctx, blocks.size());
    if(!T.p || !d_blocks.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_eigh_tangent: out of device memory");
    for(int pass = 0; pass < 2; ++pass) {
        auto const& js = pass == 0 ? mul : rot;
        if(js.empty())
            continue;
        std::vector<GemmTile> tiles;
        gemm_tiles(js.data(), int(js.size()), tiles);
        WsBuf<GemmTile> d_tiles(ctx, tiles.size());
        WsBuf<pomx_gemm_job> d_jobs(ctx, js.size());
        if(!d_tiles.p || !d_jobs.p)
            return pomx_fail(ctx, POMX_ERR_HIP, "pomx_eigh_tangent: out of device memory");
        POMX_HIP(ctx, hipMemcpyAsync(d_tiles.p, tiles.data(), sizeof(GemmTile) * tiles.size(), hipMemcpyHostToDevice,
                                     ctx->stream));
        POMX_HIP(ctx, h2d_sync(ctx, d_jobs.p, js.data(), sizeof(pomx_gemm_job) * js.size())); // (both copies done)
        int rc = pass == 0 ? launch_gemm(ctx, 0, 0, d_tiles, int(tiles.size()), d_jobs, d_dH, d_U, T)
                           : launch_gemm(ctx, 1, 0, d_tiles, int(tiles.size()), d_jobs, d_U, T, d_K);
        if(rc)
            return rc;
        // the pooled descriptors are released at the end of this scope: wait for their reader
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    POMX_HIP(ctx, h2d_sync(ctx, d_blocks.p, blocks.data(), sizeof(EighJob) * blocks.size()));
    tangent_k_kernel<<<nb, kBlock, 0, ctx->stream>>>(d_blocks, d_E, d_K, d_dE);
    POMX_CHECK_LAUNCH(ctx);
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

extern "C" int pomx_boltzmann_tangent(pomx_ctx* ctx, int64_t n, const double* d_w, const double* d_dE, double beta,
                                      double* d_dw) {
    POMX_REQUIRE(ctx, n > 0 && d_w && d_dE && d_dw, "pomx_boltzmann_tangent: bad arguments");
    int np = int(std::min<int64_t>(grid_for(n), 1024));
    WsBuf<double> part(ctx, np);
    if(!part.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_boltzmann_tangent: out of device memory");
    wde_partial_kernel<<<np, kBlock, 0, ctx->stream>>>(n, d_w, d_dE, part);
    wde_final_kernel<<<1, kBlock, 0, ctx->stream>>>(np, part);
    dw_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, d_w, d_dE, part, beta, d_dw);
    POMX_CHECK_LAUNCH(ctx);
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}
