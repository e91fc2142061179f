// Context, device memory, block assembly (K1), Boltzmann weights (K5) and sort/scan utilities.
#include "pomx_internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cmath>
#include <cstring>

int pomx_fail(pomx_ctx* ctx, int code, std::string const& msg) {
    if(ctx)
        ctx->err = msg;
    return code;
}

// Return every cached workspace block to the device (after the work that used them).
static void ws_release_cached(pomx_ctx* ctx) {
    hipStreamSynchronize(ctx->stream);
    for(auto& kv : ctx->free_blocks)
        hipFree(kv.second);
    ctx->free_blocks.clear();
}

hipError_t device_malloc(pomx_ctx* ctx, void** p, size_t bytes) {
    if(hipMalloc(p, bytes) == hipSuccess)
        return hipSuccess;
    (void)hipGetLastError();
    ws_release_cached(ctx);
    hipError_t e = hipMalloc(p, bytes);
    if(e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
    }
    return e;
}

void* ws_get(pomx_ctx* ctx, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    auto it = ctx->free_blocks.lower_bound(bytes);
    if(it != ctx->free_blocks.end() && it->first <= 2 * bytes + (size_t(64) << 20)) {
        void* p = it->second;
        size_t sz = it->first;
        ctx->free_blocks.erase(it);
        ctx->live_blocks[p] = sz;
        return p;
    }
    void* p = nullptr;
    if(device_malloc(ctx, &p, bytes) != hipSuccess)
        return nullptr;
    ctx->live_blocks[p] = bytes;
    return p;
}

void ws_put(pomx_ctx* ctx, void* p) {
    auto it = ctx->live_blocks.find(p);
    if(it == ctx->live_blocks.end())
        return;
    ctx->free_blocks.emplace(it->second, p);
    ctx->live_blocks.erase(it);
}

extern "C" {

int pomx_ctx_create(int device, pomx_ctx** out) {
    if(!out)
        return POMX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
        return POMX_ERR_NODEV;
    hipDeviceProp_t prop;
    if(hipGetDeviceProperties(&prop, device) != hipSuccess)
        return POMX_ERR_NODEV;
    if(std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return POMX_ERR_NODEV;
    auto* ctx = new pomx_ctx();
    ctx->device = device;
    if(hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
       hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return POMX_ERR_HIP;
    }
    if(pair_scratch_setup(ctx) != POMX_OK) { // the no-allocation entry points' scratch on this device
        pomx_ctx_destroy(ctx);
        return POMX_ERR_HIP;
    }
    *out = ctx;
    return POMX_OK;
}

void pomx_ctx_destroy(pomx_ctx* ctx) {
    if(!ctx)
        return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for(auto& kv : ctx->free_blocks)
        hipFree(kv.second);
    for(auto& kv : ctx->live_blocks)
        hipFree(kv.first);
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    for(auto const& r : ctx->stat_recs) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    for(hipEvent_t e : ctx->stat_pool)
        hipEventDestroy(e);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* pomx_last_error(const pomx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pomx_malloc(pomx_ctx* ctx, size_t bytes, void** d_ptr) {
    POMX_REQUIRE(ctx, d_ptr, "pomx_malloc: null output pointer");
    *d_ptr = nullptr;
    if(bytes == 0)
        bytes = 8;
    hipError_t e = device_malloc(ctx, d_ptr, bytes);
    if(e != hipSuccess)
        return pomx_fail(ctx, POMX_ERR_HIP, std::string("pomx_malloc: ") + hipGetErrorString(e));
    return POMX_OK;
}

int pomx_free(pomx_ctx* ctx, void* d_ptr) {
    if(!d_ptr)
        return POMX_OK;
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    POMX_HIP(ctx, hipFree(d_ptr));
    return POMX_OK;
}

int pomx_h2d(pomx_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
    if(!bytes)
        return POMX_OK;
    POMX_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

int pomx_d2h(pomx_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
    if(!bytes)
        return POMX_OK;
    POMX_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

} // extern "C"

__global__ void fill_zero_kernel(uint64_t* __restrict__ p, int64_t n) {
    for(int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = 0;
}
__global__ void fill_zero_bytes_kernel(unsigned char* __restrict__ p, int64_t n) {
    for(int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = 0;
}

extern "C" int pomx_fill_zero(pomx_ctx* ctx, void* d_dst, size_t bytes) {
    if(!bytes)
        return POMX_OK;
    // a fill kernel rather than hipMemsetAsync so the call is safe inside captured graphs
    if((reinterpret_cast<uintptr_t>(d_dst) % 8) == 0 && bytes % 8 == 0) {
        int64_t n = int64_t(bytes / 8);
        fill_zero_kernel<<<std::min<unsigned>(grid_for(n), 4096), kBlock, 0, ctx->stream>>>(static_cast<uint64_t*>(d_dst), n);
    } else {
        int64_t n = int64_t(bytes);
        fill_zero_bytes_kernel<<<std::min<unsigned>(grid_for(n), 4096), kBlock, 0, ctx->stream>>>(
            static_cast<unsigned char*>(d_dst), n);
    }
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

extern "C" int pomx_synchronize(pomx_ctx* ctx) {
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

// ---- K1 ------------------------------------------------------------------------------------
__global__ void scatter_kernel(int64_t nnz, const int64_t* __restrict__ pos, const double* __restrict__ val,
                               double* __restrict__ mat, int64_t mat_len) {
    int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(e < nnz) {
        int64_t p = pos[e];
        if(p >= 0 && p < mat_len)
            mat[p] = val[e]; // positions are unique (combined on the host)
    }
}

extern "C" int pomx_assemble(pomx_ctx* ctx, int64_t nnz, const int64_t* d_pos, const double* d_val, double* d_mat,
                             int64_t mat_len) {
    POMX_REQUIRE(ctx, nnz >= 0 && mat_len >= 0, "pomx_assemble: negative size");
    int rc = pomx_fill_zero(ctx, d_mat, size_t(mat_len) * sizeof(double));
    if(rc)
        return rc;
    if(nnz > 0) {
        scatter_kernel<<<grid_for(nnz), kBlock, 0, ctx->stream>>>(nnz, d_pos, d_val, d_mat, mat_len);
        POMX_CHECK_LAUNCH(ctx);
    }
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

// ---- K5 ------------------------------------------------------------------------------------
// Deterministic two-level reductions: per-workgroup partials, then one workgroup.
__device__ inline double block_reduce_min(double v, double* sh) {
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for(int s = blockDim.x / 2; s > 0; s >>= 1) {
        if(t < s)
            sh[t] = fmin(sh[t], sh[t + s]);
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}
__device__ inline double block_reduce_sum(double v, double* sh) {
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for(int s = blockDim.x / 2; s > 0; s >>= 1) {
        if(t < s)
            sh[t] += sh[t + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__global__ void min_partial_kernel(int64_t n, const double* __restrict__ E, double* __restrict__ part) {
    __shared__ double sh[kBlock];
    double m = INFINITY;
    for(int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        m = fmin(m, E[i]);
    m = block_reduce_min(m, sh);
    if(threadIdx.x == 0)
        part[blockIdx.x] = m;
}
__global__ void min_final_kernel(int np, double* __restrict__ part) {
    __shared__ double sh[kBlock];
    double m = INFINITY;
    for(int i = threadIdx.x; i < np; i += blockDim.x)
        m = fmin(m, part[i]);
    m = block_reduce_min(m, sh);
    if(threadIdx.x == 0)
        part[0] = m;
}
__global__ void exp_partial_kernel(int64_t n, const double* __restrict__ E, const double* __restrict__ E0p, double beta,
                                   double* __restrict__ w, double* __restrict__ part) {
    __shared__ double sh[kBlock];
    double E0 = E0p[0];
    double s = 0;
    for(int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        double x = exp(-beta * (E[i] - E0));
        w[i] = x;
        s += x;
    }
    s = block_reduce_sum(s, sh);
    if(threadIdx.x == 0)
        part[blockIdx.x] = s;
}
__global__ void sum_final_kernel(int np, double* __restrict__ part) {
    __shared__ double sh[kBlock];
    double s = 0;
    for(int i = threadIdx.x; i < np; i += blockDim.x)
        s += part[i];
    s = block_reduce_sum(s, sh);
    if(threadIdx.x == 0)
        part[0] = s;
}
__global__ void scale_kernel(int64_t n, double* __restrict__ w, const double* __restrict__ Zp) {
    double Z = Zp[0];
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(i < n)
        w[i] /= Z;
}

extern "C" int pomx_boltzmann(pomx_ctx* ctx, int64_t n, const double* d_E, double beta, double* d_w, double* h_E0,
                              double* h_Z) {
    POMX_REQUIRE(ctx, n > 0 && d_E && d_w, "pomx_boltzmann: bad arguments");
    int np = int(std::min<int64_t>(grid_for(n), 1024));
    WsBuf<double> p1(ctx, np), p2(ctx, np);
    min_partial_kernel<<<np, kBlock, 0, ctx->stream>>>(n, d_E, p1);
    min_final_kernel<<<1, kBlock, 0, ctx->stream>>>(np, p1);
    exp_partial_kernel<<<np, kBlock, 0, ctx->stream>>>(n, d_E, p1, beta, d_w, p2);
    sum_final_kernel<<<1, kBlock, 0, ctx->stream>>>(np, p2);
    scale_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, d_w, p2);
    POMX_CHECK_LAUNCH(ctx);
    double E0 = 0, Z = 0;
    POMX_HIP(ctx, hipMemcpyAsync(&E0, p1.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipMemcpyAsync(&Z, p2.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if(h_E0)
        *h_E0 = E0;
    if(h_Z)
        *h_Z = Z;
    return POMX_OK;
}

// ---- sort / scan utilities -------------------------------------------------------------------
int pomx_sort_pairs(pomx_ctx* ctx, int64_t n, uint64_t* keys_in, uint64_t* keys_out, uint64_t* vals_in,
                    uint64_t* vals_out, int begin_bit, int end_bit) {
    if(n <= 0)
        return POMX_OK;
    size_t tmp = 0;
    POMX_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tmp, keys_in, keys_out, vals_in, vals_out, size_t(n), begin_bit,
                                            end_bit, ctx->stream));
    WsBuf<unsigned char> t(ctx, tmp);
    if(!t.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "sort: out of device memory");
    POMX_HIP(ctx, rocprim::radix_sort_pairs(t.p, tmp, keys_in, keys_out, vals_in, vals_out, size_t(n), begin_bit,
                                            end_bit, ctx->stream));
    return POMX_OK;
}

int pomx_exclusive_scan(pomx_ctx* ctx, int64_t n, const int64_t* in, int64_t* out) {
    if(n <= 0)
        return POMX_OK;
    size_t tmp = 0;
    POMX_HIP(ctx, rocprim::exclusive_scan(nullptr, tmp, in, out, int64_t(0), size_t(n), rocprim::plus<int64_t>(),
                                          ctx->stream));
    WsBuf<unsigned char> t(ctx, tmp);
    if(!t.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "scan: out of device memory");
    POMX_HIP(ctx, rocprim::exclusive_scan(t.p, tmp, in, out, int64_t(0), size_t(n), rocprim::plus<int64_t>(),
                                          ctx->stream));
    return POMX_OK;
}

namespace {

// Fold the pending event pairs into the per-class totals (waits for the last one).
void stat_resolve(pomx_ctx* ctx) {
    if(ctx->stat_recs.empty())
        return;
    hipEventSynchronize(ctx->stat_recs.back().b);
    for(auto const& r : ctx->stat_recs) {
        float ms = 0;
        hipEventElapsedTime(&ms, r.a, r.b);
        pomx_kernel_stat& s = ctx->stats[r.cls];
        s.launches += 1;
        s.ms += ms;
        s.flop += r.flop;
        s.term_freqs += r.term_freqs;
        s.bytes += r.bytes;
        ctx->stat_pool.push_back(r.a);
        ctx->stat_pool.push_back(r.b);
    }
    ctx->stat_recs.clear();
}

hipEvent_t stat_event(pomx_ctx* ctx) {
    if(ctx->stat_pool.empty()) {
        hipEvent_t e = nullptr;
        if(hipEventCreate(&e) != hipSuccess)
            return nullptr;
        return e;
    }
    hipEvent_t e = ctx->stat_pool.back();
    ctx->stat_pool.pop_back();
    return e;
}

} // namespace

int stat_begin(pomx_ctx* ctx, int cls) {
    if(!ctx->timing || cls < 0 || cls >= POMX_STAT_NCLASS)
        return -1;
    if(ctx->stat_recs.size() >= 4096)
        stat_resolve(ctx);
    hipEvent_t a = stat_event(ctx), b = stat_event(ctx);
    if(!a || !b)
        return -1;
    hipEventRecord(a, ctx->stream);
    ctx->stat_recs.push_back(pomx_ctx::StatRec{cls, a, b, 0.0, 0.0, 0.0});
    return int(ctx->stat_recs.size()) - 1;
}

void stat_end(pomx_ctx* ctx, int handle, double flop, double term_freqs, double bytes) {
    if(handle < 0 || handle >= int(ctx->stat_recs.size()))
        return;
    auto& r = ctx->stat_recs[std::size_t(handle)];
    hipEventRecord(r.b, ctx->stream);
    r.flop = flop;
    r.term_freqs = term_freqs;
    r.bytes = bytes;
}

extern "C" int pomx_eval_timing(pomx_ctx* ctx, int enable, double* h_total_ms, int64_t* h_launches) {
    POMX_REQUIRE(ctx, ctx, "null context");
    stat_resolve(ctx);
    double ms = 0;
    int64_t n = 0;
    for(int c = 0; c < POMX_STAT_NCLASS; ++c)
        if(c != POMX_STAT_G2_GEN && c != POMX_STAT_G2_SORT && c != POMX_STAT_G2_REDUCE) {
            ms += ctx->stats[c].ms;
            n += ctx->stats[c].launches;
        }
    ms += ctx->stats[POMX_STAT_G2_REDUCE].ms;
    if(h_total_ms)
        *h_total_ms = ms;
    if(h_launches)
        *h_launches = n;
    if(enable >= 0) {
        ctx->timing = enable != 0;
        for(auto& st : ctx->stats)
            st = pomx_kernel_stat{};
    }
    return POMX_OK;
}

extern "C" int pomx_kernel_stats(pomx_ctx* ctx, pomx_kernel_stat* h_stats) {
    POMX_REQUIRE(ctx, ctx && h_stats, "pomx_kernel_stats: bad arguments");
    stat_resolve(ctx);
    for(int c = 0; c < POMX_STAT_NCLASS; ++c)
        h_stats[c] = ctx->stats[c];
    return POMX_OK;
}
