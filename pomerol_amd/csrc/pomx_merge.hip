// Fixed point of the TermList erase events shared by the term merges (declared in
// pomx_internal.hpp with the rule it settles).
#include "pomx_internal.hpp"

#include <algorithm>

namespace {

// inclusive prefix sums of the event weights: cum = exclusive + own
__global__ void event_cum_kernel(int64_t n, const uint64_t* __restrict__ w, int64_t* __restrict__ cum) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n)
        cum[k] += int64_t(w[k]);
}

// flag[0] = 1 if the sorted event lists differ
__global__ void event_diff_kernel(int64_t n, const uint64_t* __restrict__ a, const uint64_t* __restrict__ aw,
                                  const uint64_t* __restrict__ b, const uint64_t* __restrict__ bw, int* __restrict__ flag) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n && (a[k] != b[k] || aw[k] != bw[k]))
        flag[0] = 1;
}

} // namespace

int erase_fixed_point(pomx_ctx* ctx, int64_t ng, EraseRule rule, ErasePass const& pass, EraseEvents& ev,
                      const char* who) {
    constexpr int kMaxPasses = 64;
    std::string const oom = std::string(who) + ": out of device memory";
    int rc;
    WsBuf<int64_t> evn(ctx, ng), evoff(ctx, ng);
    WsBuf<int> flag(ctx, 1);
    if(!evn.p || !evoff.p || !flag.p)
        return pomx_fail(ctx, POMX_ERR_HIP, oom);
    ev = EraseEvents{};
    for(int it = 0; it < kMaxPasses; ++it) {
        rule.ev_key = ev.key.p;
        rule.ev_cum = ev.cum.p;
        rule.nev = ev.n;
        if((rc = pass(rule, false, evn, nullptr, nullptr)))
            return rc;
        if((rc = pomx_exclusive_scan(ctx, ng, evn, evoff)))
            return rc;
        int64_t la = 0, lb = 0;
        if((rc = read_i64(ctx, evn.p + ng - 1, &la)) || (rc = read_i64(ctx, evoff.p + ng - 1, &lb)))
            return rc;
        int64_t const ne = la + lb;
        if(ne == 0 && ev.n == 0)
            return POMX_OK; // no erasure anywhere: the first pass is exact
        int64_t const cap = std::max<int64_t>(ne, 1);
        WsBuf<uint64_t> k0(ctx, cap), w0(ctx, cap), k1(ctx, cap), w1(ctx, cap);
        WsBuf<int64_t> cum(ctx, cap);
        if(!k0.p || !w0.p || !k1.p || !w1.p || !cum.p)
            return pomx_fail(ctx, POMX_ERR_HIP, oom);
        if(ne > 0) {
            if((rc = pass(rule, true, evoff, k0, w0)))
                return rc;
            if((rc = pomx_sort_pairs(ctx, ne, k0, k1, w0, w1, 0, 64)))
                return rc;
        }
        if(ne == ev.n) {
            if((rc = pomx_fill_zero(ctx, flag.p, sizeof(int))))
                return rc;
            event_diff_kernel<<<grid_for(ne), kBlock, 0, ctx->stream>>>(ne, k1, w1, ev.key, ev.w, flag);
            POMX_CHECK_LAUNCH(ctx);
            int hf = 1;
            POMX_HIP(ctx, hipMemcpyAsync(&hf, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
            if(!hf)
                return POMX_OK; // this pass reproduced the events it assumed
        }
        if(ne > 0) {
            if((rc = pomx_exclusive_scan(ctx, ne, reinterpret_cast<const int64_t*>(w1.p), cum)))
                return rc;
            event_cum_kernel<<<grid_for(ne), kBlock, 0, ctx->stream>>>(ne, w1, cum);
            POMX_CHECK_LAUNCH(ctx);
        }
        ev.key = std::move(k1);
        ev.w = std::move(w1);
        ev.cum = std::move(cum);
        ev.n = ne;
    }
    return pomx_fail(ctx, POMX_ERR_RANGE, std::string(who) + ": term erasures did not settle");
}
