// Batched full-ED + G(i w_n) solves: many independent Hamiltonians in one captured hipGraph.
//
// Graph nodes (all on the context stream, device-resident descriptors, no host round trips):
//   1. zero the matrix arena, scatter the Hamiltonian entries of every instance     (K1)
//   2. eigensolve: for each block-size bucket, persistent workgroups pull blocks from a
//      largest-first queue (atomic counter, reset by a fill kernel inside the graph) (K2)
//   3. Boltzmann weights and ground energy per instance (one workgroup per instance) (K5)
//   4. rotation of c^+ for every connected block pair, FP64 MFMA tiles               (K3)
//   5. fused G: work items (block pair, term range) generate R = c^2 (w_out + w_in),
//      P = E_in - E_out on the fly (|R| > coeff_tol, GreensFunctionPart.cpp:84-94) and
//      accumulate R / (i w_n - P) for the instance's frequencies into per-item partials (K6+K9)
//   6. per-instance reduction of the partials in item order (deterministic)
// The autotuner picks the workgroup size of each eigensolver bucket and the G work-item size.
#include "pomx_batch.hpp"

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <numeric>

namespace {

__global__ void scatter_h_kernel(int64_t nnz, const int64_t* __restrict__ pos, const double* __restrict__ val,
                                 double* __restrict__ mat) {
    int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(e < nnz)
        mat[pos[e]] = val[e];
}

__global__ void zero_kernel(double* __restrict__ p, int64_t n) {
    for(int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = 0.0;
}

__global__ void zero_int_kernel(int* __restrict__ p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
        p[i] = 0;
}

// DensityMatrixPart.cpp:23-31 per instance (one workgroup per instance)
__global__ void __launch_bounds__(256) inst_weights_kernel(const int64_t* __restrict__ off, const int64_t* __restrict__ cnt,
                                                           const double* __restrict__ beta, const double* __restrict__ E,
                                                           double* __restrict__ w, double* __restrict__ E0) {
    __shared__ double red[256];
    int const k = blockIdx.x;
    int64_t const o = off[k], n = cnt[k];
    double m = INFINITY;
    for(int64_t i = threadIdx.x; i < n; i += blockDim.x)
        m = fmin(m, E[o + i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for(int s = blockDim.x / 2; s > 0; s >>= 1) {
        if(threadIdx.x < s)
            red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    double const e0 = red[0];
    __syncthreads();
    double z = 0.0;
    for(int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        double x = exp(-beta[k] * (E[o + i] - e0));
        w[o + i] = x;
        z += x;
    }
    red[threadIdx.x] = z;
    __syncthreads();
    for(int s = blockDim.x / 2; s > 0; s >>= 1) {
        if(threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    double const Z = red[0];
    for(int64_t i = threadIdx.x; i < n; i += blockDim.x)
        w[o + i] /= Z;
    if(threadIdx.x == 0)
        E0[k] = e0;
}

__device__ inline double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Fused term generation + Lehmann sum.  256 threads = 2 term lanes x 128 frequencies; the
// terms of a 256-term tile are generated once into LDS and read with uniform addresses.
__global__ void __launch_bounds__(kGThreads) fused_g_kernel(const GItem* __restrict__ items, const JobInfo* __restrict__ jobs,
                                                            const double* __restrict__ ops, const double* __restrict__ E,
                                                            const double* __restrict__ w, const double* __restrict__ beta,
                                                            int nfreq, double tol, double* __restrict__ partial) {
    __shared__ double sR[kGThreads], sP[kGThreads];
    __shared__ double red[2][kGThreads];
    GItem const it = items[blockIdx.x];
    JobInfo const J = jobs[it.job];
    int const tid = threadIdx.x;
    int const half = tid / 128, f = tid % 128;
    double const om = (f < nfreq) ? M_PI * (2 * f + 1) / beta[it.inst] : 1.0;
    double const om2 = om * om;
    double ar = 0.0, ai = 0.0;
    for(int64_t c0 = it.t0; c0 < it.t1; c0 += kGThreads) {
        int const cnt = int(min<int64_t>(kGThreads, it.t1 - c0));
        __syncthreads();
        if(tid < cnt) {
            int64_t t = c0 + tid;
            int i_out = int(t / J.n_in), i_in = int(t % J.n_in);
            // <out|c|in><in|c^+|out> = m^2 with m = <in|c^+|out>, the c^+ block is n_in x n_out
            double m = ops[J.ops_off + int64_t(i_in) * J.n_out + i_out];
            double R = m * m * (w[J.out_state + i_out] + w[J.in_state + i_in]);
            bool keep = (m != 0.0) && fabs(R) > tol;
            sR[tid] = keep ? R : 0.0;
            sP[tid] = E[J.in_state + i_in] - E[J.out_state + i_out];
        }
        __syncthreads();
        for(int k = half; k < cnt; k += 2) {
            double const R = sR[k], P = sP[k];
            double const s = R * rcp_nr(fma(P, P, om2)); // R/(i w - P) = -R (P + i w)/(P^2 + w^2)
            ar = fma(-s, P, ar);
            ai = fma(-s, om, ai);
        }
    }
    red[0][tid] = ar;
    red[1][tid] = ai;
    __syncthreads();
    if(tid < 128 && tid < nfreq) {
        int64_t o = (int64_t(blockIdx.x) * nfreq + tid) * 2;
        partial[o] = red[0][tid] + red[0][tid + 128];
        partial[o + 1] = red[1][tid] + red[1][tid + 128];
    }
}

__global__ void g_reduce_kernel(int ninst, int nfreq, const int64_t* __restrict__ item_off, const double* __restrict__ partial,
                                double* __restrict__ G) {
    int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(g >= int64_t(ninst) * nfreq)
        return;
    int k = int(g / nfreq), f = int(g % nfreq);
    double sr = 0.0, si = 0.0;
    for(int64_t i = item_off[k]; i < item_off[k + 1]; ++i) {
        sr += partial[(i * nfreq + f) * 2];
        si += partial[(i * nfreq + f) * 2 + 1];
    }
    G[g * 2] = sr;
    G[g * 2 + 1] = si;
}

template <class T> T* dev_copy(pomx_ctx* ctx, const T* h, size_t n, int& rc) {
    void* p = nullptr;
    if(device_malloc(ctx, &p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) {
        rc = pomx_fail(ctx, POMX_ERR_HIP, "pomx_batch_create: out of device memory");
        return nullptr;
    }
    if(n && h && hipMemcpyAsync(p, h, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = pomx_fail(ctx, POMX_ERR_HIP, "pomx_batch_create: copy failed");
    return static_cast<T*>(p);
}

} // namespace

namespace {

void free_batch(pomx_batch* b) {
    void* ptrs[] = {b->mat,  b->eval,     b->w,        b->ws,     b->h_val,     b->beta,   b->ops,   b->partial,
                    b->G,    b->E0,       b->h_pos,    b->inst_off, b->inst_cnt, b->item_off, b->src, b->counters,
                    b->eigh_jobs, b->jobs, b->tiles, b->jinfo, b->items, b->kmap, b->kcnt}; // status lives inside counters
    for(void* p : ptrs)
        if(p)
            (void)hipFree(p);
}

int build_items(pomx_ctx* ctx, pomx_batch* b) {
    std::vector<GItem> items;
    std::vector<int64_t> item_off(size_t(b->ninst) + 1, 0);
    for(int k = 0; k < b->ninst; ++k) {
        item_off[k] = int64_t(items.size());
        for(int64_t j = b->inst_first_job[k]; j < b->inst_first_job[k + 1]; ++j) {
            JobInfo const& J = b->h_jinfo[size_t(j)];
            int64_t nt = int64_t(J.n_out) * J.n_in;
            for(int64_t t0 = 0; t0 < nt; t0 += b->terms_per_item)
                items.push_back(GItem{int(j), k, t0, std::min(nt, t0 + b->terms_per_item)});
        }
    }
    item_off[b->ninst] = int64_t(items.size());
    if(b->items)
        (void)hipFree(b->items);
    if(b->partial)
        (void)hipFree(b->partial);
    if(b->item_off)
        (void)hipFree(b->item_off);
    int rc = POMX_OK;
    b->nitems = int(items.size());
    b->items = dev_copy(ctx, items.data(), items.size(), rc);
    b->item_off = dev_copy(ctx, item_off.data(), item_off.size(), rc);
    b->partial = dev_copy<double>(ctx, nullptr, size_t(std::max(b->nitems, 1)) * b->nfreq * 2, rc);
    if(rc)
        return rc;
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

// The whole solve as a sequence of launches on the context stream (capturable).
int enqueue_solve(pomx_ctx* ctx, pomx_batch* b) {
    hipStream_t s = ctx->stream;
    zero_kernel<<<unsigned(std::min<int64_t>((b->mat_len + 255) / 256, 65536)), 256, 0, s>>>(b->mat, b->mat_len);
    if(b->nnz)
        scatter_h_kernel<<<grid_for(b->nnz), kBlock, 0, s>>>(b->nnz, b->h_pos, b->h_val, b->mat);
    zero_int_kernel<<<grid_for(kBuckets + b->nblocks), kBlock, 0, s>>>(b->counters, kBuckets + b->nblocks);
    for(int q = 0; q < kBuckets; ++q) {
        int nj = b->bucket_begin[q + 1] - b->bucket_begin[q];
        if(nj <= 0)
            continue;
        int grid = std::min(nj, b->eigh_grid);
        int rc = launch_eigh(ctx, b->eigh_jobs + b->bucket_begin[q], nj, b->bucket_nmax[q], b->eigh_threads[q], grid,
                             b->counters + q, b->mat, b->eval, b->ws, b->status + b->bucket_begin[q]);
        if(rc)
            return rc;
    }
    if(b->big) {
        int rc = big_plan_enqueue(ctx, b->big, b->mat, b->eval);
        if(rc)
            return rc;
    }
    inst_weights_kernel<<<b->ninst, 256, 0, s>>>(b->inst_off, b->inst_cnt, b->beta, b->eval, b->w, b->E0);
    int rc = launch_rotate(ctx, b->tiles, b->ntiles, b->jobs, b->mat, b->kmap, b->kcnt, b->ops, 1e-20);
    if(rc)
        return rc;
    if(b->nitems)
        fused_g_kernel<<<b->nitems, kGThreads, 0, s>>>(b->items, b->jinfo, b->ops, b->eval, b->w, b->beta, b->nfreq, b->tol,
                                                       b->partial);
    g_reduce_kernel<<<grid_for(int64_t(b->ninst) * b->nfreq), kBlock, 0, s>>>(b->ninst, b->nfreq, b->item_off, b->partial,
                                                                                b->G);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

void drop_graph(pomx_batch* b) {
    if(b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if(b->graph)
        (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

} // namespace

extern "C" int pomx_batch_create(pomx_ctx* ctx, const pomx_batch_desc* d, pomx_batch** out) {
    POMX_REQUIRE(ctx, d && out && d->ninst > 0 && d->nfreq > 0 && d->nfreq <= 128 && d->nblocks > 0,
                 "pomx_batch_create: bad descriptor (1 <= nfreq <= 128)");
    *out = nullptr;
    for(int q = 0; q < d->njobs; ++q)
        POMX_REQUIRE(ctx, d->jobs[q].n_to >= 1 && d->jobs[q].n_from >= 1 && d->jobs[q].src_off >= 0 &&
                              d->jobs[q].src_off + d->jobs[q].n_to <= d->src_len,
                     "pomx_batch_create: job " + std::to_string(q) + " is empty or its gather map exceeds src_len");
    auto* b = new pomx_batch();
    b->ninst = d->ninst;
    b->nfreq = d->nfreq;
    b->nblocks = d->nblocks;
    b->njobs = d->njobs;
    b->mat_len = d->mat_len;
    b->nstates = d->nstates;
    b->nnz = d->nnz;
    b->tol = d->coeff_tol;
    int rc = POMX_OK;
    // eigensolver jobs: sorted by size (largest first), split into size buckets
    std::vector<int> order(size_t(d->nblocks));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d->block_n[x] > d->block_n[y]; });
    std::vector<EighJob> ej;
    int64_t ws = 0;
    // size buckets of the workgroup eigensolver; inside a bucket the jobs keep the largest-first
    // order.  Blocks of size >= kEighLargeMin go to the device-driven large-block solver.
    std::vector<std::vector<int>> byb(kBuckets);
    std::vector<int> big;
    for(int x : order) {
        int n = d->block_n[x];
        if(n >= kEighLargeMin) {
            big.push_back(x);
            continue;
        }
        int q = 0;
        while(n > kBucketMax[q])
            ++q;
        byb[q].push_back(x);
    }
    int at = 0;
    for(int q = 0; q < kBuckets; ++q) {
        b->bucket_begin[q] = at;
        for(int x : byb[q]) {
            int n = d->block_n[x];
            ej.push_back(EighJob{n, 0, d->block_mat_off[x], d->block_state_off[x], ws});
            ws += eigh_ws_doubles(n);
            b->bucket_nmax[q] = std::max(b->bucket_nmax[q], n);
            ++at;
        }
    }
    b->bucket_begin[kBuckets] = at;
    b->ws_len = ws;
    // large blocks follow the bucketed ones in the job list (the adjoint walks all blocks)
    std::vector<int> big_n;
    std::vector<int64_t> big_mat, big_val;
    for(int x : big) {
        ej.push_back(EighJob{d->block_n[x], 0, d->block_mat_off[x], d->block_state_off[x], 0});
        big_n.push_back(d->block_n[x]);
        big_mat.push_back(d->block_mat_off[x]);
        big_val.push_back(d->block_state_off[x]);
    }
    b->nbig = int(big.size());
    if(!big.empty() && (rc = big_plan_create(ctx, b->nbig, big_n.data(), big_mat.data(), big_val.data(), &b->big))) {
        delete b;
        return rc;
    }
    // rotation tiles
    std::vector<RotTile> tiles;
    rot_tiles(d->jobs, d->njobs, tiles);
    b->ntiles = int(tiles.size());
    // G jobs grouped by instance (the jobs array is in instance order)
    b->h_jinfo.resize(size_t(d->njobs));
    b->inst_first_job.assign(size_t(d->ninst) + 1, 0);
    for(int q = 0; q < d->njobs; ++q) {
        if(q > 0 && d->job_inst[q] < d->job_inst[q - 1]) {
            delete b;
            return pomx_fail(ctx, POMX_ERR_ARG, "pomx_batch_create: jobs must be ordered by instance");
        }
        b->h_jinfo[size_t(q)] = JobInfo{d->jobs[q].n_from, d->jobs[q].n_to, d->job_inst[q], 0, d->jobs[q].out_off,
                                        d->job_from_state_off[q], d->job_to_state_off[q]};
    }
    for(int k = 0, q = 0; k <= d->ninst; ++k) {
        while(q < d->njobs && d->job_inst[q] < k)
            ++q;
        b->inst_first_job[size_t(k)] = q;
    }
    b->mat = dev_copy<double>(ctx, nullptr, size_t(d->mat_len), rc);
    b->eval = dev_copy<double>(ctx, nullptr, size_t(d->nstates), rc);
    b->w = dev_copy<double>(ctx, nullptr, size_t(d->nstates), rc);
    b->ws = dev_copy<double>(ctx, nullptr, size_t(std::max<int64_t>(ws, 1)), rc);
    b->h_pos = dev_copy(ctx, d->h_pos, size_t(d->nnz), rc);
    b->h_val = dev_copy(ctx, d->h_val, size_t(d->nnz), rc);
    b->beta = dev_copy(ctx, d->inst_beta, size_t(d->ninst), rc);
    b->inst_off = dev_copy(ctx, d->inst_state_off, size_t(d->ninst), rc);
    b->inst_cnt = dev_copy(ctx, d->inst_nstates, size_t(d->ninst), rc);
    b->src = dev_copy(ctx, d->src, size_t(d->src_len), rc);
    b->ops = dev_copy<double>(ctx, nullptr, size_t(std::max<int64_t>(d->ops_len, 1)), rc);
    b->G = dev_copy<double>(ctx, nullptr, size_t(d->ninst) * d->nfreq * 2, rc);
    b->E0 = dev_copy<double>(ctx, nullptr, size_t(d->ninst), rc);
    b->counters = dev_copy<int>(ctx, nullptr, size_t(kBuckets + d->nblocks), rc);
    b->status = b->counters + kBuckets;
    b->eigh_jobs = dev_copy(ctx, ej.data(), ej.size(), rc);
    b->jobs = dev_copy(ctx, d->jobs, size_t(d->njobs), rc);
    b->tiles = dev_copy(ctx, tiles.data(), tiles.size(), rc);
    b->kmap = dev_copy<int2>(ctx, nullptr, size_t(std::max<int64_t>(d->src_len, 1)), rc);
    b->kcnt = dev_copy<int>(ctx, nullptr, size_t(std::max(d->njobs, 1)), rc);
    if(!rc)
        rc = launch_rot_compact(ctx, b->jobs, d->njobs, b->src, b->kmap, b->kcnt);
    b->jinfo = dev_copy(ctx, b->h_jinfo.data(), b->h_jinfo.size(), rc);
    b->h_eigh = ej;
    b->h_jobs.assign(d->jobs, d->jobs + d->njobs);
    b->h_src.assign(d->src, d->src + d->src_len);
    b->h_pos_host.assign(d->h_pos, d->h_pos + d->nnz);
    if(!rc)
        rc = build_items(ctx, b);
    hipDeviceProp_t prop;
    if(!rc && hipGetDeviceProperties(&prop, ctx->device) == hipSuccess)
        b->eigh_grid = 2 * prop.multiProcessorCount;
    else if(!rc)
        b->eigh_grid = 512;
    if(rc) {
        free_batch(b);
        delete b;
        return rc;
    }
    *out = b;
    return POMX_OK;
}

extern "C" int pomx_batch_solve(pomx_ctx* ctx, pomx_batch* b, int use_graph) {
    POMX_REQUIRE(ctx, b, "pomx_batch_solve: null batch");
    if(!use_graph) {
        int rc = enqueue_solve(ctx, b);
        if(rc)
            return rc;
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return POMX_OK;
    }
    if(!b->exec) {
        POMX_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_solve(ctx, b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if(rc)
            return rc;
        if(e != hipSuccess)
            return pomx_fail(ctx, POMX_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        b->graph = g;
        POMX_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    POMX_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

extern "C" int pomx_batch_read(pomx_ctx* ctx, pomx_batch* b, double* h_G, double* h_E0) {
    POMX_REQUIRE(ctx, b, "pomx_batch_read: null batch");
    std::vector<int> st(size_t(b->nblocks));
    POMX_HIP(ctx, hipMemcpyAsync(st.data(), b->status, sizeof(int) * st.size(), hipMemcpyDeviceToHost, ctx->stream));
    if(h_G)
        POMX_HIP(ctx, hipMemcpyAsync(h_G, b->G, sizeof(double) * size_t(b->ninst) * b->nfreq * 2, hipMemcpyDeviceToHost,
                                     ctx->stream));
    if(h_E0)
        POMX_HIP(ctx, hipMemcpyAsync(h_E0, b->E0, sizeof(double) * size_t(b->ninst), hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for(int v : st)
        if(v)
            return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_batch: eigensolver did not converge for a block");
    return big_plan_status(ctx, b->big);
}

extern "C" int pomx_batch_tune(pomx_ctx* ctx, pomx_batch* b, char* report, size_t report_len) {
    POMX_REQUIRE(ctx, b, "pomx_batch_tune: null batch");
    drop_graph(b);
    hipStream_t s = ctx->stream;
    hipEvent_t e0 = ctx->ev0, e1 = ctx->ev1;
    int rc = enqueue_solve(ctx, b); // valid spectrum, rotated operators and weights for the G timing
    if(rc)
        return rc;
    // one bucket's eigensolve on freshly assembled matrices, timed alone
    auto time_bucket = [&](int q, float& ms) -> int {
        int const nj = b->bucket_begin[q + 1] - b->bucket_begin[q];
        for(int rep = 0; rep < 2; ++rep) {
            zero_kernel<<<unsigned(std::min<int64_t>((b->mat_len + 255) / 256, 65536)), 256, 0, s>>>(b->mat, b->mat_len);
            if(b->nnz)
                scatter_h_kernel<<<grid_for(b->nnz), kBlock, 0, s>>>(b->nnz, b->h_pos, b->h_val, b->mat);
            zero_int_kernel<<<grid_for(kBuckets + b->nblocks), kBlock, 0, s>>>(b->counters, kBuckets + b->nblocks);
            POMX_HIP(ctx, hipEventRecord(e0, s));
            int r = launch_eigh(ctx, b->eigh_jobs + b->bucket_begin[q], nj, b->bucket_nmax[q], b->eigh_threads[q],
                                std::min(nj, b->eigh_grid), b->counters + q, b->mat, b->eval, b->ws,
                                b->status + b->bucket_begin[q]);
            if(r)
                return r;
            POMX_HIP(ctx, hipEventRecord(e1, s));
        }
        POMX_HIP(ctx, hipEventSynchronize(e1));
        POMX_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        return POMX_OK;
    };
    std::string rep;
    int const cand[3] = {128, 256, 512};
    for(int q = kBuckets - 1; q >= 0; --q) {
        if(b->bucket_begin[q + 1] == b->bucket_begin[q])
            continue;
        float best = FLT_MAX;
        int bt = b->eigh_threads[q];
        for(int c : cand) {
            if(q == 0 && c > 128)
                continue;
            b->eigh_threads[q] = c;
            float ms = 0;
            if((rc = time_bucket(q, ms)))
                return rc;
            if(ms < best) {
                best = ms;
                bt = c;
            }
        }
        b->eigh_threads[q] = bt;
        rep += "eigh[n<=" + std::to_string(std::min(kBucketMax[q], b->bucket_nmax[q])) + "]=" + std::to_string(bt) + " ";
    }
    if((rc = enqueue_solve(ctx, b))) // restore a consistent state with the chosen sizes
        return rc;
    // G work-item size: the fused kernel and its reduction, timed alone
    int64_t const gc[3] = {1024, 4096, 16384};
    float best = FLT_MAX;
    int64_t bg = b->terms_per_item;
    for(int64_t c : gc) {
        b->terms_per_item = c;
        if((rc = build_items(ctx, b)))
            return rc;
        float ms = 0;
        for(int r = 0; r < 2; ++r) {
            POMX_HIP(ctx, hipEventRecord(e0, s));
            if(b->nitems)
                fused_g_kernel<<<b->nitems, kGThreads, 0, s>>>(b->items, b->jinfo, b->ops, b->eval, b->w, b->beta, b->nfreq,
                                                               b->tol, b->partial);
            g_reduce_kernel<<<grid_for(int64_t(b->ninst) * b->nfreq), kBlock, 0, s>>>(b->ninst, b->nfreq, b->item_off,
                                                                                        b->partial, b->G);
            POMX_CHECK_LAUNCH(ctx);
            POMX_HIP(ctx, hipEventRecord(e1, s));
        }
        POMX_HIP(ctx, hipEventSynchronize(e1));
        POMX_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        if(ms < best) {
            best = ms;
            bg = c;
        }
    }
    b->terms_per_item = bg;
    if((rc = build_items(ctx, b)))
        return rc;
    rep += "g_terms_per_item=" + std::to_string(bg) + " g_ms=" + std::to_string(best);
    if(b->nbig)
        rep += " large_blocks=" + std::to_string(b->nbig);
    if(report && report_len) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return POMX_OK;
}

extern "C" void pomx_batch_destroy(pomx_ctx* ctx, pomx_batch* b) {
    if(!b)
        return;
    if(ctx)
        (void)hipStreamSynchronize(ctx->stream);
    drop_graph(b);
    adjoint_free(b);
    big_plan_destroy(b->big);
    free_batch(b);
    delete b;
}
