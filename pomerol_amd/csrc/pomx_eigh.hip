// K2: batched dense real-symmetric eigensolve, one workgroup per symmetry block.
//
// Replaces Eigen::SelfAdjointEigenSolver in HamiltonianPart::computeImpl
// (src/pomerol/HamiltonianPart.cpp:90-102).  Per block:
//   1. Householder tridiagonalisation T = Q^T A Q (LAPACK dsytd2 reflector convention);
//      the trailing update streams the block through L2 with coalesced column-parallel
//      mat-vec and rank-2 updates,
//   2. implicit QL with Wilkinson shifts on (d, e): lane 0 computes the Givens chain of one
//      QL sweep into LDS while the whole workgroup applies the previous chain to Z^T
//      rows (each thread owns one component of every eigenvector, so the rotations of a
//      sweep need no barriers between them),
//   3. back-transformation X = Q Z, reflectors applied in reverse order,
//   4. eigenvalues ranked ascending; eigenvectors written as columns (row-major block), the
//      output order of HamiltonianPart (eigenvectors overwrite the block, ascending values).
#include "pomx_internal.hpp"

#include <algorithm>
#include <cfloat>

namespace {

constexpr int kMaxLdsN = 2048;

__device__ inline double wave_sum(double v) {
    for(int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double block_sum(double v, double* red) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if(lane == 0)
        red[wave] = v;
    __syncthreads();
    double s = 0;
    int nw = blockDim.x >> 6;
    for(int i = 0; i < nw; ++i)
        s += red[i];
    __syncthreads();
    return s;
}

// One block, one workgroup (128..512 threads).  status_slot receives 1 (QL did not converge) or
// 2 (NaN eigenvalue); it is left untouched on success.
__device__ void eigh_one(EighJob const J, double* __restrict__ mat, double* __restrict__ eval, double* __restrict__ ws,
                         int* __restrict__ status_slot) {
    extern __shared__ double sh[];
    int const n = J.n;
    int const tid = threadIdx.x, nt = blockDim.x;
    double* __restrict__ A = mat + J.mat_off;
    double* __restrict__ W = eval + J.val_off;
    if(n == 1) {
        if(tid == 0) {
            W[0] = A[0];
            A[0] = 1.0;
        }
        return;
    }
    double* __restrict__ ZT = ws + J.ws_off; // n x n, row j = eigenvector j of T
    double* __restrict__ tau = ZT + int64_t(n) * n;

    double* v = sh;
    double* w = v + n;
    double* d = w + n;
    double* e = d + n;
    double* cs = e + n;
    double* sn = cs + n;
    double* red = sn + n;
    int* ctl = reinterpret_cast<int*>(red + 8);

    // ---- 1. tridiagonalisation ----------------------------------------------------------
    for(int k = 0; k < n - 1; ++k) {
        int const m = n - k - 1;
        double* rowk = A + int64_t(k) * n + k + 1;
        double s = 0;
        for(int i = 1 + tid; i < m; i += nt) {
            double x = rowk[i];
            s += x * x;
        }
        s = block_sum(s, red);
        double const alpha = rowk[0];
        double t = 0.0, beta = alpha;
        if(s > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + s), alpha);
            t = (beta - alpha) / beta;
            double const scale = 1.0 / (alpha - beta);
            for(int i = tid; i < m; i += nt)
                v[i] = (i == 0) ? 1.0 : rowk[i] * scale;
        }
        if(tid == 0) {
            d[k] = A[int64_t(k) * n + k];
            e[k] = beta;
            tau[k] = t;
        }
        __syncthreads();
        if(t != 0.0) {
            double* A22 = A + int64_t(k + 1) * n + (k + 1);
            for(int c = tid; c < m; c += nt) {
                double acc = 0.0;
                for(int r = 0; r < m; ++r)
                    acc += A22[int64_t(r) * n + c] * v[r];
                w[c] = t * acc;
            }
            __syncthreads();
            double pv = 0;
            for(int i = tid; i < m; i += nt)
                pv += w[i] * v[i];
            pv = block_sum(pv, red);
            double const K = 0.5 * t * pv;
            for(int i = tid; i < m; i += nt)
                w[i] -= K * v[i];
            __syncthreads();
            for(int r = 0; r < m; ++r) {
                double const vr = v[r], wr = w[r];
                double* row = A22 + int64_t(r) * n;
                for(int c = tid; c < m; c += nt)
                    row[c] -= vr * w[c] + wr * v[c];
            }
            for(int i = 1 + tid; i < m; i += nt)
                rowk[i] = v[i]; // reflector kept in the (dead) row k for step 3
        }
        __syncthreads();
    }
    if(tid == 0) {
        d[n - 1] = A[int64_t(n - 1) * n + (n - 1)];
        e[n - 1] = 0.0;
        ctl[3] = 0;
    }
    for(int64_t i = tid; i < int64_t(n) * n; i += nt)
        ZT[i] = ((i / n) == (i % n)) ? 1.0 : 0.0;
    __syncthreads();

    // ---- 2. implicit QL (Wilkinson shift) with Givens accumulation ------------------------
    for(int l = 0; l < n; ++l) {
        int iter = 0;
        while(true) {
            if(tid == 0) {
                int m;
                for(m = l; m < n - 1; ++m) {
                    double dd = fabs(d[m]) + fabs(d[m + 1]);
                    if(fabs(e[m]) <= DBL_EPSILON * dd)
                        break;
                }
                if(m != l && ++iter > 200) {
                    ctl[3] = 1;
                    m = l;
                }
                ctl[0] = m;
                if(m != l) {
                    double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                    double r = hypot(g, 1.0);
                    g = d[m] - d[l] + e[l] / (g + copysign(r, g));
                    double s = 1.0, c = 1.0, p = 0.0;
                    int i;
                    for(i = m - 1; i >= l; --i) {
                        double f = s * e[i], b = c * e[i];
                        r = hypot(f, g);
                        e[i + 1] = r;
                        if(r == 0.0) {
                            d[i + 1] -= p;
                            e[m] = 0.0;
                            break;
                        }
                        s = f / r;
                        c = g / r;
                        g = d[i + 1] - p;
                        r = (d[i] - g) * s + 2.0 * c * b;
                        p = s * r;
                        d[i + 1] = g + p;
                        g = c * r - b;
                        cs[i] = c;
                        sn[i] = s;
                    }
                    int lo = l;
                    if(r == 0.0 && i >= l)
                        lo = i + 1;
                    else {
                        d[l] -= p;
                        e[l] = g;
                        e[m] = 0.0;
                    }
                    ctl[1] = lo;
                }
            }
            __syncthreads();
            int const m = ctl[0];
            if(m == l)
                break;
            int const lo = ctl[1];
            for(int k = tid; k < n; k += nt) {
                double cur = ZT[int64_t(m) * n + k];
                for(int i = m - 1; i >= lo; --i) {
                    double zi = ZT[int64_t(i) * n + k];
                    ZT[int64_t(i + 1) * n + k] = sn[i] * zi + cs[i] * cur;
                    cur = cs[i] * zi - sn[i] * cur;
                }
                ZT[int64_t(lo) * n + k] = cur;
            }
            __syncthreads();
        }
    }

    // ---- 3. back-transformation X = H_0 ... H_{n-2} Z -------------------------------------
    int const lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    for(int k = n - 2; k >= 0; --k) {
        double const t = tau[k];
        if(t == 0.0)
            continue;
        int const m = n - k - 1;
        for(int i = tid; i < m; i += nt)
            v[i] = (i == 0) ? 1.0 : A[int64_t(k) * n + k + 1 + i];
        __syncthreads();
        for(int j = wave; j < n; j += nw) {
            double* z = ZT + int64_t(j) * n + (k + 1);
            double s = 0.0;
            for(int i = lane; i < m; i += 64)
                s += v[i] * z[i];
            s = wave_sum(s) * t;
            for(int i = lane; i < m; i += 64)
                z[i] -= s * v[i];
        }
        __syncthreads();
    }

    // ---- 4. ascending order, eigenvectors as columns --------------------------------------
    int* rank = reinterpret_cast<int*>(cs);
    for(int j = tid; j < n; j += nt) {
        double dj = d[j];
        int rk = 0;
        for(int i = 0; i < n; ++i) {
            double di = d[i];
            rk += (di < dj) || (di == dj && i < j);
        }
        rank[j] = rk;
        if(!(dj == dj))
            ctl[3] = 2;
    }
    __syncthreads();
    for(int j = tid; j < n; j += nt)
        W[rank[j]] = d[j];
    for(int64_t idx = tid; idx < int64_t(n) * n; idx += nt) {
        int j = int(idx / n), r = int(idx % n);
        A[int64_t(r) * n + rank[j]] = ZT[idx];
    }
    __syncthreads();
    if(tid == 0 && ctl[3] != 0)
        *status_slot = ctl[3];
}

// One workgroup per block (grid = number of blocks), or PERSISTENT workgroups that pull block
// indices from *counter until the list is exhausted (every workgroup reaches the exit test).
template <bool PERSISTENT>
__global__ void __launch_bounds__(512) eigh_kernel(const EighJob* __restrict__ jobs, int njobs, int* __restrict__ counter,
                                                   double* __restrict__ mat, double* __restrict__ eval,
                                                   double* __restrict__ ws, int* __restrict__ status) {
    if(!PERSISTENT) {
        eigh_one(jobs[blockIdx.x], mat, eval, ws, status + blockIdx.x);
        return;
    }
    __shared__ int next;
    while(true) {
        __syncthreads();
        if(threadIdx.x == 0)
            next = atomicAdd(counter, 1);
        __syncthreads();
        int const j = next;
        if(j >= njobs)
            break;
        eigh_one(jobs[j], mat, eval, ws, status + j);
    }
}

} // namespace

size_t eigh_lds_bytes(int nmax) { return (size_t(6) * nmax + 8) * sizeof(double) + 16 * sizeof(int); }
int64_t eigh_ws_doubles(int n) { return n > 1 ? int64_t(n) * n + n : 0; }
int eigh_max_n() { return kMaxLdsN; }

int launch_eigh(pomx_ctx* ctx, const EighJob* d_jobs, int njobs, int nmax, int block_threads, int persistent_grid,
                int* counter, double* d_mat, double* d_eval, double* d_ws, int* d_status) {
    if(njobs <= 0)
        return POMX_OK;
    size_t lds = eigh_lds_bytes(nmax);
    if(persistent_grid > 0) {
        POMX_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(eigh_kernel<true>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        eigh_kernel<true><<<persistent_grid, block_threads, lds, ctx->stream>>>(d_jobs, njobs, counter, d_mat, d_eval, d_ws,
                                                                                 d_status);
    } else {
        POMX_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(eigh_kernel<false>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        eigh_kernel<false><<<njobs, block_threads, lds, ctx->stream>>>(d_jobs, njobs, counter, d_mat, d_eval, d_ws, d_status);
    }
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

extern "C" int pomx_eigh_batched(pomx_ctx* ctx, int nblocks, const int* h_n, const int64_t* h_mat_off,
                                 const int64_t* h_val_off, double* d_mat, double* d_eval) {
    POMX_REQUIRE(ctx, nblocks >= 0 && (nblocks == 0 || (h_n && h_mat_off && h_val_off && d_mat && d_eval)),
                 "pomx_eigh_batched: bad arguments");
    if(nblocks == 0)
        return POMX_OK;
    std::vector<int> order(nblocks);
    for(int b = 0; b < nblocks; ++b)
        order[b] = b;
    // largest blocks first so the longest workgroups start earliest
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h_n[a] > h_n[b]; });
    std::vector<EighJob> jobs;
    std::vector<int> large;
    int64_t ws_total = 0;
    int nmax = 1;
    for(int k = 0; k < nblocks; ++k) {
        int b = order[k];
        int n = h_n[b];
        POMX_REQUIRE(ctx, n >= 1, "pomx_eigh_batched: block size must be >= 1");
        if(n >= kEighLargeMin) {
            large.push_back(b);
            continue;
        }
        jobs.push_back(EighJob{n, 0, h_mat_off[b], h_val_off[b], ws_total});
        ws_total += eigh_ws_doubles(n);
        nmax = std::max(nmax, n);
    }
    // large blocks: device-driven solver on all of them together, in groups whose scratch fits
    // a fixed budget (the largest blocks first)
    size_t const budget = size_t(64) << 30;
    for(size_t g0 = 0; g0 < large.size();) {
        size_t g1 = g0;
        size_t bytes = 0;
        std::vector<int> gn;
        std::vector<int64_t> gm, gv;
        while(g1 < large.size()) {
            size_t const need = big_block_bytes(h_n[large[g1]]);
            if(!gn.empty() && bytes + need > budget)
                break;
            bytes += need;
            gn.push_back(h_n[large[g1]]);
            gm.push_back(h_mat_off[large[g1]]);
            gv.push_back(h_val_off[large[g1]]);
            ++g1;
        }
        BigPlan* plan = nullptr;
        int rc = big_plan_create(ctx, int(gn.size()), gn.data(), gm.data(), gv.data(), &plan);
        if(rc)
            return rc;
        rc = big_plan_enqueue(ctx, plan, d_mat, d_eval);
        if(!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = pomx_fail(ctx, POMX_ERR_HIP, "pomx_eigh_batched: large-block solve failed");
        if(!rc)
            rc = big_plan_status(ctx, plan);
        big_plan_destroy(plan);
        if(rc)
            return rc;
        g0 = g1;
    }
    int const nj = int(jobs.size());
    if(nj == 0)
        return POMX_OK;
    WsBuf<EighJob> d_jobs(ctx, nj);
    WsBuf<double> d_ws(ctx, size_t(std::max<int64_t>(ws_total, 1)));
    WsBuf<int> d_status(ctx, nj);
    if(!d_jobs.p || !d_ws.p || !d_status.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_eigh_batched: out of device memory");
    POMX_HIP(ctx, h2d_sync(ctx, d_jobs.p, jobs.data(), sizeof(EighJob) * nj));
    int rc = pomx_fill_zero(ctx, d_status.p, sizeof(int) * nj);
    if(rc)
        return rc;
    if((rc = launch_eigh(ctx, d_jobs, nj, nmax, 256, 0, nullptr, d_mat, d_eval, d_ws, d_status)))
        return rc;
    std::vector<int> st(nj);
    POMX_HIP(ctx, hipMemcpyAsync(st.data(), d_status.p, sizeof(int) * nj, hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for(int k = 0; k < nj; ++k)
        if(st[k])
            return pomx_fail(ctx, POMX_ERR_RANGE,
                             std::string("pomx_eigh_batched: ") + (st[k] == 1 ? "QL did not converge" : "NaN eigenvalue") +
                                 " in block of size " + std::to_string(jobs[k].n));
    return POMX_OK;
}
