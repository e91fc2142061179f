// Adjoint of the batched G(i w_n) solve: dL/d(Hamiltonian entries) for every instance, given
// Gbar_n = dL/dRe G(i w_n) + i dL/dIm G(i w_n).  No counterpart in the reference.
//
// Forward (pomx_batch.hip), per instance:
//   H_b = U_b diag(E_b) U_b^T                              (every block b)
//   w_s = exp(-beta E_s) / sum_t exp(-beta E_t)
//   A   = U_t^T X,  X = M U_f                               (c^+ from block f to block t)
//   G_n = sum_{a in t, b in f, |R_ab| > tol} R_ab / (i w_n - P_ab),
//         R_ab = A_ab^2 (w_a + w_b),  P_ab = E_a - E_b
// Reverse, in this order (the rotated operator is recomputed from U, never stored):
//   1. Lehmann adjoint per term: Rbar = sum_n Re[conj(Gbar_n) / (z_n - P)],
//      Pbar = sum_n Re[conj(Gbar_n) R / (z_n - P)^2];  Abar = 2 A (w_a + w_b) Rbar,
//      wbar_a, wbar_b += A^2 Rbar;  Ebar_a += Pbar, Ebar_b -= Pbar
//   2. weights: Ebar_s += -beta w_s (wbar_s - sum_t w_t wbar_t)
//   3. rotation: Ubar_t = X Abar^T,  Ubar_f += M^T (U_t Abar)          (MFMA GEMMs + gather)
//   4. eigen-decomposition: K = diag(Ebar) + F o (U^T Ubar), F_ij = 1/(E_j - E_i) for
//      non-degenerate pairs and 0 inside a degenerate subspace (G is invariant under rotations
//      inside such subspaces, so those pairs carry no gradient), Hbar = U K U^T
//   5. dL/dh_val[u] = Hbar[pos_u] = sum_i (U K)[r][i] U[c][i]
#include "pomx_batch.hpp"

#include <algorithm>

namespace {

struct Line { // one row (i_in fixed) or one column (i_out fixed) of a job's term matrix
    int job, idx, is_col, pad;
};

struct RowMove { // Ubar_f[b][:] += sign * Y[a][:]
    int job, a, b, sign;
};

__device__ inline double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Step 1.  One thread per term (terms in the c^+ tile order, t = i_in * n_out + i_out).
__global__ void __launch_bounds__(256) lehmann_adjoint_kernel(const GItem* __restrict__ items, const JobInfo* __restrict__ jobs,
                                                              const double* __restrict__ ops, const double* __restrict__ E,
                                                              const double* __restrict__ w, const double* __restrict__ beta,
                                                              const double* __restrict__ gbar, int nfreq, double tol,
                                                              double* __restrict__ abar, double* __restrict__ cw,
                                                              double* __restrict__ cp) {
    __shared__ double s_om[128], s_gr[128], s_gi[128];
    GItem const it = items[blockIdx.x];
    JobInfo const J = jobs[it.job];
    for(int f = threadIdx.x; f < nfreq; f += blockDim.x) {
        s_om[f] = M_PI * (2 * f + 1) / beta[it.inst];
        s_gr[f] = gbar[(int64_t(it.inst) * nfreq + f) * 2];
        s_gi[f] = gbar[(int64_t(it.inst) * nfreq + f) * 2 + 1];
    }
    __syncthreads();
    for(int64_t t = it.t0 + threadIdx.x; t < it.t1; t += blockDim.x) {
        int const i_in = int(t / J.n_out), i_out = int(t % J.n_out);
        int64_t const idx = J.ops_off + t;
        double const m = ops[idx];
        double const wsum = w[J.out_state + i_out] + w[J.in_state + i_in];
        double const R = m * m * wsum;
        double const P = E[J.in_state + i_in] - E[J.out_state + i_out];
        double rb = 0.0, pb = 0.0;
        if(m != 0.0 && fabs(R) > tol) {
            double const P2 = P * P;
            for(int f = 0; f < nfreq; ++f) {
                double const om = s_om[f], gr = s_gr[f], gi = s_gi[f];
                double const inv = rcp_nr(fma(om, om, P2)); // 1/|z - P|^2
                rb = fma(-(gr * P + gi * om), inv, rb);
                pb = fma(fma(gr, P2 - om * om, 2.0 * gi * P * om), inv * inv, pb);
            }
            pb *= R;
        }
        abar[idx] = 2.0 * m * wsum * rb;
        cw[idx] = m * m * rb;
        cp[idx] = pb;
    }
}

// Row sums (into the 'in' state) and column sums (into the 'out' state) of cw and cp.
__global__ void line_sum_kernel(int nlines, const Line* __restrict__ lines, const JobInfo* __restrict__ jobs,
                                const double* __restrict__ cw, const double* __restrict__ cp, double* __restrict__ in_w,
                                double* __restrict__ in_e, double* __restrict__ out_w, double* __restrict__ out_e) {
    int const q = blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= nlines)
        return;
    Line const L = lines[q];
    JobInfo const J = jobs[L.job];
    double sw = 0.0, se = 0.0;
    if(!L.is_col) {
        const double* a = cw + J.ops_off + int64_t(L.idx) * J.n_out;
        const double* b = cp + J.ops_off + int64_t(L.idx) * J.n_out;
        for(int k = 0; k < J.n_out; ++k) {
            sw += a[k];
            se += b[k];
        }
        in_w[J.in_state + L.idx] = sw;
        in_e[J.in_state + L.idx] = se;
    } else {
        const double* a = cw + J.ops_off + L.idx;
        const double* b = cp + J.ops_off + L.idx;
        for(int k = 0; k < J.n_in; ++k) {
            sw += a[int64_t(k) * J.n_out];
            se += b[int64_t(k) * J.n_out];
        }
        out_w[J.out_state + L.idx] = sw;
        out_e[J.out_state + L.idx] = -se;
    }
}

// Step 2, one workgroup per instance.
__global__ void __launch_bounds__(256) weights_adjoint_kernel(const int64_t* __restrict__ off, const int64_t* __restrict__ cnt,
                                                              const double* __restrict__ beta, const double* __restrict__ w,
                                                              const double* __restrict__ in_w, const double* __restrict__ in_e,
                                                              const double* __restrict__ out_w,
                                                              const double* __restrict__ out_e, double* __restrict__ ebar) {
    __shared__ double red[256];
    int const k = blockIdx.x;
    int64_t const o = off[k], n = cnt[k];
    double s = 0.0;
    for(int64_t i = threadIdx.x; i < n; i += blockDim.x)
        s += w[o + i] * (in_w[o + i] + out_w[o + i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for(int h = blockDim.x / 2; h > 0; h >>= 1) {
        if(threadIdx.x < h)
            red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    double const mean = red[0];
    for(int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        double const wb = in_w[o + i] + out_w[o + i];
        ebar[o + i] = in_e[o + i] + out_e[o + i] - beta[k] * w[o + i] * (wb - mean);
    }
}

// X = M U_f into the operator layout (n_t x n_f per job): X[a][:] = sign * U_f[src][:]
__global__ void gather_x_kernel(int njobs, const pomx_rot_job* __restrict__ jobs, const double* __restrict__ U,
                                const int* __restrict__ src, double* __restrict__ X) {
    int const q = blockIdx.y;
    if(q >= njobs)
        return;
    pomx_rot_job const J = jobs[q];
    int64_t const n = int64_t(J.n_to) * J.n_from;
    for(int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
        int const a = int(e / J.n_from), j = int(e % J.n_from);
        int const sv = src[J.src_off + a];
        double x = 0.0;
        if(sv != 0) {
            double u = U[J.u_from_off + int64_t((sv > 0 ? sv : -sv) - 1) * J.n_from + j];
            x = sv > 0 ? u : -u;
        }
        X[J.out_off + e] = x;
    }
}

// Ubar_f[b][:] += sign * Y[a][:]   (one workgroup per moved row)
__global__ void scatter_rows_kernel(const RowMove* __restrict__ moves, const pomx_rot_job* __restrict__ jobs,
                                    const double* __restrict__ Y, double* __restrict__ Ubar) {
    RowMove const mv = moves[blockIdx.x];
    pomx_rot_job const J = jobs[mv.job];
    const double* y = Y + J.out_off + int64_t(mv.a) * J.n_from;
    double* u = Ubar + J.u_from_off + int64_t(mv.b) * J.n_from;
    double const s = double(mv.sign);
    for(int j = threadIdx.x; j < J.n_from; j += blockDim.x)
        u[j] = fma(s, y[j], u[j]);
}

// K = diag(Ebar) + F o S in place of S, one workgroup per block.
__global__ void kmat_kernel(const EighJob* __restrict__ blocks, const double* __restrict__ E, const double* __restrict__ ebar,
                            double* __restrict__ S) {
    EighJob const B = blocks[blockIdx.x];
    int const n = B.n;
    const double* e = E + B.val_off;
    double* s = S + B.mat_off;
    for(int64_t q = threadIdx.x; q < int64_t(n) * n; q += blockDim.x) {
        int const i = int(q / n), j = int(q % n);
        double v;
        if(i == j)
            v = ebar[B.val_off + i];
        else {
            double const d = e[j] - e[i];
            double const scale = fmax(1.0, fmax(fabs(e[i]), fabs(e[j])));
            v = (fabs(d) <= 1e-10 * scale) ? 0.0 : s[q] / d;
        }
        s[q] = v;
    }
}

// dval[u] = sum_i V[r][i] U[c][i], one wavefront per Hamiltonian entry.
__global__ void __launch_bounds__(256) entry_grad_kernel(int64_t nnz, const int64_t* __restrict__ pos,
                                                         const int* __restrict__ pos_block, const EighJob* __restrict__ blocks,
                                                         const double* __restrict__ V, const double* __restrict__ U,
                                                         double* __restrict__ dval) {
    int64_t const u = (blockIdx.x * int64_t(blockDim.x) + threadIdx.x) >> 6;
    int const lane = threadIdx.x & 63;
    if(u >= nnz)
        return;
    EighJob const B = blocks[pos_block[u]];
    int64_t const rel = pos[u] - B.mat_off;
    int const r = int(rel / B.n), c = int(rel % B.n);
    const double* v = V + B.mat_off + int64_t(r) * B.n;
    const double* x = U + B.mat_off + int64_t(c) * B.n;
    double s = 0.0;
    for(int i = lane; i < B.n; i += 64)
        s = fma(v[i], x[i], s);
    for(int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o, 64);
    if(lane == 0)
        dval[u] = s;
}

__global__ void zero_kernel(double* __restrict__ p, int64_t n) {
    for(int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = 0.0;
}

unsigned zero_grid(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536))); }

template <class T> int upload(pomx_ctx* ctx, T*& d, std::vector<T> const& h) {
    if(device_malloc(ctx, reinterpret_cast<void**>(&d), sizeof(T) * std::max<size_t>(h.size(), 1)) != hipSuccess)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_batch_adjoint: out of device memory");
    if(!h.empty())
        POMX_HIP(ctx, h2d_sync(ctx, d, h.data(), sizeof(T) * h.size()));
    return POMX_OK;
}

int alloc_doubles(pomx_ctx* ctx, double*& d, int64_t n) {
    if(device_malloc(ctx, reinterpret_cast<void**>(&d), sizeof(double) * size_t(std::max<int64_t>(n, 1))) != hipSuccess)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_batch_adjoint: out of device memory");
    return POMX_OK;
}

} // namespace

struct AdjointState {
    double *abar = nullptr, *cw = nullptr, *cp = nullptr, *xs = nullptr, *ubar = nullptr, *sk = nullptr;
    double *in_w = nullptr, *in_e = nullptr, *out_w = nullptr, *out_e = nullptr, *ebar = nullptr, *gbar = nullptr,
           *dval = nullptr;
    Line* lines = nullptr;
    RowMove* moves = nullptr;
    int* pos_block = nullptr;
    int nlines = 0, nmoves = 0;
    // GEMM launches: to-blocks Ubar_t = X Abar^T (NT), Y = U_t Abar (NN), S = U^T Ubar (TN), V = U K (NN)
    pomx_gemm_job *g_to = nullptr, *g_y = nullptr, *g_s = nullptr, *g_v = nullptr;
    GemmTile *t_to = nullptr, *t_y = nullptr, *t_s = nullptr, *t_v = nullptr;
    int n_to = 0, n_y = 0, n_s = 0, n_v = 0;
};

void adjoint_free(pomx_batch* b) {
    AdjointState* a = b->adj;
    if(!a)
        return;
    void* ptrs[] = {a->abar, a->cw,   a->cp,   a->xs,    a->ubar,  a->sk,    a->in_w, a->in_e, a->out_w, a->out_e,
                    a->ebar, a->gbar, a->dval, a->lines, a->moves, a->pos_block, a->g_to, a->g_y,  a->g_s,   a->g_v,
                    a->t_to, a->t_y,  a->t_s,  a->t_v};
    for(void* p : ptrs)
        if(p)
            (void)hipFree(p);
    delete a;
    b->adj = nullptr;
}

namespace {

int adjoint_init(pomx_ctx* ctx, pomx_batch* b) {
    if(b->njobs > 65535)
        return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_batch_adjoint: more than 65535 connected block pairs");
    auto* a = new AdjointState();
    b->adj = a;
    int64_t const ops_len = [&] {
        int64_t m = 1;
        for(auto const& j : b->h_jobs)
            m = std::max<int64_t>(m, j.out_off + int64_t(j.n_to) * j.n_from);
        return m;
    }();
    int rc = POMX_OK;
    for(double** p : {&a->abar, &a->cw, &a->cp, &a->xs})
        if(!rc)
            rc = alloc_doubles(ctx, *p, ops_len);
    for(double** p : {&a->ubar, &a->sk})
        if(!rc)
            rc = alloc_doubles(ctx, *p, b->mat_len);
    for(double** p : {&a->in_w, &a->in_e, &a->out_w, &a->out_e, &a->ebar})
        if(!rc)
            rc = alloc_doubles(ctx, *p, b->nstates);
    if(!rc)
        rc = alloc_doubles(ctx, a->gbar, int64_t(b->ninst) * b->nfreq * 2);
    if(!rc)
        rc = alloc_doubles(ctx, a->dval, b->nnz);
    if(rc)
        return rc;
    // lines and row moves
    std::vector<Line> lines;
    std::vector<RowMove> moves;
    for(int q = 0; q < b->njobs; ++q) {
        JobInfo const& J = b->h_jinfo[size_t(q)];
        for(int i = 0; i < J.n_in; ++i)
            lines.push_back(Line{q, i, 0, 0});
        for(int i = 0; i < J.n_out; ++i)
            lines.push_back(Line{q, i, 1, 0});
        pomx_rot_job const& R = b->h_jobs[size_t(q)];
        for(int x = 0; x < R.n_to; ++x) {
            int sv = b->h_src[size_t(R.src_off + x)];
            if(sv != 0)
                moves.push_back(RowMove{q, x, (sv > 0 ? sv : -sv) - 1, sv > 0 ? 1 : -1});
        }
    }
    a->nlines = int(lines.size());
    a->nmoves = int(moves.size());
    // block of every Hamiltonian entry
    std::vector<int> order(b->h_eigh.size());
    for(size_t i = 0; i < order.size(); ++i)
        order[i] = int(i);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return b->h_eigh[size_t(x)].mat_off < b->h_eigh[size_t(y)].mat_off; });
    std::vector<int64_t> starts(order.size());
    for(size_t i = 0; i < order.size(); ++i)
        starts[i] = b->h_eigh[size_t(order[i])].mat_off;
    std::vector<int> pos_block(size_t(b->nnz));
    for(int64_t u = 0; u < b->nnz; ++u) {
        int64_t p = b->h_pos_host[size_t(u)];
        size_t k = size_t(std::upper_bound(starts.begin(), starts.end(), p) - starts.begin()) - 1;
        EighJob const& B = b->h_eigh[size_t(order[k])];
        if(p < B.mat_off || p >= B.mat_off + int64_t(B.n) * B.n)
            return pomx_fail(ctx, POMX_ERR_ARG, "pomx_batch_adjoint: Hamiltonian entry outside every block");
        pos_block[size_t(u)] = order[k];
    }
    // GEMM job lists
    std::vector<pomx_gemm_job> g_to, g_y, g_s, g_v;
    for(auto const& R : b->h_jobs) {
        int const nt = R.n_to, nf = R.n_from;
        g_to.push_back(pomx_gemm_job{nt, nt, nf, 0, 1, nf, nf, nt, R.out_off, R.out_off, R.u_to_off, 1.0, 0.0});
        g_y.push_back(pomx_gemm_job{nt, nf, nt, 0, 0, nt, nf, nf, R.u_to_off, R.out_off, R.out_off, 1.0, 0.0});
    }
    for(auto const& B : b->h_eigh) {
        int const n = B.n;
        g_s.push_back(pomx_gemm_job{n, n, n, 1, 0, n, n, n, B.mat_off, B.mat_off, B.mat_off, 1.0, 0.0});
        g_v.push_back(pomx_gemm_job{n, n, n, 0, 0, n, n, n, B.mat_off, B.mat_off, B.mat_off, 1.0, 0.0});
    }
    auto mk = [&](std::vector<pomx_gemm_job> const& js, pomx_gemm_job*& dj, GemmTile*& dt, int& nt) -> int {
        std::vector<GemmTile> tiles;
        gemm_tiles(js.data(), int(js.size()), tiles);
        nt = int(tiles.size());
        int r = upload(ctx, dj, js);
        return r ? r : upload(ctx, dt, tiles);
    };
    if((rc = upload(ctx, a->lines, lines)) || (rc = upload(ctx, a->moves, moves)) || (rc = upload(ctx, a->pos_block, pos_block)) ||
       (rc = mk(g_to, a->g_to, a->t_to, a->n_to)) || (rc = mk(g_y, a->g_y, a->t_y, a->n_y)) ||
       (rc = mk(g_s, a->g_s, a->t_s, a->n_s)) || (rc = mk(g_v, a->g_v, a->t_v, a->n_v)))
        return rc;
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

} // namespace

extern "C" int pomx_batch_set_values(pomx_ctx* ctx, pomx_batch* b, const double* h_val) {
    POMX_REQUIRE(ctx, b && (h_val || b->nnz == 0), "pomx_batch_set_values: bad arguments");
    if(b->nnz)
        POMX_HIP(ctx, hipMemcpyAsync(b->h_val, h_val, sizeof(double) * size_t(b->nnz), hipMemcpyHostToDevice, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

extern "C" int pomx_batch_adjoint(pomx_ctx* ctx, pomx_batch* b, const double* h_Gbar, double* h_dval) {
    POMX_REQUIRE(ctx, b && h_Gbar && (h_dval || b->nnz == 0), "pomx_batch_adjoint: bad arguments");
    int rc = POMX_OK;
    if(!b->adj && (rc = adjoint_init(ctx, b)))
        return rc;
    AdjointState* a = b->adj;
    hipStream_t s = ctx->stream;
    POMX_HIP(ctx, hipMemcpyAsync(a->gbar, h_Gbar, sizeof(double) * size_t(b->ninst) * b->nfreq * 2, hipMemcpyHostToDevice, s));
    // 1. Lehmann adjoint
    if(b->nitems)
        lehmann_adjoint_kernel<<<b->nitems, 256, 0, s>>>(b->items, b->jinfo, b->ops, b->eval, b->w, b->beta, a->gbar, b->nfreq,
                                                         b->tol, a->abar, a->cw, a->cp);
    for(double* p : {a->in_w, a->in_e, a->out_w, a->out_e})
        zero_kernel<<<zero_grid(b->nstates), 256, 0, s>>>(p, b->nstates);
    if(a->nlines)
        line_sum_kernel<<<grid_for(a->nlines), kBlock, 0, s>>>(a->nlines, a->lines, b->jinfo, a->cw, a->cp, a->in_w, a->in_e,
                                                               a->out_w, a->out_e);
    // 2. weights
    weights_adjoint_kernel<<<b->ninst, 256, 0, s>>>(b->inst_off, b->inst_cnt, b->beta, b->w, a->in_w, a->in_e, a->out_w,
                                                    a->out_e, a->ebar);
    // 3. rotation
    zero_kernel<<<zero_grid(b->mat_len), 256, 0, s>>>(a->ubar, b->mat_len);
    if(b->njobs) {
        gather_x_kernel<<<dim3(64, unsigned(b->njobs)), 256, 0, s>>>(b->njobs, b->jobs, b->mat, b->src, a->xs);
        if((rc = launch_gemm(ctx, 0, 1, a->t_to, a->n_to, a->g_to, a->xs, a->abar, a->ubar)))
            return rc;
        if((rc = launch_gemm(ctx, 0, 0, a->t_y, a->n_y, a->g_y, b->mat, a->abar, a->xs)))
            return rc;
        if(a->nmoves)
            scatter_rows_kernel<<<a->nmoves, 256, 0, s>>>(a->moves, b->jobs, a->xs, a->ubar);
    }
    // 4. eigen-decomposition
    if((rc = launch_gemm(ctx, 1, 0, a->t_s, a->n_s, a->g_s, b->mat, a->ubar, a->sk)))
        return rc;
    kmat_kernel<<<b->nblocks, 256, 0, s>>>(b->eigh_jobs, b->eval, a->ebar, a->sk);
    if((rc = launch_gemm(ctx, 0, 0, a->t_v, a->n_v, a->g_v, b->mat, a->sk, a->ubar)))
        return rc;
    // 5. entries
    if(b->nnz)
        entry_grad_kernel<<<grid_for(b->nnz * 64), 256, 0, s>>>(b->nnz, b->h_pos, a->pos_block, b->eigh_jobs, a->ubar, b->mat,
                                                                a->dval);
    POMX_CHECK_LAUNCH(ctx);
    if(b->nnz)
        POMX_HIP(ctx, hipMemcpyAsync(h_dval, a->dval, sizeof(double) * size_t(b->nnz), hipMemcpyDeviceToHost, s));
    POMX_HIP(ctx, hipStreamSynchronize(s));
    return POMX_OK;
}
