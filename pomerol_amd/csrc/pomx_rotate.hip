// K3: rotation of a monomial operator into the eigenbasis, batched over connected block pairs.
//
// Replaces MonomialOperatorPart::computeImpl (src/pomerol/MonomialOperatorPart.cpp:48-92):
//   OURight = M * U_R   (M has at most one +-1 per column: a signed row gather)
//   O       = U_L^T * OURight,  then sparseView(Tol) (entries |x| <= Tol*1e-12 dropped).
// Rows t of OURight with no source state are zero and add nothing to O, so the product runs
// over the mapped rows only: O = sum_q U_L[t_q, :]^T (+-U_R[s_q, :]), q over the (t, +-(s+1))
// pairs of the gather map with a source, kept in ascending t (a compaction pass per job).
// Both operands are then row gathers, fused into the tile loads of an FP64 MFMA GEMM
// (v_mfma_f64_16x16x4f64): 128x128 output tile per workgroup, eight waves of 32x64, K staged
// through LDS 16 pairs at a time, the next slab's rows loaded while the current one is
// multiplied.
#include "pomx_internal.hpp"

#include <algorithm>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// One workgroup per job: the pairs (t, src[t]) with src[t] != 0, in ascending t, to
// kmap[src_off ..]; their number to kcnt[job].
__global__ void __launch_bounds__(256) rot_compact_kernel(const pomx_rot_job* __restrict__ jobs, const int* __restrict__ src,
                                                          int2* __restrict__ kmap, int* __restrict__ kcnt) {
    __shared__ int wsum[4];
    pomx_rot_job const J = jobs[blockIdx.x];
    const int* __restrict__ sr = src + J.src_off;
    int2* __restrict__ km = kmap + J.src_off;
    int const tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    int base = 0;
    for(int c0 = 0; c0 < J.n_to; c0 += 256) {
        int const t = c0 + tid;
        int const sv = t < J.n_to ? sr[t] : 0;
        unsigned long long const m = __ballot(sv != 0);
        int const before = __popcll(m & ((1ull << lane) - 1ull));
        __syncthreads(); // wsum of the previous chunk consumed
        if(lane == 0)
            wsum[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for(int w = 0; w < wave; ++w)
            off += wsum[w];
        if(sv != 0)
            km[off + before] = int2{t, sv};
        base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    }
    if(tid == 0)
        kcnt[blockIdx.x] = base;
}

// 128 x 128 output tile per workgroup of 8 waves (4 x 2, 32 x 64 each): against 64 x 64 tiles
// of 4 waves it halves the operand rows each multiply-add reads from L2 (measured 60.6 against
// 48.5 TFLOP/s on a 15876 x 15876 pair), at 112 registers, four waves per SIMD.
constexpr int kRotTile = 128;
constexpr int kRotWgm = 4; // wave rows of the workgroup (two wave columns)

// Output tile TT x TT of one job per workgroup of WGM x 2 waves, each wave (TT/WGM) x (TT/2)
// as 16 x 16 accumulators of v_mfma_f64_16x16x4f64.
template <int TT, int WGM>
__global__ void __launch_bounds__(128 * WGM) rotate_kernel(const RotTile* __restrict__ tiles, const pomx_rot_job* __restrict__ jobs,
                                                           const double* __restrict__ U, const int2* __restrict__ kmap,
                                                           const int* __restrict__ kcnt, double* __restrict__ out, double cutoff) {
    constexpr int NT = 128 * WGM;                 // threads
    constexpr int WM = TT / WGM, WN = TT / 2;     // wave tile
    constexpr int NBM = WM / 16, NBN = WN / 16;   // 16 x 16 blocks per wave
    constexpr int RS = NT / TT, NR = 16 / RS;     // slab row stride of the loads, rows per thread
    static_assert(TT >= 64 && NT % TT == 0 && 16 % RS == 0 && WM % 16 == 0, "tile shape");
    __shared__ double As[1][16][TT + 4];
    __shared__ double Bs[1][16][TT + 4];
    RotTile const T = tiles[blockIdx.x];
    if(T.job < 0) // padding of the tile list (uniform per workgroup)
        return;
    pomx_rot_job const J = jobs[T.job];
    int const M = J.n_to, N = J.n_from, K = kcnt[T.job];
    const double* __restrict__ Ut = U + J.u_to_off;
    const double* __restrict__ Uf = U + J.u_from_off;
    const int2* __restrict__ km = kmap + J.src_off;
    int const tid = threadIdx.x, lane = tid & 63;
    int const wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
    int const col = tid % TT, row0 = __builtin_amdgcn_readfirstlane(tid / TT);
    int const i = T.i0 + col, j = T.j0 + col;
    bool const iok = i < M, jok = j < N;
    // column read by this lane (clamped into the block: the loads need no branch)
    int const ic = iok ? i : M - 1, jc = jok ? j : N - 1;
    // slab rows of this thread: row0 + RS r (r < NR), uniform per wave; one lane per tile
    // column.  kp: map entries of the slab after the one in flight; ra/rb: the slab in flight,
    // unsigned; sg: its signs (0 for rows past the end of the map).  The sign and the masks are
    // applied at the LDS store, so no arithmetic waits on the loads before the products of the
    // current slab.
    int2 kp[NR];
    int sg[NR];
    double ra[NR], rb[NR];
    auto load_map = [&](int t0) {
#pragma unroll
        for(int r = 0; r < NR; ++r) {
            int const q = t0 + row0 + RS * r;
            kp[r] = q < K ? km[q] : int2{0, 0};
        }
    };
    auto load_vals = [&]() {
#pragma unroll
        for(int r = 0; r < NR; ++r) {
            int const sv = kp[r].y;
            int const s = sv > 0 ? sv - 1 : (sv < 0 ? -sv - 1 : 0);
            sg[r] = sv;
            ra[r] = Ut[int64_t(kp[r].x) * M + ic];
            rb[r] = Uf[int64_t(s) * N + jc];
        }
    };
    dbl4 acc[NBM][NBN];
#pragma unroll
    for(int a = 0; a < NBM; ++a)
#pragma unroll
        for(int b = 0; b < NBN; ++b)
            acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
    auto store = [&](int buf) {
#pragma unroll
        for(int r = 0; r < NR; ++r) {
            As[buf][row0 + RS * r][col] = (sg[r] != 0 && iok) ? ra[r] : 0.0;
            Bs[buf][row0 + RS * r][col] = (sg[r] != 0 && jok) ? (sg[r] > 0 ? rb[r] : -rb[r]) : 0.0;
        }
    };
    auto mult = [&](int buf) {
#pragma unroll
        for(int kq = 0; kq < 4; ++kq) {
            int const kr = kq * 4 + (lane >> 4);
            double av[NBM], bv[NBN];
#pragma unroll
            for(int m = 0; m < NBM; ++m)
                av[m] = As[buf][kr][wm + 16 * m + (lane & 15)];
#pragma unroll
            for(int n = 0; n < NBN; ++n)
                bv[n] = Bs[buf][kr][wn + 16 * n + (lane & 15)];
#pragma unroll
            for(int m = 0; m < NBM; ++m)
#pragma unroll
                for(int n = 0; n < NBN; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bv[n], acc[m][n], 0, 0, 0);
        }
    };
    load_map(0);
    load_vals();
    load_map(16);
    for(int t0 = 0; t0 < K; t0 += 16) {
        __syncthreads();
        store(0);
        __syncthreads();
        if(t0 + 16 < K) {
            load_vals();
            load_map(t0 + 32);
        }
        mult(0);
    }
    double* __restrict__ O = out + J.out_off;
#pragma unroll
    for(int m = 0; m < NBM; ++m)
#pragma unroll
        for(int n = 0; n < NBN; ++n)
#pragma unroll
            for(int r = 0; r < 4; ++r) {
                // v_mfma_f64_16x16x4f64 result layout: D[i][j] in lane j + 16*(i%4), register i/4
                int const oi = T.i0 + wm + m * 16 + (lane >> 4) + 4 * r;
                int const oj = T.j0 + wn + n * 16 + (lane & 15);
                if(oi < M && oj < N) {
                    double const x = acc[m][n][r];
                    O[int64_t(oi) * N + oj] = fabs(x) <= cutoff ? 0.0 : x;
                }
            }
}

} // namespace


void rot_tiles(const pomx_rot_job* jobs, int njobs, std::vector<RotTile>& tiles) {
    // Walk of every job's output tiles in groups of kGroup tile rows, column by column inside a
    // group: workgroups that run at the same time share the rows of U_L and of U_R they read.
    constexpr int kGroup = 8;
    std::vector<RotTile> walk;
    constexpr int TT = kRotTile;
    for(int q = 0; q < njobs; ++q) {
        int const tm = (jobs[q].n_to + TT - 1) / TT, tn = (jobs[q].n_from + TT - 1) / TT;
        for(int g0 = 0; g0 < tm; g0 += kGroup)
            for(int c = 0; c < tn; ++c)
                for(int a = g0; a < std::min(tm, g0 + kGroup); ++a)
                    walk.push_back(RotTile{q, TT * a, TT * c, 0});
    }
    // Workgroup p of a launch runs on XCD p % 8, each XCD with its own L2: XCD x gets the
    // contiguous slice x of the walk, in order (the list is padded to a multiple of 8 with
    // tiles of job -1, which return at once).
    constexpr int kXcd = 8;
    size_t const per = (walk.size() + kXcd - 1) / kXcd;
    tiles.assign(per * kXcd, RotTile{-1, 0, 0, 0});
    for(size_t r = 0; r < walk.size(); ++r)
        tiles[(r % per) * kXcd + r / per] = walk[r];
}

int launch_rot_compact(pomx_ctx* ctx, const pomx_rot_job* d_jobs, int njobs, const int* d_src, int2* d_kmap, int* d_kcnt) {
    if(njobs <= 0)
        return POMX_OK;
    rot_compact_kernel<<<unsigned(njobs), 256, 0, ctx->stream>>>(d_jobs, d_src, d_kmap, d_kcnt);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

int launch_rotate(pomx_ctx* ctx, const RotTile* d_tiles, int ntiles, const pomx_rot_job* d_jobs, const double* d_U,
                  const int2* d_kmap, const int* d_kcnt, double* d_out, double cutoff) {
    if(ntiles <= 0)
        return POMX_OK;
    rotate_kernel<kRotTile, kRotWgm><<<unsigned(ntiles), 128 * kRotWgm, 0, ctx->stream>>>(d_tiles, d_jobs, d_U, d_kmap, d_kcnt, d_out, cutoff);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

extern "C" int pomx_rotate_batched(pomx_ctx* ctx, int njobs, const pomx_rot_job* h_jobs, const double* d_U,
                                   const int* d_src, double* d_out, double cutoff) {
    POMX_REQUIRE(ctx, njobs >= 0 && (njobs == 0 || (h_jobs && d_U && d_src && d_out)), "pomx_rotate_batched: bad arguments");
    if(njobs == 0)
        return POMX_OK;
    int64_t src_len = 0;
    for(int q = 0; q < njobs; ++q) {
        POMX_REQUIRE(ctx, h_jobs[q].n_to >= 1 && h_jobs[q].n_from >= 1 && h_jobs[q].src_off >= 0,
                     "pomx_rotate_batched: empty block or negative gather-map offset");
        src_len = std::max<int64_t>(src_len, h_jobs[q].src_off + h_jobs[q].n_to);
    }
    std::vector<RotTile> tiles;
    rot_tiles(h_jobs, njobs, tiles);
    WsBuf<RotTile> d_tiles(ctx, tiles.size());
    WsBuf<pomx_rot_job> d_jobs(ctx, njobs);
    WsBuf<int2> d_kmap(ctx, size_t(src_len));
    WsBuf<int> d_kcnt(ctx, size_t(njobs));
    if(!d_tiles.p || !d_jobs.p || !d_kmap.p || !d_kcnt.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_rotate_batched: out of device memory");
    POMX_HIP(ctx, hipMemcpyAsync(d_tiles.p, tiles.data(), sizeof(RotTile) * tiles.size(), hipMemcpyHostToDevice, ctx->stream));
    POMX_HIP(ctx, h2d_sync(ctx, d_jobs.p, h_jobs, sizeof(pomx_rot_job) * njobs)); // (both copies done)
    int rc = launch_rot_compact(ctx, d_jobs, njobs, d_src, d_kmap, d_kcnt);
    if(!rc)
        rc = launch_rotate(ctx, d_tiles, int(tiles.size()), d_jobs, d_U, d_kmap, d_kcnt, d_out, cutoff);
    if(rc) {
        (void)hipStreamSynchronize(ctx->stream); // launches already queued read the pooled buffers
        return rc;
    }
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}
