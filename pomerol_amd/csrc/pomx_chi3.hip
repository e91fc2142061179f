// K12 (three-point susceptibility term generation and merge) and K13 (Lehmann sum on a list of
// frequency pairs) for one ThreePointSusceptibilityPart.
//
// Reference: ThreePointSusceptibilityPart::computeImpl and addMultiterm
// (ThreePointSusceptibilityPart.cpp:243-348), the NonResonantFF/NonResonantFB/Resonant TermLists
// (ThreePointSusceptibilityPart.hpp:66-370, .cpp:34-149) and their evaluation
// (ThreePointSusceptibilityPart.hpp:527-540, .cpp:350-359).
//
// A part is <1|Fa|2><2|Fb|3><3|B|1> with B = B1 B2 over the blocks (S1, S2, S3).  Pipeline (all
// on the device, one part per call):
//  1. contiguous copies: Bt[i1][i3] = B(i3, i1) (LDS-tiled product of the two operator views) and
//     Fbc[i2][i3] = Fb(i2, i3), so the lanes of a wave running over i3 read coalesced rows;
//  2. key fields: the pole bins round(P/res) of E1-E2, E2-E3 and E1-E3 themselves when they fit
//     31 bits (every |P|/res < 2^30), else their dense ranks (rank tables r12, r23, r13), so a
//     term's TermList identity becomes one 64-bit key kind (2 bits) | field (31) | field (31);
//  3. one wave per (i1, i2) row with the lanes over i3: count the terms of every row, prefix-sum,
//     emit (key, packed indices) in the reference's (i1, i2, i3) visiting order, a triple's own
//     terms in the order addMultiterm adds them; the packed value (i1, i2, i3, slot) is
//     numerically ordered like that visiting order;
//  4. stable radix sort by key, one thread per key group: coefficients summed in emission order,
//     poles averaged over the group's members, live groups compacted into kind-segmented SoA.
//     TermList erases a merged term whose coefficient is <= tol/(size+1), size+1 being the
//     number of terms in the list at that moment: the groups of the same kind whose first
//     member precedes the merge in visiting order (binary search in their sorted first-member
//     positions) less those erased and not restarted before it.  The erase/restart events are
//     found by iterating the reduce to a fixed point: a pass uses the previous pass's events
//     (none in the first) and the iteration stops when a pass reproduces them; by induction over
//     the visiting order every decision of that pass is then the sequential one;
//  5. evaluation: fermion-fermion and fermion-boson terms at every pair (terms staged through LDS,
//     several pairs per thread, per-slice partial sums reduced in a fixed order); resonant terms
//     only at the pairs with |z1 - xi z2| <= res, compacted and summed by the pole-sum kernel.
#include "pomx_internal.hpp"

#include <algorithm>
#include <cmath>

namespace {

struct Chi3Dev {
    int n1, n2, n3;
    pomx_view Fa;       // <1|Fa|2>, n1 x n2
    const double* Fb;   // <2|Fb|3>, n2 x n3 contiguous
    const double* Bt;   // Bt[i1 * n3 + i3] = <3|B1 B2|1>
    const double *E1, *E2, *E3, *w1, *w2, *w3;
    double prefactor, beta, tol, res;
    int pp, sw;
};

__device__ inline double vmat(pomx_view const& v, int r, int c) { return v.base[r * v.rs + c * v.cs]; }

// ---- 1. contiguous copies -------------------------------------------------------------------
// Bt[i1][i3] = sum_m B1(i3, m) B2(m, i1), 16 x 16 output tiles, m ascending.
__global__ void __launch_bounds__(256) chi3_bt_kernel(int n1, int n3, int n4, pomx_view B1, pomx_view B2,
                                                      double* __restrict__ Bt) {
    __shared__ double sX[16][17], sY[16][17];
    int const tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    int const r0 = blockIdx.y * 16, c0 = blockIdx.x * 16;
    double acc = 0.0;
    for(int m0 = 0; m0 < n4; m0 += 16) {
        int const m = m0 + ty;
        sX[ty][tx] = (m < n4 && r0 + tx < n1) ? vmat(B2, m, r0 + tx) : 0.0; // B2(m, i1)
        sY[ty][tx] = (m < n4 && c0 + tx < n3) ? vmat(B1, c0 + tx, m) : 0.0; // B1(i3, m)
        __syncthreads();
#pragma unroll
        for(int k = 0; k < 16; ++k)
            acc = fma(sX[k][ty], sY[k][tx], acc);
        __syncthreads();
    }
    if(r0 + ty < n1 && c0 + tx < n3)
        Bt[int64_t(r0 + ty) * n3 + c0 + tx] = acc;
}

__global__ void chi3_copy_kernel(int nr, int nc, pomx_view V, double* __restrict__ out) {
    int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(g < int64_t(nr) * nc)
        out[g] = vmat(V, int(g / nc), int(g % nc));
}

// ---- 3. counting / emission -----------------------------------------------------------------
// The matrix element (Fa Fb) B * prefactor of ThreePointSusceptibilityPart.cpp:272-277.
__device__ inline double chi3_M(Chi3Dev const& D, double fa, int i1, int i2, int i3) {
    return fa * D.Fb[int64_t(i2) * D.n3 + i3] * D.Bt[int64_t(i1) * D.n3 + i3] * D.prefactor;
}

// Terms addMultiterm adds for one triple (ThreePointSusceptibilityPart.cpp:298-348), in its
// order: bit 0 fermion-fermion, bit 1 fermion-boson, bit 2 resonant, bit 3 the second
// fermion-fermion term of an unswapped part.
__device__ inline unsigned chi3_flags(Chi3Dev const& D, double M, int i1, int i2, int i3) {
    double const Wi = D.w1[i1], Wj = D.w2[i2], Wk = D.w3[i3];
    unsigned f = 0;
    if(fabs(M * (Wi + Wj)) > D.tol)
        f |= 1u;
    if(fabs(D.E1[i1] - D.E3[i3]) <= D.res) {
        if(fabs(-M * D.beta * Wi) > D.tol)
            f |= 4u;
    } else if(fabs(M * (Wk - Wi)) > D.tol) {
        f |= D.sw ? 2u : 10u;
    }
    return f;
}

__global__ void __launch_bounds__(256) chi3_count_kernel(Chi3Dev D, int64_t nrows, int* __restrict__ cnt) {
    int const lane = threadIdx.x & 63;
    for(int64_t row = blockIdx.x * 4 + (threadIdx.x >> 6); row < nrows; row += int64_t(gridDim.x) * 4) {
        int const i1 = int(row / D.n2), i2 = int(row % D.n2);
        double const fa = vmat(D.Fa, i1, i2);
        int c = 0;
        if(fa != 0.0)
            for(int i3 = lane; i3 < D.n3; i3 += 64)
                c += __popc(chi3_flags(D, chi3_M(D, fa, i1, i2, i3), i1, i2, i3));
#pragma unroll
        for(int o = 32; o > 0; o >>= 1)
            c += __shfl_xor(c, o);
        if(lane == 0)
            cnt[row] = c;
    }
}

__global__ void chi3_widen_kernel(int64_t n, const int* __restrict__ in, int64_t* __restrict__ out) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n)
        out[k] = in[k];
}

constexpr uint64_t kRankMask = (uint64_t(1) << 31) - 1;
constexpr int64_t kDirectOff = int64_t(1) << 30;

// the pole bin round(P/res) itself as a 31-bit key field (when every |P|/res < 2^30)
__device__ inline uint64_t chi3_bin(double P, double res) { return uint64_t(int64_t(round(P / res)) + kDirectOff); }

// direct: key fields are the bins themselves (r12, r23, r13 unused); otherwise their dense ranks
__global__ void __launch_bounds__(256) chi3_emit_kernel(Chi3Dev D, int64_t nrows, int64_t ntot, const int64_t* __restrict__ off,
                                                        int direct, const uint32_t* __restrict__ r12,
                                                        const uint32_t* __restrict__ r23, const uint32_t* __restrict__ r13,
                                                        uint64_t* __restrict__ keys, uint64_t* __restrict__ vals) {
    int const lane = threadIdx.x & 63;
    for(int64_t row = blockIdx.x * 4 + (threadIdx.x >> 6); row < nrows; row += int64_t(gridDim.x) * 4) {
        int const i1 = int(row / D.n2), i2 = int(row % D.n2);
        double const fa = vmat(D.Fa, i1, i2);
        if(fa == 0.0)
            continue; // wave-uniform
        int64_t base = off[row];
        int64_t const end = (row + 1 < nrows) ? off[row + 1] : ntot; // never write past this row's slots
        uint64_t const a12 = direct ? chi3_bin(D.E1[i1] - D.E2[i2], D.res) : uint64_t(r12[row]);
        for(int c0 = 0; c0 < D.n3; c0 += 64) {
            int const i3 = c0 + lane;
            unsigned f = 0;
            if(i3 < D.n3)
                f = chi3_flags(D, chi3_M(D, fa, i1, i2, i3), i1, i2, i3);
            int const c = __popc(f);
            int incl = c;
#pragma unroll
            for(int o = 1; o < 64; o <<= 1) {
                int const t = __shfl_up(incl, o);
                if(lane >= o)
                    incl += t;
            }
            int const total = __shfl(incl, 63);
            if(f && base + incl <= end) {
                int64_t pos = base + incl - c;
                uint64_t const a23 = direct ? chi3_bin(D.E2[i2] - D.E3[i3], D.res) : uint64_t(r23[int64_t(i2) * D.n3 + i3]);
                uint64_t const a13 = direct ? chi3_bin(D.E1[i1] - D.E3[i3], D.res) : uint64_t(r13[int64_t(i1) * D.n3 + i3]);
                uint64_t const v = (uint64_t(i1) << 40) | (uint64_t(i2) << 24) | (uint64_t(i3) << 8);
                uint64_t const kff = ((a12 & kRankMask) << 31) | (a23 & kRankMask);
                if(f & 1u) {
                    keys[pos] = kff;
                    vals[pos++] = v;
                }
                if(f & 2u) {
                    keys[pos] = (uint64_t(1) << 62) | (((D.sw ? a23 : a12) & kRankMask) << 31) | (a13 & kRankMask);
                    vals[pos++] = v | 1u;
                }
                if(f & 4u) {
                    keys[pos] = (uint64_t(2) << 62) | (a23 & kRankMask);
                    vals[pos++] = v | 1u;
                }
                if(f & 8u) {
                    keys[pos] = kff;
                    vals[pos] = v | 2u;
                }
            }
            base += total;
        }
    }
}

// out = (min E1, max E1, min E2, max E2, min E3, max E3), one block of 256 threads
__global__ void __launch_bounds__(256) chi3_range_kernel(int n1, int n2, int n3, const double* __restrict__ E1,
                                                         const double* __restrict__ E2, const double* __restrict__ E3,
                                                         double* __restrict__ out) {
    __shared__ double lo[3][256], hi[3][256];
    int const t = threadIdx.x;
    const double* E[3] = {E1, E2, E3};
    int const n[3] = {n1, n2, n3};
    for(int v = 0; v < 3; ++v) {
        double a = E[v][0], b = E[v][0];
        for(int k = t; k < n[v]; k += 256) {
            a = fmin(a, E[v][k]);
            b = fmax(b, E[v][k]);
        }
        lo[v][t] = a;
        hi[v][t] = b;
    }
    __syncthreads();
    for(int w = 128; w > 0; w >>= 1) {
        if(t < w)
            for(int v = 0; v < 3; ++v) {
                lo[v][t] = fmin(lo[v][t], lo[v][t + w]);
                hi[v][t] = fmax(hi[v][t], hi[v][t + w]);
            }
        __syncthreads();
    }
    if(t < 3) {
        out[2 * t] = lo[t][0];
        out[2 * t + 1] = hi[t][0];
    }
}

// ---- 4. merge -------------------------------------------------------------------------------
__global__ void chi3_heads_kernel(int64_t n, const uint64_t* __restrict__ keys, int64_t* __restrict__ head) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n)
        head[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}
__global__ void chi3_group_start_kernel(int64_t n, const int64_t* __restrict__ head, const int64_t* __restrict__ gid,
                                        int64_t* __restrict__ start) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n && head[k])
        start[gid[k]] = k;
}

// (kind, visiting position of the group's first member) per group, for the erase divisor.
constexpr uint64_t kKindMask = uint64_t(3) << 62;
__global__ void chi3_first_kernel(int64_t ngroups, const int64_t* __restrict__ start, const uint64_t* __restrict__ keys,
                                  const uint64_t* __restrict__ vals, uint64_t* __restrict__ fk, uint64_t* __restrict__ gi_out) {
    int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(g < ngroups) {
        fk[g] = (keys[start[g]] & kKindMask) | vals[start[g]];
        gi_out[g] = uint64_t(g);
    }
}

// One thread per key group.  Coefficient and poles of each member as addMultiterm forms them:
//   FF:  c = +-M (Wi + Wj), poles (Eij, -+Ejk) or, swapped, (Ejk, -+Eij);
//   FF2: c = +-M (Wk - Wi), poles (Eij, -+Ejk);
//   FB:  c = +-M (Wk - Wi), poles (Eij or Ejk, Eik);
//   R:   c = +-(-M beta Wi), pole -Ejk or Ejk.
// g holds (C, mean P0, mean P1) per group.  ev_key/ev_cum: the assumed erase (+1) and restart
// (-1) events, sorted by (kind, position), with inclusive prefix sums of their weights.
// EMIT = false: evn[gi] = number of events of the group; EMIT = true: the events are written
// at the offsets evn[gi].
template <bool EMIT>
__global__ void chi3_reduce_kernel(Chi3Dev D, int64_t ngroups, int64_t n, const int64_t* __restrict__ start,
                                   const uint64_t* __restrict__ keys, const uint64_t* __restrict__ vals, EraseRule R,
                                   double* __restrict__ g, int64_t* __restrict__ live, int64_t* __restrict__ evn,
                                   uint64_t* __restrict__ ev_key_out, uint64_t* __restrict__ ev_w_out) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups)
        return;
    int64_t const s = start[gi], e = (gi + 1 < ngroups) ? start[gi + 1] : n;
    int const kind = int(keys[s] >> 62);
    uint64_t const kbits = keys[s] & kKindMask; // each kind is its own TermList
    EraseCursor const cur = erase_cursor(R, kbits);
    double const pm = D.pp ? 1.0 : -1.0;
    bool alive = false;
    double c = 0, p0 = 0, p1 = 0;
    long cnt = 0;
    int64_t nevents = 0, at = EMIT ? evn[gi] : 0;
    for(int64_t k = s; k < e; ++k) {
        uint64_t const v = vals[k];
        int const i1 = int((v >> 40) & 0xffff), i2 = int((v >> 24) & 0xffff), i3 = int((v >> 8) & 0xffff);
        bool const second = (v & 0xff) == 2;
        double const M = chi3_M(D, vmat(D.Fa, i1, i2), i1, i2, i3);
        double const Ei = D.E1[i1], Ej = D.E2[i2], Ek = D.E3[i3];
        double const Wi = D.w1[i1], Wj = D.w2[i2], Wk = D.w3[i3];
        double const Eij = Ei - Ej, Ejk = Ej - Ek, Eik = Ei - Ek;
        double a, P0, P1 = 0.0;
        if(kind == 0) {
            if(!second) {
                a = M * (Wi + Wj);
                P0 = D.sw ? Ejk : Eij;
                P1 = (D.sw ? Eij : Ejk) * pm;
            } else {
                a = M * (Wk - Wi);
                P0 = Eij;
                P1 = D.pp ? Ejk : -Ejk;
            }
            if(D.pp)
                a = -a;
        } else if(kind == 1) {
            a = M * (Wk - Wi);
            if(D.sw)
                a = -a;
            P0 = D.sw ? Ejk : Eij;
            P1 = Eik;
        } else {
            a = -M * D.beta * Wi;
            if(D.sw)
                a = -a;
            P0 = D.sw ? Ejk : -Ejk;
        }
        if(!alive) {
            if(k > s) { // restart of an erased group
                if(EMIT) {
                    ev_key_out[at] = kbits | v;
                    ev_w_out[at++] = uint64_t(int64_t(-1));
                }
                ++nevents;
            }
            alive = true;
            c = a;
            p0 = P0;
            p1 = P1;
            cnt = 1;
        } else {
            c += a;
            p0 += P0;
            p1 += P1;
            ++cnt;
            uint64_t const q = kbits | v;
            if(fabs(c) <= erase_threshold(R, cur, q)) {
                alive = false;
                if(EMIT) {
                    ev_key_out[at] = q;
                    ev_w_out[at++] = 1;
                }
                ++nevents;
            }
        }
    }
    if(!EMIT) {
        double const inv = 1.0 / double(cnt);
        g[gi * 3 + 0] = c;
        g[gi * 3 + 1] = p0 * inv;
        g[gi * 3 + 2] = p1 * inv;
        live[gi] = alive ? 1 : 0;
        evn[gi] = nevents;
    }
}

// kinfo[0..2]: live index of the first group of each kind; kinfo[3..5]: live groups per kind.
// Groups are sorted by key and the kind sits in the top bits: binary-search the kind boundaries.
__global__ void chi3_kinds_kernel(int64_t ngroups, const int64_t* __restrict__ start, const uint64_t* __restrict__ keys,
                                  const int64_t* __restrict__ live, const int64_t* __restrict__ lpos,
                                  int64_t* __restrict__ kinfo) {
    if(threadIdx.x != 0 || blockIdx.x != 0)
        return;
    int64_t const total = lpos[ngroups - 1] + live[ngroups - 1];
    int64_t first[4];
    for(int kind = 0; kind < 3; ++kind) {
        int64_t lo = 0, hi = ngroups;
        while(lo < hi) {
            int64_t mid = (lo + hi) / 2;
            if(int(keys[start[mid]] >> 62) < kind)
                lo = mid + 1;
            else
                hi = mid;
        }
        first[kind] = lo < ngroups ? lpos[lo] : total;
    }
    first[3] = total;
    for(int kind = 0; kind < 3; ++kind) {
        kinfo[kind] = first[kind];
        kinfo[3 + kind] = first[kind + 1] - first[kind];
    }
}

// Live groups into [FF: C, P0, P1][FB: C, P1, P12][R: C, P], each field count doubles long.
__global__ void chi3_compact_kernel(int64_t ngroups, const int64_t* __restrict__ live, const int64_t* __restrict__ lpos,
                                    const double* __restrict__ g, const int64_t* __restrict__ start,
                                    const uint64_t* __restrict__ keys, const int64_t* __restrict__ kinfo,
                                    double* __restrict__ out) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups || !live[gi])
        return;
    int const kind = int(keys[start[gi]] >> 62);
    int64_t const at = lpos[gi] - kinfo[kind];
    int64_t const c0 = kinfo[3], c1 = kinfo[4], c2 = kinfo[5];
    if(kind == 0) {
        out[at] = g[gi * 3];
        out[c0 + at] = g[gi * 3 + 1];
        out[2 * c0 + at] = g[gi * 3 + 2];
    } else if(kind == 1) {
        double* b = out + 3 * c0;
        b[at] = g[gi * 3];
        b[c1 + at] = g[gi * 3 + 1];
        b[2 * c1 + at] = g[gi * 3 + 2];
    } else {
        double* b = out + 3 * c0 + 3 * c1;
        b[at] = g[gi * 3];
        b[c2 + at] = g[gi * 3 + 1];
    }
}

// ---- 5. evaluation --------------------------------------------------------------------------
constexpr int kPF = 4;       // frequency pairs per thread
constexpr int kStage = 256;  // terms staged through LDS at a time

__device__ inline double rcp_refined(double x) {
    // v_rcp_f64 estimate refined by one Newton-Raphson step r' = r + r (1 - x r)
    double const r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// KIND 0: C / ((z1 - P0)(z2 - P1))          (ThreePointSusceptibilityPart.hpp:527-529)
// KIND 1: C / ((z1 - P1)(z1 - xi z2 - P12))  (ThreePointSusceptibilityPart.hpp:531-533)
// Pairs f = blockIdx.x * 256 kPF + u * 256 + threadIdx.x; partial[slice][f] accumulates.
template <int KIND>
__global__ void __launch_bounds__(256) chi3_eval_kernel(int64_t nterms, int64_t slice_len, const double* __restrict__ T,
                                                        int nfreq, const double* __restrict__ fz, double xi,
                                                        double* __restrict__ partial) {
    __shared__ double sC[kStage], sA[kStage], sB[kStage];
    int64_t const t0 = int64_t(blockIdx.y) * slice_len;
    int64_t const t1 = min(nterms, t0 + slice_len);
    int const f0 = blockIdx.x * (256 * kPF) + threadIdx.x;
    double x1[kPF], y1[kPF], x2[kPF], y2[kPF], ar[kPF], ai[kPF];
#pragma unroll
    for(int u = 0; u < kPF; ++u) {
        int const f = min(f0 + u * 256, nfreq - 1);
        x1[u] = fz[4 * f];
        y1[u] = fz[4 * f + 1];
        if(KIND == 0) {
            x2[u] = fz[4 * f + 2];
            y2[u] = fz[4 * f + 3];
        } else { // z1 - xi z2
            x2[u] = x1[u] - xi * fz[4 * f + 2];
            y2[u] = y1[u] - xi * fz[4 * f + 3];
        }
        ar[u] = 0.0;
        ai[u] = 0.0;
    }
    for(int64_t c0 = t0; c0 < t1; c0 += kStage) {
        int const cnt = int(min<int64_t>(kStage, t1 - c0));
        __syncthreads();
        if(threadIdx.x < cnt) {
            sC[threadIdx.x] = T[c0 + threadIdx.x];
            sA[threadIdx.x] = T[nterms + c0 + threadIdx.x];
            sB[threadIdx.x] = T[2 * nterms + c0 + threadIdx.x];
        }
        __syncthreads();
#pragma unroll 2
        for(int k = 0; k < cnt; ++k) {
            double const C = sC[k], A = sA[k], B = sB[k];
#pragma unroll
            for(int u = 0; u < kPF; ++u) {
                double const a1 = x1[u] - A, a2 = x2[u] - B;
                double const Dr = fma(a1, a2, -y1[u] * y2[u]);
                double const Di = fma(a1, y2[u], y1[u] * a2);
                double const r = C * rcp_refined(fma(Dr, Dr, Di * Di));
                ar[u] = fma(Dr, r, ar[u]);
                ai[u] = fma(-Di, r, ai[u]);
            }
        }
    }
    double* p = partial + size_t(blockIdx.y) * nfreq * 2;
#pragma unroll
    for(int u = 0; u < kPF; ++u) {
        int const f = f0 + u * 256;
        if(f < nfreq) {
            p[2 * f] += ar[u];
            p[2 * f + 1] += ai[u];
        }
    }
}

// Every z purely imaginary (z1 = i y1, z2 = i y2; the Matsubara grids): with u = y2 (KIND 0) or
// u = y1 - xi y2 (KIND 1), and A, B the term's two poles (P0, P1 or P1, P12), the denominator is
// D = (i y1 - A)(i u - B) = (A B - y1 u) - i (A u + B y1), so a pair costs D, |D|^2, the two
// accumulating FMAs and a quarter of a reciprocal shared by four pairs
// (1/q_h = (product of the other three) / (q0 q1 q2 q3)); |D|^2 stays within ~1e-20..1e60 for
// Matsubara frequencies, so the product of four is far inside the double range.
constexpr int kPI = 8; // pairs per thread
template <int KIND>
__global__ void __launch_bounds__(256) chi3_eval_imag_kernel(int64_t nterms, int64_t slice_len,
                                                             const double* __restrict__ T, int nfreq,
                                                             const double* __restrict__ fz, double xi,
                                                             double* __restrict__ partial) {
    __shared__ double sC[kStage], sA[kStage], sB[kStage], sAB[kStage];
    int64_t const t0 = int64_t(blockIdx.y) * slice_len;
    int64_t const t1 = min(nterms, t0 + slice_len);
    int const f0 = blockIdx.x * (256 * kPI) + threadIdx.x;
    double y1[kPI], u[kPI], yu[kPI], ar[kPI], ai[kPI];
#pragma unroll
    for(int h = 0; h < kPI; ++h) {
        int const f = min(f0 + h * 256, nfreq - 1);
        y1[h] = fz[4 * f + 1];
        u[h] = KIND == 0 ? fz[4 * f + 3] : y1[h] - xi * fz[4 * f + 3];
        yu[h] = y1[h] * u[h];
        ar[h] = 0.0;
        ai[h] = 0.0;
    }
    for(int64_t c0 = t0; c0 < t1; c0 += kStage) {
        int const cnt = int(min<int64_t>(kStage, t1 - c0));
        __syncthreads();
        if(threadIdx.x < cnt) {
            double const A = T[nterms + c0 + threadIdx.x], B = T[2 * nterms + c0 + threadIdx.x];
            sC[threadIdx.x] = T[c0 + threadIdx.x];
            sA[threadIdx.x] = A;
            sB[threadIdx.x] = B;
            sAB[threadIdx.x] = A * B;
        }
        __syncthreads();
#pragma unroll 2
        for(int k = 0; k < cnt; ++k) {
            double const C = sC[k], A = sA[k], B = sB[k], AB = sAB[k];
#pragma unroll
            for(int h0 = 0; h0 < kPI; h0 += 4) {
                double Dr[4], Di[4], q[4];
#pragma unroll
                for(int h = 0; h < 4; ++h) {
                    Dr[h] = AB - yu[h0 + h];
                    Di[h] = fma(A, u[h0 + h], B * y1[h0 + h]); // -Im D
                    q[h] = fma(Dr[h], Dr[h], Di[h] * Di[h]);
                }
                double const q01 = q[0] * q[1], q23 = q[2] * q[3];
                double const r = C * rcp_refined(q01 * q23);
                double const r01 = q23 * r, r23 = q01 * r;
                double const sc[4] = {q[1] * r01, q[0] * r01, q[3] * r23, q[2] * r23};
#pragma unroll
                for(int h = 0; h < 4; ++h) { // C conj(D) / |D|^2 with conj(D) = Dr + i Di
                    ar[h0 + h] = fma(Dr[h], sc[h], ar[h0 + h]);
                    ai[h0 + h] = fma(Di[h], sc[h], ai[h0 + h]);
                }
            }
        }
    }
    double* p = partial + size_t(blockIdx.y) * nfreq * 2;
#pragma unroll
    for(int h = 0; h < kPI; ++h) {
        int const f = f0 + h * 256;
        if(f < nfreq) {
            p[2 * f] += ar[h];
            p[2 * f + 1] += ai[h];
        }
    }
}

// flag[0] = 1 if some pair has a non-zero real part
__global__ void chi3_any_real_kernel(int nfreq, const double* __restrict__ fz, int* __restrict__ flag) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f < nfreq && (fz[4 * f] != 0.0 || fz[4 * f + 2] != 0.0))
        flag[0] = 1;
}

__global__ void chi3_eval_reduce_kernel(int nfreq, int nslices, const double* __restrict__ partial,
                                        double* __restrict__ acc) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= nfreq)
        return;
    double sr = 0.0, si = 0.0;
    for(int s = 0; s < nslices; ++s) {
        sr += partial[(size_t(s) * nfreq + f) * 2];
        si += partial[(size_t(s) * nfreq + f) * 2 + 1];
    }
    acc[2 * f] += sr;
    acc[2 * f + 1] += si;
}

// 1 where |z1 - xi z2| <= res (ThreePointSusceptibilityPart.hpp:535-540)
__global__ void chi3_res_flags_kernel(int nfreq, const double* __restrict__ fz, double xi, double res,
                                      int64_t* __restrict__ flag) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f < nfreq)
        flag[f] = hypot(fz[4 * f] - xi * fz[4 * f + 2], fz[4 * f + 1] - xi * fz[4 * f + 3]) <= res ? 1 : 0;
}
__global__ void chi3_res_gather_kernel(int nfreq, const double* __restrict__ fz, const int64_t* __restrict__ flag,
                                       const int64_t* __restrict__ pos, double* __restrict__ z1, int* __restrict__ idx) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f < nfreq && flag[f]) {
        int64_t const k = pos[f];
        z1[2 * k] = fz[4 * f];
        z1[2 * k + 1] = fz[4 * f + 1];
        idx[k] = f;
    }
}
__global__ void chi3_res_scatter_kernel(int64_t nres, const int* __restrict__ idx, const double* __restrict__ v,
                                        double* __restrict__ acc) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < nres) {
        acc[2 * idx[k]] += v[2 * k];
        acc[2 * idx[k] + 1] += v[2 * k + 1];
    }
}

#define CHI3_WS(ctx, buf)                                                                            \
    do {                                                                                             \
        if(!(buf).p)                                                                                 \
            return pomx_fail(ctx, POMX_ERR_HIP, "three-point susceptibility: out of device memory"); \
    } while(0)

int chi3_eval(pomx_ctx* ctx, const double* data, int64_t const count[3], int xi, double res, int nfreq,
              const double* d_freq, double* d_acc) {
    int rc;
    int64_t const c0 = count[0], c1 = count[1], c2 = count[2];
    if(nfreq <= 0)
        return POMX_OK;
    if(c0 + c1 > 0) {
        WsBuf<int> anyre(ctx, 1);
        CHI3_WS(ctx, anyre);
        if((rc = pomx_fill_zero(ctx, anyre.p, sizeof(int))))
            return rc;
        chi3_any_real_kernel<<<grid_for(nfreq), kBlock, 0, ctx->stream>>>(nfreq, d_freq, anyre);
        POMX_CHECK_LAUNCH(ctx);
        int has_real = 1;
        POMX_HIP(ctx, hipMemcpyAsync(&has_real, anyre.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        int const per = has_real ? kPF : kPI; // pairs per thread
        int const ftiles = (nfreq + 256 * per - 1) / (256 * per);
        int64_t const want = std::max<int64_t>(1, 2048 / ftiles);
        int64_t const big = std::max(c0, c1);
        int64_t slice = std::max<int64_t>(kStage, (big + want - 1) / want);
        slice = (slice + kStage - 1) / kStage * kStage;
        int const nslices = int(std::max<int64_t>(1, (big + slice - 1) / slice));
        WsBuf<double> partial(ctx, size_t(nslices) * nfreq * 2);
        CHI3_WS(ctx, partial);
        if((rc = pomx_fill_zero(ctx, partial.p, sizeof(double) * size_t(nslices) * nfreq * 2)))
            return rc;
        dim3 const g0(ftiles, unsigned((c0 + slice - 1) / slice)), g1(ftiles, unsigned((c1 + slice - 1) / slice));
        if(c0 > 0) {
            if(has_real)
                chi3_eval_kernel<0><<<g0, 256, 0, ctx->stream>>>(c0, slice, data, nfreq, d_freq, double(xi), partial);
            else
                chi3_eval_imag_kernel<0><<<g0, 256, 0, ctx->stream>>>(c0, slice, data, nfreq, d_freq, double(xi), partial);
        }
        if(c1 > 0) {
            if(has_real)
                chi3_eval_kernel<1><<<g1, 256, 0, ctx->stream>>>(c1, slice, data + 3 * c0, nfreq, d_freq, double(xi),
                                                                 partial);
            else
                chi3_eval_imag_kernel<1><<<g1, 256, 0, ctx->stream>>>(c1, slice, data + 3 * c0, nfreq, d_freq,
                                                                      double(xi), partial);
        }
        chi3_eval_reduce_kernel<<<grid_for(nfreq), kBlock, 0, ctx->stream>>>(nfreq, nslices, partial, d_acc);
        POMX_CHECK_LAUNCH(ctx);
    }
    if(c2 > 0) {
        WsBuf<int64_t> flag(ctx, nfreq), pos(ctx, nfreq);
        CHI3_WS(ctx, flag);
        CHI3_WS(ctx, pos);
        chi3_res_flags_kernel<<<grid_for(nfreq), kBlock, 0, ctx->stream>>>(nfreq, d_freq, double(xi), res, flag);
        POMX_CHECK_LAUNCH(ctx);
        if((rc = pomx_exclusive_scan(ctx, nfreq, flag, pos)))
            return rc;
        int64_t lf = 0, lp = 0;
        if((rc = read_i64(ctx, flag.p + nfreq - 1, &lf)) || (rc = read_i64(ctx, pos.p + nfreq - 1, &lp)))
            return rc;
        int64_t const nres = lf + lp;
        if(nres > 0) {
            WsBuf<double> z1(ctx, size_t(nres) * 2), v(ctx, size_t(nres) * 2);
            WsBuf<int> idx(ctx, nres);
            CHI3_WS(ctx, z1);
            CHI3_WS(ctx, v);
            CHI3_WS(ctx, idx);
            chi3_res_gather_kernel<<<grid_for(nfreq), kBlock, 0, ctx->stream>>>(nfreq, d_freq, flag, pos, z1, idx);
            if((rc = pomx_fill_zero(ctx, v.p, sizeof(double) * size_t(nres) * 2)))
                return rc;
            const double* R = data + 3 * c0 + 3 * c1;
            if((rc = pomx_pair_eval(ctx, POMX_KIND_GF, c2, R, R + c2, int(nres), z1, v)))
                return rc;
            chi3_res_scatter_kernel<<<grid_for(nres), kBlock, 0, ctx->stream>>>(nres, idx, v, d_acc);
            POMX_CHECK_LAUNCH(ctx);
        }
    }
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

int chi3_part_run(pomx_ctx* ctx, pomx_chi3_part const& P, double beta, double res, double tol, int nfreq,
                  const double* d_freq, double* d_acc, pomx_chi3_terms* keep, int64_t* h_counts, int64_t* h_raw) {
    POMX_REQUIRE(ctx, P.n1 >= 1 && P.n2 >= 1 && P.n3 >= 1 && P.n4 >= 1, "pomx_chi3_part_compute: empty block");
    POMX_REQUIRE(ctx, P.channel >= 0 && P.channel <= 2, "pomx_chi3_part_compute: channel must be 0 (PP), 1 (PH) or 2 (xPH)");
    POMX_REQUIRE(ctx, res > 0.0 && tol >= 0.0, "pomx_chi3_part_compute: pole_res must be > 0 and coeff_tol >= 0");
    POMX_REQUIRE(ctx, nfreq >= 0 && (nfreq == 0 || (d_freq && d_acc)), "pomx_chi3_part_compute: bad frequency list");
    if(P.n1 > 65535 || P.n2 > 65535 || P.n3 > 65535)
        return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_chi3_part_compute: block larger than 65535");
    if(keep) {
        keep->data = nullptr;
        keep->count[0] = keep->count[1] = keep->count[2] = 0;
        keep->xi = P.channel == 0 ? -1 : 1;
    }
    int rc;
    // 1. contiguous copies
    WsBuf<double> Bt(ctx, int64_t(P.n1) * P.n3), Fbc(ctx, int64_t(P.n2) * P.n3);
    CHI3_WS(ctx, Bt);
    CHI3_WS(ctx, Fbc);
    chi3_bt_kernel<<<dim3(unsigned((P.n3 + 15) / 16), unsigned((P.n1 + 15) / 16)), 256, 0, ctx->stream>>>(
        P.n1, P.n3, P.n4, P.B1, P.B2, Bt);
    chi3_copy_kernel<<<grid_for(int64_t(P.n2) * P.n3), kBlock, 0, ctx->stream>>>(P.n2, P.n3, P.Fb, Fbc);
    POMX_CHECK_LAUNCH(ctx);
    // 2. key fields: the pole bins themselves when every |E_a - E_b|/res < 2^30 (one range
    //    reduction), else dense ranks of the bins (rank tables of E_y - E_x: ranks of the negated
    //    poles, which group alike)
    WsBuf<double> rng(ctx, 6);
    CHI3_WS(ctx, rng);
    chi3_range_kernel<<<1, 256, 0, ctx->stream>>>(P.n1, P.n2, P.n3, P.E1, P.E2, P.E3, rng);
    POMX_CHECK_LAUNCH(ctx);
    double hr[6];
    POMX_HIP(ctx, hipMemcpyAsync(hr, rng.p, sizeof(hr), hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    auto span = [&](int a, int b) { return std::max(std::fabs(hr[2 * a + 1] - hr[2 * b]), std::fabs(hr[2 * b + 1] - hr[2 * a])); };
    double const widest = std::max({span(0, 1), span(1, 2), span(0, 2)});
    bool const direct = widest / res < 1.0e9; // |bin| <= 1e9 + 1 < 2^30
    WsBuf<uint32_t> r12, r23, r13;
    if(!direct) {
        WsBuf<int> err(ctx, 1);
        CHI3_WS(ctx, err);
        if((rc = pomx_fill_zero(ctx, err.p, sizeof(int))))
            return rc;
        int64_t nr12 = 0, nr23 = 0, nr13 = 0;
        if((rc = rank_table(ctx, P.n1, P.n2, P.E1, P.E2, res, r12, nullptr, &nr12, err)) ||
           (rc = rank_table(ctx, P.n2, P.n3, P.E2, P.E3, res, r23, nullptr, &nr23, err)) ||
           (rc = rank_table(ctx, P.n1, P.n3, P.E1, P.E3, res, r13, nullptr, &nr13, err)))
            return rc;
        CHI3_WS(ctx, r12);
        CHI3_WS(ctx, r23);
        CHI3_WS(ctx, r13);
        int herr = 0;
        POMX_HIP(ctx, hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if(herr)
            return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_chi3_part_compute: |pole/pole_res| out of range");
        if(nr12 > int64_t(kRankMask) || nr23 > int64_t(kRankMask) || nr13 > int64_t(kRankMask))
            return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_chi3_part_compute: too many distinct pole bins for the 64-bit key");
    }
    Chi3Dev D{P.n1, P.n2, P.n3, P.Fa, Fbc, Bt, P.E1, P.E2, P.E3, P.w1, P.w2, P.w3,
              ((P.channel == 1) == (P.swapped != 0)) ? 1.0 : -1.0, beta, tol, res, P.channel == 0, P.swapped != 0};
    // 3. count, scan, emit
    int64_t const nrows = int64_t(P.n1) * P.n2;
    unsigned const rblocks = unsigned(std::min<int64_t>((nrows + 3) / 4, 1 << 16));
    WsBuf<int> rowcnt(ctx, nrows);
    WsBuf<int64_t> wide(ctx, nrows), off(ctx, nrows);
    CHI3_WS(ctx, rowcnt);
    CHI3_WS(ctx, wide);
    CHI3_WS(ctx, off);
    chi3_count_kernel<<<rblocks, 256, 0, ctx->stream>>>(D, nrows, rowcnt);
    chi3_widen_kernel<<<grid_for(nrows), kBlock, 0, ctx->stream>>>(nrows, rowcnt, wide);
    POMX_CHECK_LAUNCH(ctx);
    if((rc = pomx_exclusive_scan(ctx, nrows, wide, off)))
        return rc;
    int64_t lw = 0, lo = 0;
    if((rc = read_i64(ctx, wide.p + nrows - 1, &lw)) || (rc = read_i64(ctx, off.p + nrows - 1, &lo)))
        return rc;
    int64_t const n = lw + lo;
    if(h_raw)
        *h_raw = n;
    int64_t counts[3] = {0, 0, 0};
    double* terms = nullptr;
    if(n > 0) {
        if(n > (int64_t(1) << 33))
            return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_chi3_part_compute: more than 2^33 raw terms in one part");
        WsBuf<uint64_t> k0(ctx, n), k1(ctx, n), v0(ctx, n), v1(ctx, n);
        CHI3_WS(ctx, k0);
        CHI3_WS(ctx, k1);
        CHI3_WS(ctx, v0);
        CHI3_WS(ctx, v1);
        chi3_emit_kernel<<<rblocks, 256, 0, ctx->stream>>>(D, nrows, n, off, direct ? 1 : 0, r12, r23, r13, k0, v0);
        POMX_CHECK_LAUNCH(ctx);
        if((rc = pomx_sort_pairs(ctx, n, k0, k1, v0, v1, 0, 64)))
            return rc;
        // 4. groups, reduce, compact
        WsBuf<int64_t> head(ctx, n), gid(ctx, n);
        CHI3_WS(ctx, head);
        CHI3_WS(ctx, gid);
        chi3_heads_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, k1, head);
        POMX_CHECK_LAUNCH(ctx);
        if((rc = pomx_exclusive_scan(ctx, n, head, gid)))
            return rc;
        int64_t lh = 0, lg = 0;
        if((rc = read_i64(ctx, head.p + n - 1, &lh)) || (rc = read_i64(ctx, gid.p + n - 1, &lg)))
            return rc;
        int64_t const ng = lh + lg;
        WsBuf<int64_t> start(ctx, ng), live(ctx, ng), lpos(ctx, ng), kinfo(ctx, 6);
        WsBuf<double> g(ctx, size_t(ng) * 3);
        CHI3_WS(ctx, start);
        CHI3_WS(ctx, live);
        CHI3_WS(ctx, lpos);
        CHI3_WS(ctx, kinfo);
        CHI3_WS(ctx, g);
        chi3_group_start_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, head, gid, start);
        // first-member positions of the groups, sorted by (kind, position), for the erase divisor
        {
            WsBuf<uint64_t> f0(ctx, ng), f1(ctx, ng), i0(ctx, ng), i1(ctx, ng);
            CHI3_WS(ctx, f0);
            CHI3_WS(ctx, f1);
            CHI3_WS(ctx, i0);
            CHI3_WS(ctx, i1);
            chi3_first_kernel<<<grid_for(ng), kBlock, 0, ctx->stream>>>(ng, start, k1, v1, f0, i0);
            POMX_CHECK_LAUNCH(ctx);
            if((rc = pomx_sort_pairs(ctx, ng, f0, f1, i0, i1, 0, 64)))
                return rc;
            EraseRule const rule{tol, f1.p, ng, nullptr, nullptr, 0};
            EraseEvents ev;
            auto pass = [&](EraseRule const& R, bool emit, int64_t* evn, uint64_t* ko, uint64_t* wo) -> int {
                if(emit)
                    chi3_reduce_kernel<true><<<grid_for(ng), kBlock, 0, ctx->stream>>>(D, ng, n, start, k1, v1, R, g,
                                                                                         live, evn, ko, wo);
                else
                    chi3_reduce_kernel<false><<<grid_for(ng), kBlock, 0, ctx->stream>>>(D, ng, n, start, k1, v1, R, g,
                                                                                          live, evn, ko, wo);
                POMX_CHECK_LAUNCH(ctx);
                return POMX_OK;
            };
            if((rc = erase_fixed_point(ctx, ng, rule, pass, ev, "pomx_chi3_part_compute")))
                return rc;
        }
        if((rc = pomx_exclusive_scan(ctx, ng, live, lpos)))
            return rc;
        chi3_kinds_kernel<<<1, 64, 0, ctx->stream>>>(ng, start, k1, live, lpos, kinfo);
        POMX_CHECK_LAUNCH(ctx);
        int64_t hk[6];
        POMX_HIP(ctx, hipMemcpyAsync(hk, kinfo.p, sizeof(hk), hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        counts[0] = hk[3];
        counts[1] = hk[4];
        counts[2] = hk[5];
        int64_t const ndata = 3 * counts[0] + 3 * counts[1] + 2 * counts[2];
        if(ndata > 0) {
            if(keep) {
                void* p = nullptr;
                POMX_HIP(ctx, device_malloc(ctx, &p, sizeof(double) * size_t(ndata)));
                terms = static_cast<double*>(p);
            } else {
                terms = static_cast<double*>(ws_get(ctx, sizeof(double) * size_t(ndata)));
                if(!terms)
                    return pomx_fail(ctx, POMX_ERR_HIP, "three-point susceptibility: out of device memory");
            }
            chi3_compact_kernel<<<grid_for(ng), kBlock, 0, ctx->stream>>>(ng, live, lpos, g, start, k1, kinfo, terms);
            POMX_CHECK_LAUNCH(ctx);
        }
    }
    for(int k = 0; k < 3; ++k)
        h_counts[k] = counts[k];
    rc = chi3_eval(ctx, terms, counts, P.channel == 0 ? -1 : 1, res, nfreq, d_freq, d_acc);
    if(keep) {
        keep->data = terms;
        for(int k = 0; k < 3; ++k)
            keep->count[k] = counts[k];
    } else if(terms) {
        ws_put(ctx, terms);
    }
    if(rc)
        return rc;
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

} // namespace

extern "C" int pomx_chi3_part_compute(pomx_ctx* ctx, const pomx_chi3_part* part, double beta, double pole_res,
                                      double coeff_tol, int nfreq, const double* d_freq, double* d_acc,
                                      pomx_chi3_terms* keep, int64_t* h_counts, int64_t* h_raw_terms) {
    if(!ctx || !part || !h_counts)
        return ctx ? pomx_fail(ctx, POMX_ERR_ARG, "pomx_chi3_part_compute: null argument") : POMX_ERR_ARG;
    return chi3_part_run(ctx, *part, beta, pole_res, coeff_tol, nfreq, d_freq, d_acc, keep, h_counts, h_raw_terms);
}

extern "C" int pomx_chi3_terms_eval(pomx_ctx* ctx, const pomx_chi3_terms* terms, double pole_res, int nfreq,
                                    const double* d_freq, double* d_acc) {
    if(!ctx || !terms)
        return ctx ? pomx_fail(ctx, POMX_ERR_ARG, "pomx_chi3_terms_eval: null argument") : POMX_ERR_ARG;
    POMX_REQUIRE(ctx, nfreq >= 0 && (nfreq == 0 || (d_freq && d_acc)), "pomx_chi3_terms_eval: bad frequency list");
    POMX_REQUIRE(ctx, terms->count[0] >= 0 && terms->count[1] >= 0 && terms->count[2] >= 0,
                 "pomx_chi3_terms_eval: negative term count");
    POMX_REQUIRE(ctx, terms->data || terms->count[0] + terms->count[1] + terms->count[2] == 0,
                 "pomx_chi3_terms_eval: terms without data");
    return chi3_eval(ctx, terms->data, terms->count, terms->xi, pole_res, nfreq, d_freq, d_acc);
}
