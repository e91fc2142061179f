// K7 (two-particle GF term generation), K8 (pole merge) and K9 (Lehmann sum on a frequency
// triplet list) for one TwoParticleGFPart.
//
// Reference: TwoParticleGFPart::computeImpl and addMultiterm (TwoParticleGFPart.cpp:189-298),
// the NonResonantTerm/ResonantTerm TermLists (TwoParticleGFPart.hpp:83-142,208-268, .cpp:35-43,
// 86-95) and their evaluation (TwoParticleGFPart.hpp:475-495, .cpp:309-322).
//
// Pipeline (all on the device):
//  1. rank tables: every pole is a difference of two block energies, P1 = E2[i2]-E1[i1],
//     P2 = E3[i3]-E2[i2], P3 = E4[i4]-E3[i3]; the bins round(P/res) of each block pair are
//     sorted once and replaced by dense ranks r12, r23, r34, so a term's TermList identity
//     (kind, bin1, bin2, bin3) becomes one 64-bit key (2 + 20 + 21 + 21 bits);
//  2. chunking: (i1,i2) pairs are visited in r12 order and cut into chunks only at r12 group
//     boundaries, so every group of mergeable terms lives in one chunk;
//  3. per chunk: count terms per (i1,i2,i3) row, prefix-sum, emit (key, packed indices),
//     stable radix sort, and reduce every key group (coefficients summed, poles averaged,
//     dropped if the sum is negligible);
//  4. evaluate the merged terms on the frequency list (LDS-staged terms, frequencies in
//     registers, partial sums per term slice reduced in a fixed order).
#include "pomx_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace {

struct G2Dev {
    int n1, n2, n3, n4;
    pomx_view O1, O2, O3, CX4;
    const double *E1, *E2, *E3, *E4;
    const double *w1, *w2, *w3, *w4;
    double sign, beta, tol;
    // bounds of the quadruple sweep (row_bound): wsm[i] = max_{j >= i} w4[j], cdm[i1 n3 + i3] =
    // max_i4 |O3(i3, i4) CX4(i4, i1)|, bf = max(beta, 1) (1 + 1e-9)
    const double* wsm = nullptr;
    const double* cdm = nullptr;
    double bf = 0.0;
};

// derivatives of a part's inputs along a Hamiltonian direction (pomx_g2_tangent)
struct G2Tan {
    pomx_view dO1, dO2, dO3, dCX4;
    const double *dE1, *dE2, *dE3, *dE4;
    const double *dw1, *dw2, *dw3, *dw4;
};

__device__ inline double mat(pomx_view const& v, int r, int c) { return v.base[r * v.rs + c * v.cs]; }

// ---- 1. rank tables -------------------------------------------------------------------------
constexpr int64_t kBinOff = int64_t(1) << 61;

__global__ void bins_kernel(int nx, int ny, const double* __restrict__ Ex, const double* __restrict__ Ey, double res,
                            uint64_t* __restrict__ keys, uint64_t* __restrict__ idx, int* __restrict__ err) {
    int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(g >= int64_t(nx) * ny)
        return;
    int x = int(g / ny), y = int(g % ny);
    double q = round((Ey[y] - Ex[x]) / res);
    if(!(fabs(q) < 1e18)) {
        err[0] = 1;
        q = 0;
    }
    keys[g] = uint64_t(int64_t(q) + kBinOff);
    idx[g] = uint64_t(g);
}

__global__ void rank_heads_kernel(int64_t n, const uint64_t* __restrict__ keys, int64_t* __restrict__ head) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n)
        head[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}

// uniq[idx[k]] = 1 when the sorted key k differs from both neighbours (its bin is held once)
__global__ void unique_bin_kernel(int64_t n, const uint64_t* __restrict__ keys, const uint64_t* __restrict__ idx,
                                  uint8_t* __restrict__ uniq) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n)
        uniq[idx[k]] = ((k == 0 || keys[k] != keys[k - 1]) && (k + 1 == n || keys[k] != keys[k + 1])) ? 1 : 0;
}

__global__ void rank_scatter_kernel(int64_t n, const int64_t* __restrict__ head, const int64_t* __restrict__ excl,
                                    const uint64_t* __restrict__ idx, uint32_t* __restrict__ rank) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n)
        rank[idx[k]] = uint32_t(excl[k] + head[k] - 1);
}

// ---- 3. term counting / emission ------------------------------------------------------------
// The reference's per-quadruple logic (TwoParticleGFPart.cpp:241-297): weights summed in the
// order w1+w2+w3+w4, matrix element ((O1*O2)*O3)*CX4*sign, four candidate terms with their
// own coefficient thresholds.  One wave per (i1,i2,i3) row with the lanes over i4, reading the
// contiguous copies O3c[i3][i4] and C4t[i1][i4] (coalesced).

// bit k set: term of kind k (0 non-resonant z2, 1 non-resonant z4, 2 resonant z2z3,
// 3 resonant z1z2) is emitted for this quadruple
__device__ inline unsigned kinds_of(G2Dev const& D, double ab, double s3, double wi, double wj, double wk, double c,
                                    double d, double wl) {
    if(c == 0.0 || d == 0.0)
        return 0u;
    if(!(s3 + wl > D.tol))
        return 0u;
    double const C = ab * c * d * D.sign;
    double const tol = D.tol;
    unsigned m = 0u;
    if(fabs(C * (wj + wk)) > tol)
        m |= 1u;
    if(fabs(C * (wi + wl)) > tol)
        m |= 2u;
    if(fabs(C * D.beta * wj) > tol || fabs(C * (wj - wl)) > tol)
        m |= 4u;
    if(fabs(C * D.beta * wi) > tol || fabs(C * (wk - wi)) > tol)
        m |= 8u;
    return m;
}

struct RowCtx {
    int i1, i2, i3;
    double ab, s3, wi, wj, wk;
    bool active;
};

// pr = i1 * n2 + i2
__device__ inline RowCtx row_ctx(G2Dev const& D, uint64_t pr, int i3) {
    RowCtx r;
    r.i1 = int(pr / uint64_t(D.n2));
    r.i2 = int(pr % uint64_t(D.n2));
    r.i3 = i3;
    double a = mat(D.O1, r.i1, r.i2), b = mat(D.O2, r.i2, i3);
    r.ab = a * b;
    r.wi = D.w1[r.i1];
    r.wj = D.w2[r.i2];
    r.wk = D.w3[i3];
    r.s3 = r.wi + r.wj + r.wk;
    r.active = (a != 0.0 && b != 0.0) && (r.s3 + D.wsm[0] > D.tol); // wsm[0] = max w4
    return r;
}

// Pruning of the sweep over i4.  Every coefficient kinds_of tests (C (w2 + w3), C (w1 + w4),
// C beta w2, C (w2 - w4), C beta w1, C (w3 - w1)) is at most |C| max(beta, 1) (s3 + w4) for
// non-negative weights, and |C| <= |O1 O2| cdm[i1][i3].  With Bs = |O1 O2| cdm[i1][i3] bf (the
// factor 1 + 1e-9 in bf covers every rounding of those tests and of this bound), i4 can give a
// term only if Bs (s3 + w4[i4]) > tol, hence only if Bs (s3 + wsm[i4]) > tol; wsm is
// non-increasing, so once that fails for some i4 it fails for every later one.  With tol = 0
// nothing is pruned.
__device__ inline double row_bound(G2Dev const& D, double ab, int i1, int i3) {
    return fabs(ab) * D.cdm[int64_t(i1) * D.n3 + i3] * D.bf;
}
__device__ inline bool i4_may_pass(G2Dev const& D, double Bs, double s3, int i4) {
    return !(D.tol > 0.0) || Bs * (s3 + D.wsm[i4]) > D.tol;
}

// Rows of one chunk in the reference's visiting order (i1, i3, i2): the chunk's (i1, i2) pairs
// sorted ascending, each entry x = i1 * n2 + i2, y = the pair's index in the r12-sorted pair list
// (the order of the row counts), z = first entry of its i1 segment, w = entries in that segment.
// Chunk row R lies in the segment of entry R / n3; with local = R - n3 z, i3 = local / w and the
// pair is entry z + local % w.
__device__ inline void chunk_row(const uint4* __restrict__ rmap, int n3, int64_t R, uint4& pr, int& i3) {
    uint4 const s = rmap[R / n3];
    int64_t const local = R - int64_t(n3) * s.z;
    i3 = int(local / s.w);
    pr = rmap[s.z + local % s.w];
}

__device__ inline int wave_sum_int(int v) {
    for(int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline int wave_excl_scan_int(int v, int lane, int& total) {
    int x = v;
    for(int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if(lane >= o)
            x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// contiguous copies: O3c[i3][i4] = O3(i3, i4), C4t[i1][i4] = CX4(i4, i1)
__global__ void g2_stage_kernel(G2Dev D, double* __restrict__ O3c, double* __restrict__ C4t) {
    int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    int64_t n34 = int64_t(D.n3) * D.n4, n14 = int64_t(D.n1) * D.n4;
    if(g < n34)
        O3c[g] = mat(D.O3, int(g / D.n4), int(g % D.n4));
    else if(g < n34 + n14) {
        int64_t h = g - n34;
        C4t[h] = mat(D.CX4, int(h % D.n4), int(h / D.n4));
    }
}

// wsm[i] = max_{j >= i} w4[j] (one block of kBlock threads, each over a contiguous segment)
__global__ void __launch_bounds__(256) suffix_max_kernel(int n, const double* __restrict__ w, double* __restrict__ wsm) {
    __shared__ double seg_max[256];
    int const t = threadIdx.x, seg = (n + blockDim.x - 1) / blockDim.x;
    int const lo = min(n, t * seg), hi = min(n, lo + seg);
    double m = 0.0; // weights are non-negative
    for(int i = hi - 1; i >= lo; --i)
        m = fmax(m, w[i]);
    seg_max[t] = m;
    __syncthreads();
    if(t == 0) // maxima of the segments to the right, from the last segment down
        for(int u = int(blockDim.x) - 2; u >= 0; --u)
            seg_max[u] = fmax(seg_max[u], seg_max[u + 1]);
    __syncthreads();
    double r = (t + 1 < int(blockDim.x)) ? seg_max[t + 1] : 0.0;
    for(int i = hi - 1; i >= lo; --i) {
        r = fmax(r, w[i]);
        wsm[i] = r;
    }
}

// cdm[i1 n3 + i3] = max_i4 |O3c[i3][i4] C4t[i1][i4]| (one wave per entry, lanes over i4)
__global__ void __launch_bounds__(256) cd_max_kernel(G2Dev D, const double* __restrict__ O3c,
                                                     const double* __restrict__ C4t, double* __restrict__ cdm) {
    int const lane = threadIdx.x & 63;
    int64_t const n = int64_t(D.n1) * D.n3;
    int64_t const nwaves = int64_t(gridDim.x) * (blockDim.x >> 6);
    for(int64_t e = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6; e < n; e += nwaves) {
        int const i1 = int(e / D.n3), i3 = int(e % D.n3);
        const double* o3 = O3c + int64_t(i3) * D.n4;
        const double* c4 = C4t + int64_t(i1) * D.n4;
        double m = 0.0;
        for(int i4 = lane; i4 < D.n4; i4 += 64)
            m = fmax(m, fabs(o3[i4] * c4[i4]));
        for(int o = 32; o > 0; o >>= 1)
            m = fmax(m, __shfl_xor(m, o, 64));
        if(lane == 0)
            cdm[e] = m;
    }
}

// Direct terms.  A term's (kind, bin(P1), bin(P2), bin(P3)) can equal another term's only if its
// (i1, i2) shares the P1 bin with another pair or its (i3, i4) shares the P3 bin with another pair
// (equal bins of both pin (i1, i2) and (i3, i4), hence the quadruple).  Terms whose (i1, i2) and
// (i3, i4) bins are both unique in their rank tables (u12, u34) are therefore groups of one: they
// skip the sort and the reduce and are written straight into the merged term lists (direct);
// the others (shared bins; 0.3 % of the config-2 terms) go through the sort (sorted).

// The live rows of one (i1, i2) pair, 64 at a time: lane l tests row i3 = b3 + l (non-zero
// O1 O2, and some i4 that may pass, i4_may_pass at i4 = 0), and the wave then visits the live rows
// in ascending i3 with its lanes over i4, up to the first 64 i4 of which none may pass.
struct PairRows {
    int i1, i2;
    double a, wi, wj;
};
__device__ inline PairRows pair_rows(G2Dev const& D, uint64_t pr) {
    PairRows p;
    p.i1 = int(pr / uint64_t(D.n2));
    p.i2 = int(pr % uint64_t(D.n2));
    p.a = mat(D.O1, p.i1, p.i2);
    p.wi = D.w1[p.i1];
    p.wj = D.w2[p.i2];
    return p;
}
// lane's row i3: ab, s3, wk, Bs; returns whether it is live
__device__ inline bool lane_row(G2Dev const& D, PairRows const& p, int i3, double& ab, double& s3, double& wk,
                                double& Bs) {
    ab = s3 = wk = Bs = 0.0;
    if(p.a == 0.0 || i3 >= D.n3)
        return false;
    double const b = mat(D.O2, p.i2, i3);
    ab = p.a * b;
    wk = D.w3[i3];
    s3 = p.wi + p.wj + wk;
    Bs = row_bound(D, ab, p.i1, i3);
    return b != 0.0 && (s3 + D.wsm[0] > D.tol) && i4_may_pass(D, Bs, s3, 0);
}

// One wave per (i1, i2) pair (position pos of the r12-sorted pair list), live rows i3 in turn,
// lanes over i4: the sorted terms per row (rowcnt[pos n3 + i3]) and per pair (paircnt[pos]), and
// the direct terms per pair and kind (dpc[k npairs + pos]).
__global__ void __launch_bounds__(256) g2_count_kernel(G2Dev D, const double* __restrict__ O3c,
                                                       const double* __restrict__ C4t, const uint64_t* __restrict__ pairs,
                                                       int64_t npairs, const uint8_t* __restrict__ u12,
                                                       const uint8_t* __restrict__ u34, int* __restrict__ rowcnt,
                                                       int64_t* __restrict__ paircnt, int64_t* __restrict__ dpc) {
    int const lane = threadIdx.x & 63;
    int64_t const nwaves = int64_t(gridDim.x) * (blockDim.x >> 6);
    for(int64_t pos = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6; pos < npairs; pos += nwaves) {
        uint64_t const pr = pairs[pos];
        PairRows const P = pair_rows(D, pr);
        bool const d12 = u12 && u12[pr];
        int64_t sorted_pair = 0;
        int dk[4] = {0, 0, 0, 0};
        const double* c4 = C4t + int64_t(P.i1) * D.n4;
        for(int b3 = 0; b3 < D.n3; b3 += 64) {
            double lab, ls3, lwk, lBs;
            uint64_t live = __ballot(lane_row(D, P, b3 + lane, lab, ls3, lwk, lBs));
            int mine = 0; // sorted terms of this lane's row
            while(live) {
                int const l = __ffsll((unsigned long long)live) - 1;
                live &= live - 1;
                int const i3 = b3 + l;
                double const ab = __shfl(lab, l, 64), s3 = __shfl(ls3, l, 64), wk = __shfl(lwk, l, 64);
                double const Bs = __shfl(lBs, l, 64);
                const double* o3 = O3c + int64_t(i3) * D.n4;
                const uint8_t* u34row = u34 ? u34 + int64_t(i3) * D.n4 : nullptr;
                int c = 0;
                for(int b4 = 0; b4 < D.n4; b4 += 64) {
                    int const i4 = b4 + lane;
                    bool const may = i4 < D.n4 && i4_may_pass(D, Bs, s3, i4);
                    if(!__ballot(may))
                        break;
                    if(may) {
                        unsigned const m = kinds_of(D, ab, s3, P.wi, P.wj, wk, o3[i4], c4[i4], D.w4[i4]);
                        if(d12 && u34row[i4]) {
                            dk[0] += int(m & 1u);
                            dk[1] += int((m >> 1) & 1u);
                            dk[2] += int((m >> 2) & 1u);
                            dk[3] += int((m >> 3) & 1u);
                        } else
                            c += __popc(m);
                    }
                }
                c = wave_sum_int(c);
                sorted_pair += c;
                if(lane == l)
                    mine = c;
            }
            if(b3 + lane < D.n3)
                rowcnt[pos * D.n3 + b3 + lane] = mine;
        }
#pragma unroll
        for(int k = 0; k < 4; ++k)
            dk[k] = wave_sum_int(dk[k]);
        if(lane == 0) {
            paircnt[pos] = sorted_pair;
#pragma unroll
            for(int k = 0; k < 4; ++k)
                dpc[k * npairs + pos] = dk[k];
        }
    }
}

// The merged term lists of one part, one buffer per kind: non-resonant kinds (C, P1, P2, P3), resonant
// kinds (Rc, Nc, P1, P2, P3), field f of term t at base[k][f cap[k] + t].
struct KindLists {
    double* base[4];
    int64_t cap[4];
};

// row counts of a chunk in visiting order (cnt is in r12-sorted pair order)
__global__ void chunk_rowcnt_kernel(int64_t nrows, int n3, const uint4* __restrict__ rmap, const int* __restrict__ cnt,
                                    int64_t* __restrict__ out) {
    int64_t R = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(R >= nrows)
        return;
    uint4 pr;
    int i3;
    chunk_row(rmap, n3, R, pr, i3);
    out[R] = cnt[int64_t(pr.y) * n3 + i3];
}

// Emission in visiting order: rows of the chunk as chunk_row orders them, lanes over i4, kinds in
// addMultiterm's order, so after a stable sort by key the members of every group are in the
// reference's visiting order.
__global__ void __launch_bounds__(256) g2_emit_kernel(G2Dev D, const double* __restrict__ O3c,
                                                      const double* __restrict__ C4t, const uint4* __restrict__ rmap,
                                                      int64_t nrows, int64_t ntot,
                                                      const int64_t* __restrict__ off, const uint32_t* __restrict__ r12,
                                                      const uint32_t* __restrict__ r23, const uint32_t* __restrict__ r34,
                                                      const uint8_t* __restrict__ u12, const uint8_t* __restrict__ u34,
                                                      uint64_t* __restrict__ keys, uint64_t* __restrict__ vals) {
    int const lane = threadIdx.x & 63;
    int64_t const nwaves = int64_t(gridDim.x) * (blockDim.x >> 6);
    for(int64_t lr = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6; lr < nrows; lr += nwaves) {
        uint4 pr;
        int i3;
        chunk_row(rmap, D.n3, lr, pr, i3);
        int64_t at = off[lr];
        int64_t const end = (lr + 1 < nrows) ? off[lr + 1] : ntot; // never write past this row's slots
        if(at >= end) // no sorted term in this row
            continue;
        RowCtx r = row_ctx(D, pr.x, i3);
        bool const d12 = u12 && u12[pr.x]; // direct terms of this row are written elsewhere
        if(!r.active)
            continue;
        double const Bs = row_bound(D, r.ab, r.i1, r.i3);
        uint64_t const k12 = uint64_t(r12[int64_t(r.i1) * D.n2 + r.i2]) << 42;
        uint64_t const k23 = uint64_t(r23[int64_t(r.i2) * D.n3 + r.i3]) << 21;
        uint64_t const base = uint64_t(r.i1) | (uint64_t(r.i2) << 16) | (uint64_t(r.i3) << 32);
        const double* o3 = O3c + int64_t(r.i3) * D.n4;
        const double* c4 = C4t + int64_t(r.i1) * D.n4;
        const uint32_t* r34row = r34 + int64_t(r.i3) * D.n4;
        for(int b4 = 0; b4 < D.n4; b4 += 64) {
            int const i4 = b4 + lane;
            bool const may = i4 < D.n4 && i4_may_pass(D, Bs, r.s3, i4);
            if(!__ballot(may))
                break;
            unsigned m = 0u;
            if(may && !(d12 && u34[int64_t(r.i3) * D.n4 + i4]))
                m = kinds_of(D, r.ab, r.s3, r.wi, r.wj, r.wk, o3[i4], c4[i4], D.w4[i4]);
            int total;
            int64_t p = at + wave_excl_scan_int(__popc(m), lane, total);
            uint64_t const k34 = (i4 < D.n4) ? uint64_t(r34row[i4]) : 0;
            for(int kind = 0; kind < 4; ++kind)
                if((m >> kind) & 1u) {
                    if(p < end) {
                        keys[p] = (uint64_t(kind) << 62) | k12 | k23 | k34;
                        vals[p] = base | (uint64_t(i4) << 48);
                    }
                    ++p;
                }
            at += total;
        }
    }
}

__global__ void heads_kernel(int64_t n, const uint64_t* __restrict__ keys, int64_t* __restrict__ head) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n)
        head[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}
__global__ void group_start_kernel(int64_t n, const int64_t* __restrict__ head, const int64_t* __restrict__ gid,
                                   int64_t* __restrict__ start) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n && head[k])
        start[gid[k]] = k;
}

// Visiting position of a raw term in the reference's loop (TwoParticleGFPart.cpp:209-250:
// i1, i3, i2, i4, then addMultiterm's order NR z2, NR z4 | R z1z2, R z2z3), with the TermList
// (non-resonant 0, resonant 1) in the top bit.
__device__ inline uint64_t g2_position(G2Dev const& D, int kind, int i1, int i2, int i3, int i4) {
    uint64_t const q = ((uint64_t(i1) * D.n3 + i3) * D.n2 + i2) * D.n4 + i4;
    uint64_t const slot = (kind == 1 || kind == 2) ? 1u : 0u;
    return (uint64_t(kind >= 2) << 63) | (q << 1) | slot;
}
__device__ inline void g2_unpack(uint64_t v, int& i1, int& i2, int& i3, int& i4) {
    i1 = int(v & 0xffff);
    i2 = int((v >> 16) & 0xffff);
    i3 = int((v >> 32) & 0xffff);
    i4 = int(v >> 48);
}

// Group members are in visiting order (g2_emit_kernel + a stable sort), so a group's first
// member starts it.  Visiting position of every group's first member:
__global__ void g2_first_pos_kernel(G2Dev D, int64_t ngroups, const int64_t* __restrict__ start,
                                    const uint64_t* __restrict__ keys, const uint64_t* __restrict__ vals,
                                    uint64_t* __restrict__ fk) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups)
        return;
    int64_t const s = start[gi];
    int i1, i2, i3, i4;
    g2_unpack(vals[s], i1, i2, i3, i4);
    fk[gi] = g2_position(D, int(keys[s] >> 62), i1, i2, i3, i4);
}

// Positions of every merge (every member but the first of a group): group gi's merges go to
// q[start[gi] - gi ...] (the number of merges of the groups before it).
__global__ void g2_merge_pos_kernel(G2Dev D, int64_t n, int64_t ngroups, const int64_t* __restrict__ start,
                                    const uint64_t* __restrict__ keys, const uint64_t* __restrict__ vals,
                                    uint64_t* __restrict__ q) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups)
        return;
    int64_t const s = start[gi], e = (gi + 1 < ngroups) ? start[gi + 1] : n;
    int const kind = int(keys[s] >> 62);
    uint64_t* out = q + (s - gi);
    for(int64_t k = s + 1; k < e; ++k) {
        int i1, i2, i3, i4;
        g2_unpack(vals[k], i1, i2, i3, i4);
        out[k - s - 1] = g2_position(D, kind, i1, i2, i3, i4);
    }
}

// hist[j] += number of groups whose first position fk lies between q[j-1] and q[j] (q ascending,
// m entries; j = number of q below fk), so that the inclusive prefix sum at j counts the groups
// started before the merge q[j].  Workgroup histograms in LDS when the bins fit.
constexpr int kHistLds = 8192;
__global__ void __launch_bounds__(256) g2_first_hist_kernel(int64_t nfk, const uint64_t* __restrict__ fk,
                                                            const uint64_t* __restrict__ q, int64_t m,
                                                            unsigned long long* __restrict__ hist) {
    __shared__ unsigned int lh[kHistLds];
    bool const local = m + 1 <= kHistLds;
    if(local)
        for(int k = threadIdx.x; k <= m; k += blockDim.x)
            lh[k] = 0u;
    __syncthreads();
    for(int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nfk; i += int64_t(gridDim.x) * blockDim.x) {
        int64_t const j = upper_index(q, 0, m, fk[i]);
        if(local)
            atomicAdd(&lh[j], 1u);
        else
            atomicAdd(&hist[j], 1ull);
    }
    __syncthreads();
    if(local)
        for(int k = threadIdx.x; k <= m; k += blockDim.x)
            if(lh[k])
                atomicAdd(&hist[k], (unsigned long long)lh[k]);
}

// inclusive = exclusive + own
__global__ void add_own_kernel(int64_t n, const int64_t* __restrict__ own, int64_t* __restrict__ incl) {
    int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(k < n)
        incl[k] += own[k];
}

// out = started[j] at the entry q[j] == key
__global__ void lookup_started_kernel(const uint64_t* __restrict__ q, int64_t m, const int64_t* __restrict__ started,
                                      uint64_t key, int64_t* __restrict__ out) {
    if(blockIdx.x == 0 && threadIdx.x == 0)
        out[0] = started[upper_index(q, 0, m, key) - 1];
}

// Direct terms: one wave per pair with unique bins, rows i3 in turn, lanes over i4; the terms of
// each kind are placed by ballot prefix counts from the pair's offset doff[k npairs + pos].  With
// q (the part's merge positions, m entries) the visiting position of every direct term, a group of
// one, is counted into the gap between merges it falls in (as g2_first_hist_kernel does).  TAN:
// the tangent kind lists, each term also carrying the derivatives of its coefficients and poles
// along TD (as g2_reduce_tangent_kernel forms them for a group of one).
template <bool TAN>
__global__ void __launch_bounds__(256) g2_direct_emit_kernel(G2Dev D, G2Tan TD, const double* __restrict__ O3c,
                                                             const double* __restrict__ C4t,
                                                             const uint64_t* __restrict__ pairs, int64_t npairs,
                                                             const uint8_t* __restrict__ u12,
                                                             const uint8_t* __restrict__ u34,
                                                             const int64_t* __restrict__ doff, KindLists out,
                                                             const uint64_t* __restrict__ q, int64_t m,
                                                             unsigned long long* __restrict__ hist) {
    __shared__ unsigned int lh[kHistLds];
    bool const local = q && m + 1 <= kHistLds;
    if(local)
        for(int k = threadIdx.x; k <= m; k += blockDim.x)
            lh[k] = 0u;
    __syncthreads();
    int const lane = threadIdx.x & 63;
    uint64_t const below = (uint64_t(1) << lane) - 1u;
    int64_t const nwaves = int64_t(gridDim.x) * (blockDim.x >> 6);
    for(int64_t pos = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6; pos < npairs; pos += nwaves) {
        uint64_t const pr = pairs[pos];
        if(!u12[pr])
            continue;
        PairRows const PR = pair_rows(D, pr);
        int64_t at[4];
#pragma unroll
        for(int k = 0; k < 4; ++k)
            at[k] = doff[k * npairs + pos];
        const double* c4 = C4t + int64_t(PR.i1) * D.n4;
        double const P1 = D.E2[PR.i2] - D.E1[PR.i1];
        for(int b3 = 0; b3 < D.n3; b3 += 64) {
            double lab, ls3, lwk, lBs;
            uint64_t live = __ballot(lane_row(D, PR, b3 + lane, lab, ls3, lwk, lBs));
            while(live) {
            int const l = __ffsll((unsigned long long)live) - 1;
            live &= live - 1;
            int const i3 = b3 + l;
            double const ab = __shfl(lab, l, 64), s3 = __shfl(ls3, l, 64), wk = __shfl(lwk, l, 64);
            double const Bs = __shfl(lBs, l, 64);
            const double* o3 = O3c + int64_t(i3) * D.n4;
            double const P2 = D.E3[i3] - D.E2[PR.i2];
            // tangent: row constants (o2, dw1..dw3, dP1, dP2)
            double o2 = 0.0, dwi = 0.0, dwj = 0.0, dwk = 0.0, dP1 = 0.0, dP2 = 0.0;
            if constexpr(TAN) {
                o2 = mat(D.O2, PR.i2, i3);
                dwi = TD.dw1[PR.i1];
                dwj = TD.dw2[PR.i2];
                dwk = TD.dw3[i3];
                dP1 = TD.dE2[PR.i2] - TD.dE1[PR.i1];
                dP2 = TD.dE3[i3] - TD.dE2[PR.i2];
            }
            for(int b4 = 0; b4 < D.n4; b4 += 64) {
                int const i4 = b4 + lane;
                bool const may = i4 < D.n4 && i4_may_pass(D, Bs, s3, i4);
                if(!__ballot(may))
                    break;
                unsigned mk = 0u;
                double C = 0.0, wl = 0.0, P3 = 0.0, dC = 0.0, dwl = 0.0, dP3 = 0.0;
                if(may && u34[int64_t(i3) * D.n4 + i4]) {
                    double const c = o3[i4], d = c4[i4];
                    wl = D.w4[i4];
                    mk = kinds_of(D, ab, s3, PR.wi, PR.wj, wk, c, d, wl);
                    C = ab * c * d * D.sign; // ((O1 O2) O3) CX4 sign, as the reduce forms it
                    P3 = D.E4[i4] - D.E3[i3];
                    if constexpr(TAN) { // as g2_reduce_tangent_kernel: d(o1 o2 o3 o4) sign
                        double const o1 = PR.a;
                        dC = (mat(TD.dO1, PR.i1, PR.i2) * o2 * c * d + o1 * mat(TD.dO2, PR.i2, i3) * c * d +
                              o1 * o2 * mat(TD.dO3, i3, i4) * d + o1 * o2 * c * mat(TD.dCX4, i4, PR.i1)) *
                             D.sign;
                        dwl = TD.dw4[i4];
                        dP3 = TD.dE4[i4] - TD.dE3[i3];
                    }
                }
#pragma unroll
                for(int k = 0; k < 4; ++k) {
                    bool const bit = (mk >> k) & 1u;
                    uint64_t const ball = __ballot(bit);
                    if(bit) {
                        int64_t const t = at[k] + __popcll(ball & below);
                        double* B = out.base[k];
                        int64_t const cap = out.cap[k];
                        if(k < 2) {
                            B[t] = (k == 0) ? -C * (PR.wj + wk) : C * (PR.wi + wl);
                            B[cap + t] = P1;
                            B[2 * cap + t] = P2;
                            B[3 * cap + t] = P3;
                            if constexpr(TAN) {
                                B[4 * cap + t] = (k == 0) ? -(dC * (PR.wj + wk) + C * (dwj + dwk))
                                                          : dC * (PR.wi + wl) + C * (dwi + dwl);
                                B[5 * cap + t] = dP1;
                                B[6 * cap + t] = dP2;
                                B[7 * cap + t] = dP3;
                            }
                        } else {
                            B[t] = (k == 2) ? -C * D.beta * PR.wj : C * D.beta * PR.wi;
                            B[cap + t] = (k == 2) ? C * (PR.wj - wl) : C * (wk - PR.wi);
                            B[2 * cap + t] = P1;
                            B[3 * cap + t] = P2;
                            B[4 * cap + t] = P3;
                            if constexpr(TAN) {
                                B[5 * cap + t] = (k == 2) ? -D.beta * (dC * PR.wj + C * dwj)
                                                          : D.beta * (dC * PR.wi + C * dwi);
                                B[6 * cap + t] = (k == 2) ? dC * (PR.wj - wl) + C * (dwj - dwl)
                                                          : dC * (wk - PR.wi) + C * (dwk - dwi);
                                B[7 * cap + t] = dP1;
                                B[8 * cap + t] = dP2;
                                B[9 * cap + t] = dP3;
                            }
                        }
                        if(q) {
                            int64_t const j = upper_index(q, 0, m, g2_position(D, k, PR.i1, PR.i2, i3, i4));
                            if(local)
                                atomicAdd(&lh[j], 1u);
                            else
                                atomicAdd(&hist[j], 1ull);
                        }
                    }
                    at[k] += __popcll(ball);
                }
            }
            }
        }
    }
    __syncthreads();
    if(local)
        for(int k = threadIdx.x; k <= m; k += blockDim.x)
            if(lh[k])
                atomicAdd(&hist[k], (unsigned long long)lh[k]);
}

// One thread per key group: sum the members' coefficients (pole = count-weighted mean).
// The erase threshold is thr, or, with R.fk set, the reference's tol/(size+1) (EraseRule); EMIT
// as for erase_fixed_point (events written at evn[gi], group results untouched).  With the fixed
// thr, ambig (if given) is set when a merge's max(|c0|, |c1|) lies in (thr, R.tol]: only such a
// merge can be decided differently by the reference's threshold, which lies in that interval.
template <bool EMIT>
__global__ void g2_reduce_kernel(G2Dev D, int64_t n, int64_t ngroups, const int64_t* __restrict__ start,
                                 const uint64_t* __restrict__ keys, const uint64_t* __restrict__ vals, double thr,
                                 EraseRule R, double* __restrict__ g, int64_t* __restrict__ live,
                                 int64_t* __restrict__ evn, uint64_t* __restrict__ ev_key_out,
                                 uint64_t* __restrict__ ev_w_out, int* __restrict__ ambig) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups)
        return;
    int64_t s = start[gi], e = (gi + 1 < ngroups) ? start[gi + 1] : n;
    int kind = int(keys[s] >> 62);
    EraseCursor const cur = R.fk ? erase_cursor(R, uint64_t(kind >= 2) << 63) : EraseCursor{0, 0, 0};
    int64_t nevents = 0, at = EMIT ? evn[gi] : 0;
    bool alive = false;
    double c0 = 0, c1 = 0, p1 = 0, p2 = 0, p3 = 0;
    long cnt = 0;
    for(int64_t k = s; k < e; ++k) {
        uint64_t v = vals[k];
        int i1 = int(v & 0xffff), i2 = int((v >> 16) & 0xffff), i3 = int((v >> 32) & 0xffff), i4 = int(v >> 48);
        double C = mat(D.O1, i1, i2) * mat(D.O2, i2, i3) * mat(D.O3, i3, i4) * mat(D.CX4, i4, i1) * D.sign;
        double wi = D.w1[i1], wj = D.w2[i2], wk = D.w3[i3], wl = D.w4[i4];
        double a, b = 0.0;
        if(kind == 0)
            a = -C * (wj + wk);
        else if(kind == 1)
            a = C * (wi + wl);
        else if(kind == 2) {
            a = -C * D.beta * wj;
            b = C * (wj - wl);
        } else {
            a = C * D.beta * wi;
            b = C * (wk - wi);
        }
        double P1 = D.E2[i2] - D.E1[i1], P2 = D.E3[i3] - D.E2[i2], P3 = D.E4[i4] - D.E3[i3];
        if(!alive) {
            if(k > s && R.fk) { // restart of an erased group
                if(EMIT) {
                    ev_key_out[at] = g2_position(D, kind, i1, i2, i3, i4);
                    ev_w_out[at++] = uint64_t(int64_t(-1));
                }
                ++nevents;
            }
            alive = true;
            c0 = a;
            c1 = b;
            p1 = P1;
            p2 = P2;
            p3 = P3;
            cnt = 1;
        } else {
            c0 += a;
            c1 += b;
            p1 += P1;
            p2 += P2;
            p3 += P3;
            ++cnt;
            double const t = R.fk ? erase_threshold(R, cur, g2_position(D, kind, i1, i2, i3, i4)) : thr;
            if(ambig) {
                double const x = fmax(fabs(c0), fabs(c1));
                if(x > thr && x <= R.tol)
                    ambig[0] = 1;
            }
            if(fabs(c0) <= t && fabs(c1) <= t) {
                alive = false;
                if(R.fk) {
                    if(EMIT) {
                        ev_key_out[at] = g2_position(D, kind, i1, i2, i3, i4);
                        ev_w_out[at++] = 1;
                    }
                    ++nevents;
                }
            }
        }
    }
    if(!EMIT) {
        double inv = 1.0 / double(cnt);
        g[gi * 5 + 0] = c0;
        g[gi * 5 + 1] = c1;
        g[gi * 5 + 2] = p1 * inv;
        g[gi * 5 + 3] = p2 * inv;
        g[gi * 5 + 4] = p3 * inv;
        live[gi] = alive ? 1 : 0;
        if(evn)
            evn[gi] = nevents;
    }
}

// Where a chunk's live groups go in the part's kind lists: the live groups of kind k (contiguous in
// the live order, from live index first[k]) to positions at[k] on.
struct KindSlots {
    int64_t first[4], at[4];
};

// scatter live groups into the kind lists (NR kinds: C, P1, P2, P3; R kinds: Rc, Nc, P1, P2, P3)
__global__ void g2_compact_kernel(int64_t ngroups, const int64_t* __restrict__ live, const int64_t* __restrict__ lpos,
                                  const double* __restrict__ g, const int64_t* __restrict__ start,
                                  const uint64_t* __restrict__ keys, KindLists out, KindSlots slot) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups || !live[gi])
        return;
    int const kind = int(keys[start[gi]] >> 62);
    int64_t const t = slot.at[kind] + (lpos[gi] - slot.first[kind]);
    double* B = out.base[kind];
    int64_t const cap = out.cap[kind];
    const double* x = g + gi * 5;
    if(kind < 2) {
        B[t] = x[0];
        B[cap + t] = x[2];
        B[2 * cap + t] = x[3];
        B[3 * cap + t] = x[4];
    } else {
        B[t] = x[0];
        B[cap + t] = x[1];
        B[2 * cap + t] = x[2];
        B[3 * cap + t] = x[3];
        B[4 * cap + t] = x[4];
    }
}

// Tangent of g2_reduce_kernel: the same groups, sums, erasures and pole means (computed in the
// same order, so the values match the forward path bit for bit), each group also carrying the
// derivative of its summed coefficients and of its mean poles.  g holds 10 doubles per group:
// c0, c1, P1, P2, P3, dc0, dc1, dP1, dP2, dP3.
__global__ void g2_reduce_tangent_kernel(G2Dev D, G2Tan T, int64_t n, int64_t ngroups, const int64_t* __restrict__ start,
                                         const uint64_t* __restrict__ keys, const uint64_t* __restrict__ vals, double thr,
                                         EraseRule R, double* __restrict__ g, int64_t* __restrict__ live) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups)
        return;
    int64_t s = start[gi], e = (gi + 1 < ngroups) ? start[gi + 1] : n;
    int kind = int(keys[s] >> 62);
    EraseCursor const cur = R.fk ? erase_cursor(R, uint64_t(kind >= 2) << 63) : EraseCursor{0, 0, 0};
    bool alive = false;
    double c0 = 0, c1 = 0, p1 = 0, p2 = 0, p3 = 0;
    double dc0 = 0, dc1 = 0, dp1 = 0, dp2 = 0, dp3 = 0;
    long cnt = 0;
    for(int64_t k = s; k < e; ++k) {
        uint64_t v = vals[k];
        int i1 = int(v & 0xffff), i2 = int((v >> 16) & 0xffff), i3 = int((v >> 32) & 0xffff), i4 = int(v >> 48);
        double const o1 = mat(D.O1, i1, i2), o2 = mat(D.O2, i2, i3), o3 = mat(D.O3, i3, i4), o4 = mat(D.CX4, i4, i1);
        double const C = o1 * o2 * o3 * o4 * D.sign;
        double const dC = (mat(T.dO1, i1, i2) * o2 * o3 * o4 + o1 * mat(T.dO2, i2, i3) * o3 * o4 +
                           o1 * o2 * mat(T.dO3, i3, i4) * o4 + o1 * o2 * o3 * mat(T.dCX4, i4, i1)) *
                          D.sign;
        double wi = D.w1[i1], wj = D.w2[i2], wk = D.w3[i3], wl = D.w4[i4];
        double dwi = T.dw1[i1], dwj = T.dw2[i2], dwk = T.dw3[i3], dwl = T.dw4[i4];
        double a, b = 0.0, da, db = 0.0;
        if(kind == 0) {
            a = -C * (wj + wk);
            da = -(dC * (wj + wk) + C * (dwj + dwk));
        } else if(kind == 1) {
            a = C * (wi + wl);
            da = dC * (wi + wl) + C * (dwi + dwl);
        } else if(kind == 2) {
            a = -C * D.beta * wj;
            b = C * (wj - wl);
            da = -D.beta * (dC * wj + C * dwj);
            db = dC * (wj - wl) + C * (dwj - dwl);
        } else {
            a = C * D.beta * wi;
            b = C * (wk - wi);
            da = D.beta * (dC * wi + C * dwi);
            db = dC * (wk - wi) + C * (dwk - dwi);
        }
        double P1 = D.E2[i2] - D.E1[i1], P2 = D.E3[i3] - D.E2[i2], P3 = D.E4[i4] - D.E3[i3];
        double dP1 = T.dE2[i2] - T.dE1[i1], dP2 = T.dE3[i3] - T.dE2[i2], dP3 = T.dE4[i4] - T.dE3[i3];
        if(!alive) {
            alive = true;
            c0 = a;
            c1 = b;
            p1 = P1;
            p2 = P2;
            p3 = P3;
            dc0 = da;
            dc1 = db;
            dp1 = dP1;
            dp2 = dP2;
            dp3 = dP3;
            cnt = 1;
        } else {
            c0 += a;
            c1 += b;
            p1 += P1;
            p2 += P2;
            p3 += P3;
            dc0 += da;
            dc1 += db;
            dp1 += dP1;
            dp2 += dP2;
            dp3 += dP3;
            ++cnt;
            double const t = R.fk ? erase_threshold(R, cur, g2_position(D, kind, i1, i2, i3, i4)) : thr;
            if(fabs(c0) <= t && fabs(c1) <= t)
                alive = false;
        }
    }
    double inv = 1.0 / double(cnt);
    double* o = g + gi * 10;
    o[0] = c0;
    o[1] = c1;
    o[2] = p1 * inv;
    o[3] = p2 * inv;
    o[4] = p3 * inv;
    o[5] = dc0;
    o[6] = dc1;
    o[7] = dp1 * inv;
    o[8] = dp2 * inv;
    o[9] = dp3 * inv;
    live[gi] = alive ? 1 : 0;
}

// Live tangent groups (g2_reduce_tangent_kernel) into the part's tangent kind lists: NR (c, P1, P2,
// P3, dc, dP1, dP2, dP3), R (Rc, Nc, P1, P2, P3, dRc, dNc, dP1, dP2, dP3).
__global__ void g2_compact_tangent_kernel(int64_t ngroups, const int64_t* __restrict__ live,
                                          const int64_t* __restrict__ lpos, const double* __restrict__ g,
                                          const int64_t* __restrict__ start, const uint64_t* __restrict__ keys,
                                          KindLists out, KindSlots slot) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups || !live[gi])
        return;
    int const kind = int(keys[start[gi]] >> 62);
    int64_t const t = slot.at[kind] + (lpos[gi] - slot.first[kind]);
    double* B = out.base[kind];
    int64_t const cap = out.cap[kind];
    const double* x = g + gi * 10;
    if(kind < 2) {
        int const f[8] = {0, 2, 3, 4, 5, 7, 8, 9};
#pragma unroll
        for(int q = 0; q < 8; ++q)
            B[q * cap + t] = x[f[q]];
    } else {
#pragma unroll
        for(int q = 0; q < 10; ++q)
            B[q * cap + t] = x[q];
    }
}

// Live groups per kind.  Groups are sorted by key and the kind sits in the top key bits, so the
// groups of one kind are contiguous: binary-search the kind boundaries and difference the
// exclusive scan of the live flags (one thread; no atomics).
__global__ void kind_count_kernel(int64_t ngroups, const int64_t* __restrict__ live, const int64_t* __restrict__ lpos,
                                  const int64_t* __restrict__ start, const uint64_t* __restrict__ keys,
                                  int64_t* __restrict__ cnt) {
    if(blockIdx.x != 0 || threadIdx.x != 0)
        return;
    int64_t before[5];
    for(int k = 0; k <= 4; ++k) {
        int64_t lo = 0, hi = ngroups; // first group with kind >= k
        while(lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if(int(keys[start[mid]] >> 62) >= k)
                hi = mid;
            else
                lo = mid + 1;
        }
        before[k] = (lo < ngroups) ? lpos[lo] : (lpos[ngroups - 1] + live[ngroups - 1]);
    }
    for(int k = 0; k < 4; ++k)
        cnt[k] = before[k + 1] - before[k];
}

// ---- 4. evaluation --------------------------------------------------------------------------
// One workgroup = 256 threads x kFPT frequency triplets, one slice of the term list.  Terms are
// staged through LDS 256 at a time (with their per-term constants precomputed at staging);
// every wave then reads each term with a uniform LDS address and evaluates it for all its
// lanes' frequencies.  Partial sums per slice are reduced in a fixed order (deterministic).
constexpr int kChunk = 256;
constexpr int kFPT = 8; // frequencies per thread (8 independent accumulator pairs per lane; multiple of 4)

__device__ inline double rcp_nr(double x) {
    // v_rcp_f64 estimate (relative error <= 2^-24.4 on gfx950, measured) refined by one
    // Newton-Raphson step: relative error <= ~2.2e-15, far below the 1e-10 bar of the summed
    // correlators, for one FMA pair less per term and frequency
    double const r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

__constant__ int c_perm3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

struct Freq {
    double z1r, z1i, z2r, z2i, z3r, z3i;
};

__device__ inline Freq load_freq(const double* __restrict__ fz, int f, int perm) {
    // TwoParticleGFPart.cpp:309-315: F = {z1, z2, -z3}, permuted by the part's permutation
    double F[3][2] = {{fz[6 * f + 0], fz[6 * f + 1]}, {fz[6 * f + 2], fz[6 * f + 3]}, {-fz[6 * f + 4], -fz[6 * f + 5]}};
    Freq q;
    int a = c_perm3[perm][0], b = c_perm3[perm][1], c = c_perm3[perm][2];
    q.z1r = F[a][0];
    q.z1i = F[a][1];
    q.z2r = F[b][0];
    q.z2i = F[b][1];
    q.z3r = F[c][0];
    q.z3i = F[c][1];
    return q;
}

// Stage term (fields of the SoA term list) into LDS slots.
//  general path: NR (C, P1, P2, P3); R (Rc, Nc, P1, P2, P3)
//  imaginary-frequency path (every z purely imaginary, z_k = i w_k, a = -P1, c = -P3):
//    kind 0: C, a, b=-P2, c, a*b         kind 1: C, a, b'=-P1-P2-P3, c, a*b'
//    kind 2: Rc, Nc, a, c, a*c, e=-P2-P3, e*e     kind 3: same with e=-P1-P2
template <int KIND, bool IMAG> struct Stage {
    static constexpr int NF = IMAG ? (KIND < 2 ? 5 : 7) : (KIND < 2 ? 4 : 5);
    __device__ static inline void load(const double* __restrict__ T, int64_t cap, int64_t t, double (*sT)[kChunk],
                                       int slot) {
        if(!IMAG) {
#pragma unroll
            for(int k = 0; k < NF; ++k)
                sT[k][slot] = T[k * cap + t];
            return;
        }
        if(KIND < 2) {
            double C = T[t], P1 = T[cap + t], P2 = T[2 * cap + t], P3 = T[3 * cap + t];
            double a = -P1, c = -P3;
            double b = (KIND == 0) ? -P2 : ((-P1) - P2) - P3;
            sT[0][slot] = C;
            sT[1][slot] = a;
            sT[2][slot] = b;
            sT[3][slot] = c;
            sT[4][slot] = a * b;
        } else {
            double Rc = T[t], Nc = T[cap + t], P1 = T[2 * cap + t], P2 = T[3 * cap + t], P3 = T[4 * cap + t];
            double a = -P1, c = -P3;
            double e = (KIND == 3) ? (-P1) - P2 : (-P2) - P3;
            sT[0][slot] = Rc;
            sT[1][slot] = Nc;
            sT[2][slot] = a;
            sT[3][slot] = c;
            sT[4][slot] = a * c;
            sT[5][slot] = e;
            sT[6][slot] = e * e;
        }
    }
};

// KIND 0/1: NonResonantTerm::operator() (TwoParticleGFPart.hpp:475-479)
// KIND 2/3: ResonantTerm::operator() (TwoParticleGFPart.hpp:481-495), resonance |D| <= res
template <int KIND, bool IMAG>
__global__ void __launch_bounds__(256) g2_eval_kernel(int64_t nterms, int64_t slice_len, const double* __restrict__ T,
                                                      int64_t cap, int nfreq, const double* __restrict__ fz, int perm,
                                                      double res, double* __restrict__ partial) {
    using S = Stage<KIND, IMAG>;
    __shared__ double sT[S::NF][kChunk];
    int const fbase = (blockIdx.x * blockDim.x + threadIdx.x) * kFPT;
    int64_t const t0 = int64_t(blockIdx.y) * slice_len;
    int64_t const t1 = min(nterms, t0 + slice_len);
    double const res2 = res * res;
    // per-frequency constants
    double x[kFPT], y[kFPT], w[kFPT], fa[kFPT], fb[kFPT], fc[kFPT], fd[kFPT];
    double zr1[kFPT], zr2[kFPT], zr3[kFPT];
    double ar[kFPT], ai[kFPT];
#pragma unroll
    for(int u = 0; u < kFPT; ++u) {
        Freq q = load_freq(fz, min(fbase + u, nfreq - 1), perm);
        x[u] = q.z1i;
        y[u] = q.z2i;
        w[u] = q.z3i;
        zr1[u] = q.z1r;
        zr2[u] = q.z2r;
        zr3[u] = q.z3r;
        if(IMAG) {
            if(KIND == 0) {
                fa[u] = x[u] * y[u];
            } else if(KIND == 1) {
                fb[u] = (x[u] + y[u]) + w[u]; // Im(z1+z2+z3)
                fa[u] = x[u] * fb[u];
            } else {
                fa[u] = x[u] * w[u];
                fb[u] = (KIND == 3) ? x[u] + y[u] : y[u] + w[u];
                fc[u] = fb[u] * fb[u];
            }
        } else {
            fa[u] = q.z1r + q.z2r;                 // Re(z1+z2)
            fb[u] = q.z1i + q.z2i;                 // Im(z1+z2)
            fc[u] = q.z2r + q.z3r;                 // Re(z2+z3)
            fd[u] = q.z2i + q.z3i;                 // Im(z2+z3)
        }
        ar[u] = 0.0;
        ai[u] = 0.0;
    }
    for(int64_t c0 = t0; c0 < t1; c0 += kChunk) {
        int const cnt = int(min<int64_t>(kChunk, t1 - c0));
        __syncthreads();
        if(threadIdx.x < cnt)
            S::load(T, cap, c0 + threadIdx.x, sT, threadIdx.x);
        __syncthreads();
#pragma unroll 2
        for(int k = 0; k < cnt; ++k) {
            if(IMAG && KIND < 2) {
                double const C = sT[0][k], a = sT[1][k], b = sT[2][k], c = sT[3][k], ab = sT[4][k];
                // four frequencies share one reciprocal: with r = 1/(q0 q1 q2 q3),
                // 1/q0 = q1 (q2 q3) r and so on (|D|^2 lies within ~1e-20..1e10 here, so the
                // product of four stays far inside the double range)
#pragma unroll
                for(int u = 0; u < kFPT; u += 4) {
                    double Dr[4], Di[4], q[4];
#pragma unroll
                    for(int h = 0; h < 4; ++h) {
                        double const yy = (KIND == 0) ? y[u + h] : fb[u + h];
                        double const X = ab - fa[u + h];           // Re((a+ix)(b+iy))
                        double const Y = fma(a, yy, b * x[u + h]); // Im
                        Dr[h] = fma(X, c, -Y * w[u + h]);
                        Di[h] = fma(X, w[u + h], Y * c);
                        q[h] = fma(Dr[h], Dr[h], Di[h] * Di[h]);
                    }
                    double const q01 = q[0] * q[1], q23 = q[2] * q[3];
                    double const r = C * rcp_nr(q01 * q23);
                    double const r01 = q23 * r, r23 = q01 * r;
                    double const s[4] = {q[1] * r01, q[0] * r01, q[3] * r23, q[2] * r23};
#pragma unroll
                    for(int h = 0; h < 4; ++h) {
                        ar[u + h] = fma(s[h], Dr[h], ar[u + h]);
                        ai[u + h] = fma(-s[h], Di[h], ai[u + h]);
                    }
                }
            } else if(IMAG) {
                double const Rc = sT[0][k], Nc = sT[1][k], a = sT[2][k], c = sT[3][k], ac = sT[4][k], e = sT[5][k],
                             e2 = sT[6][k];
                if(e2 > res2) {
                    // |E|^2 = e^2 + Im(E)^2 > res^2 at every frequency: no resonance for this term
                    // (the common case; the branch is uniform because every lane reads term k)
#pragma unroll
                    for(int u = 0; u < kFPT; u += 4) { // four frequencies per reciprocal, as above
                        double Gr[4], Gi[4], q[4];
#pragma unroll
                        for(int h = 0; h < 4; ++h) {
                            double const Qr = ac - fa[u + h];                 // Re((a+ix)(c+iw))
                            double const Qi = fma(a, w[u + h], c * x[u + h]); // Im
                            Gr[h] = fma(e, Qr, -fb[u + h] * Qi);
                            Gi[h] = fma(e, Qi, fb[u + h] * Qr);
                            q[h] = fma(Gr[h], Gr[h], Gi[h] * Gi[h]);
                        }
                        double const q01 = q[0] * q[1], q23 = q[2] * q[3];
                        double const r = Nc * rcp_nr(q01 * q23);
                        double const r01 = q23 * r, r23 = q01 * r;
                        double const sv[4] = {q[1] * r01, q[0] * r01, q[3] * r23, q[2] * r23};
#pragma unroll
                        for(int h = 0; h < 4; ++h) {
                            ar[u + h] = fma(sv[h], Gr[h], ar[u + h]);
                            ai[u + h] = fma(-sv[h], Gi[h], ai[u + h]);
                        }
                    }
                } else {
#pragma unroll
                    for(int u = 0; u < kFPT; ++u) {
                        double const Qr = ac - fa[u];
                        double const Qi = fma(a, w[u], c * x[u]);
                        bool const resonant = e2 + fc[u] <= res2; // |E|^2, E = e + i fb
                        double const Gr = resonant ? Qr : fma(e, Qr, -fb[u] * Qi);
                        double const Gi = resonant ? Qi : fma(e, Qi, fb[u] * Qr);
                        double const K = resonant ? Rc : Nc;
                        double const s = K * rcp_nr(fma(Gr, Gr, Gi * Gi));
                        ar[u] = fma(s, Gr, ar[u]);
                        ai[u] = fma(-s, Gi, ai[u]);
                    }
                }
            } else if(KIND < 2) {
                double const C = sT[0][k], P1 = sT[1][k], P2 = sT[2][k], P3 = sT[3][k];
#pragma unroll
                for(int u = 0; u < kFPT; ++u) {
                    double d1r = zr1[u] - P1, d1i = x[u];
                    double d3r = zr3[u] - P3, d3i = w[u];
                    double d2r, d2i;
                    if(KIND == 0) {
                        d2r = zr2[u] - P2;
                        d2i = y[u];
                    } else {
                        d2r = fa[u] + zr3[u] - P1 - P2 - P3;
                        d2i = fb[u] + w[u];
                    }
                    double xr = d1r * d2r - d1i * d2i, xi = d1r * d2i + d1i * d2r;
                    double Dr = xr * d3r - xi * d3i, Di = xr * d3i + xi * d3r;
                    double s = C * rcp_nr(fma(Dr, Dr, Di * Di));
                    ar[u] = fma(s, Dr, ar[u]);
                    ai[u] = fma(-s, Di, ai[u]);
                }
            } else {
                double const Rc = sT[0][k], Nc = sT[1][k], P1 = sT[2][k], P2 = sT[3][k], P3 = sT[4][k];
#pragma unroll
                for(int u = 0; u < kFPT; ++u) {
                    double d1r = zr1[u] - P1, d1i = x[u];
                    double d3r = zr3[u] - P3, d3i = w[u];
                    double Qr = d1r * d3r - d1i * d3i, Qi = d1r * d3i + d1i * d3r;
                    double Er, Ei;
                    if(KIND == 3) {
                        Er = fa[u] - P1 - P2;
                        Ei = fb[u];
                    } else {
                        Er = fc[u] - P2 - P3;
                        Ei = fd[u];
                    }
                    bool const resonant = fma(Er, Er, Ei * Ei) <= res2;
                    double Gr = resonant ? Qr : (Er * Qr - Ei * Qi);
                    double Gi = resonant ? Qi : (Er * Qi + Ei * Qr);
                    double K = resonant ? Rc : Nc;
                    double s = K * rcp_nr(fma(Gr, Gr, Gi * Gi));
                    ar[u] = fma(s, Gr, ar[u]);
                    ai[u] = fma(-s, Gi, ai[u]);
                }
            }
        }
    }
#pragma unroll
    for(int u = 0; u < kFPT; ++u) {
        int f = fbase + u;
        if(f < nfreq) {
            int64_t o = (int64_t(blockIdx.y) * nfreq + f) * 2;
            partial[o] += ar[u];
            partial[o + 1] += ai[u];
        }
    }
}

// ---- evaluation on grids of shared fixed slots ----------------------------------------------
// When every aligned group of kFPT consecutive triplets shares z1 and z3 (the Matsubara grid of
// the reference's programs: nu outermost, w3, then w2 innermost), a thread's kFPT frequencies
// differ only in F2 = z2.  Under the part's permutation F2 lands in one slot S of the term, and
// for most (kind, S) only one denominator factor V = a + i y depends on it; the others form a
// fixed product G per term and thread.  Then
//   term = Kc / (G V) = K conj(V) / |V|^2,  K = Kc / G  (once per term and thread),
// so a frequency costs |V|^2, a quarter of a shared reciprocal, two FMAs for K conj(V) and two
// for the accumulation (about 9 FP64 instructions instead of 14).  Kinds and slots:
//   kind 0 (D1 D2 D3): any S, V = D_S;               kind 1 (D1 D2' D3): S = 1, V = D2';
//   kind 2 (D1 D3 [E = z2+z3-P2-P3]): S = 0, V = D1 (E fixed); S = 1, V = E;
//   kind 3 (D1 D3 [E = z1+z2-P1-P2]): S = 1, V = E; S = 2, V = D3 (E fixed).
// Resonance (|E| <= res): with E fixed it selects the coefficient and whether E enters G; with E
// varying, terms with (P-sum)^2 > res^2 can never resonate (uniform fast path), the others take
// a per-frequency path.
__host__ __device__ constexpr bool g2_structured(int kind, int s) {
    return kind == 0 || (kind == 1 && s == 1) || (kind == 2 && s <= 1) || (kind == 3 && s >= 1);
}

struct Cplx {
    double r, i;
};
__device__ inline Cplx cx_mul(Cplx a, Cplx b) { return Cplx{fma(a.r, b.r, -a.i * b.i), fma(a.r, b.i, a.i * b.r)}; }

// STRIDED: the thread's frequencies are f = t + u * stride (u < kFPT) and they share z2 and z3
// (grids whose outer index is the bosonic frequency): then F1 = z1 is the varying one, and S is
// its slot.  Otherwise f = kFPT t + u, sharing z1 and z3, with F2 varying.
template <int KIND, int S, bool STRIDED>
__global__ void __launch_bounds__(256) g2_eval_s_kernel(int64_t nterms, int64_t slice_len, const double* __restrict__ T,
                                                        int64_t cap, int nfreq, int stride, const double* __restrict__ fz,
                                                        int perm, double res, double* __restrict__ partial) {
    constexpr int NF = KIND < 2 ? 4 : 5;           // NR (C, P1, P2, P3); R (Rc, Nc, P1, P2, P3)
    constexpr bool EVAR = (KIND == 2 && S == 1) || (KIND == 3 && S == 1); // V = E
    __shared__ double sT[NF][kChunk];
    int const tid = blockIdx.x * blockDim.x + threadIdx.x;
    int const fbase = STRIDED ? tid : tid * kFPT, fstep = STRIDED ? stride : 1;
    bool const live = STRIDED ? tid < stride : fbase < nfreq;
    int64_t const t0 = int64_t(blockIdx.y) * slice_len;
    int64_t const t1 = min(nterms, t0 + slice_len);
    double const res2 = res * res;
    // imaginary parts of the fixed slots (shared by the group) and of the varying factor
    Freq const q0 = load_freq(fz, live ? fbase : 0, perm);
    double const x1 = q0.z1i, x2 = q0.z2i, x3 = q0.z3i;
    double yv[kFPT], ar[kFPT], ai[kFPT];
#pragma unroll
    for(int u = 0; u < kFPT; ++u) {
        Freq const q = load_freq(fz, live ? min(fbase + u * fstep, nfreq - 1) : 0, perm);
        if(KIND == 1)
            yv[u] = (q.z1i + q.z2i) + q.z3i; // Im(z1 + z2 + z3)
        else if(KIND == 2 && S == 1)
            yv[u] = q.z2i + q.z3i; // Im E
        else if(KIND == 3 && S == 1)
            yv[u] = q.z1i + q.z2i; // Im E
        else
            yv[u] = (S == 0) ? q.z1i : (S == 1) ? q.z2i : q.z3i;
        ar[u] = 0.0;
        ai[u] = 0.0;
    }
    for(int64_t c0 = t0; c0 < t1; c0 += kChunk) {
        int const cnt = int(min<int64_t>(kChunk, t1 - c0));
        __syncthreads();
        if(int(threadIdx.x) < cnt)
#pragma unroll
            for(int f = 0; f < NF; ++f)
                sT[f][threadIdx.x] = T[f * cap + c0 + threadIdx.x];
        __syncthreads();
#pragma unroll 2
        for(int k = 0; k < cnt; ++k) {
            double const P1 = sT[NF - 3][k], P2 = sT[NF - 2][k], P3 = sT[NF - 1][k];
            Cplx const D1{-P1, x1}, D3{-P3, x3};
            double Kc, av;
            Cplx G;
            if(KIND == 0) {
                Cplx const D2{-P2, x2};
                Kc = sT[0][k];
                av = (S == 0) ? -P1 : (S == 1) ? -P2 : -P3;
                G = (S == 0) ? cx_mul(D2, D3) : (S == 1) ? cx_mul(D1, D3) : cx_mul(D1, D2);
            } else if(KIND == 1) {
                Kc = sT[0][k];
                av = ((-P1) - P2) - P3;
                G = cx_mul(D1, D3);
            } else if(EVAR) {
                Kc = sT[1][k]; // non-resonant coefficient; resonant terms are handled below
                av = (KIND == 3) ? (-P1) - P2 : (-P2) - P3;
                G = cx_mul(D1, D3);
            } else {
                // E fixed per thread: resonance selects the coefficient and whether E enters G
                Cplx const E = (KIND == 2) ? Cplx{(-P2) - P3, x2 + x3} : Cplx{(-P1) - P2, x1 + x2};
                Cplx const F = (KIND == 2) ? D3 : D1;
                bool const resonant = fma(E.r, E.r, E.i * E.i) <= res2;
                Kc = resonant ? sT[0][k] : sT[1][k];
                av = (KIND == 2) ? -P1 : -P3;
                G = resonant ? F : cx_mul(F, E);
            }
            // K = Kc / G
            double const ig = Kc * rcp_nr(fma(G.r, G.r, G.i * G.i));
            double const Kr = G.r * ig, Ki = -G.i * ig;
            double const a2 = av * av, Kra = Kr * av, Kia = Ki * av;
            if(!EVAR || a2 > res2) {
#pragma unroll
                for(int u = 0; u < kFPT; u += 4) {
                    double q[4];
#pragma unroll
                    for(int h = 0; h < 4; ++h)
                        q[h] = fma(yv[u + h], yv[u + h], a2); // |V|^2
                    double const q01 = q[0] * q[1], q23 = q[2] * q[3];
                    double const r = rcp_nr(q01 * q23);
                    double const r01 = q23 * r, r23 = q01 * r;
                    double const s[4] = {q[1] * r01, q[0] * r01, q[3] * r23, q[2] * r23};
#pragma unroll
                    for(int h = 0; h < 4; ++h) {
                        ar[u + h] = fma(s[h], fma(Ki, yv[u + h], Kra), ar[u + h]);  // Re(K conj(V)) / |V|^2
                        ai[u + h] = fma(s[h], fma(-Kr, yv[u + h], Kia), ai[u + h]); // Im(K conj(V)) / |V|^2
                    }
                }
            } else {
                // V = E may resonate at some frequency: Rc / G there, Nc / (G E) elsewhere
                double const igr = sT[0][k] * rcp_nr(fma(G.r, G.r, G.i * G.i));
                double const Rr = G.r * igr, Ri = -G.i * igr;
#pragma unroll
                for(int u = 0; u < kFPT; ++u) {
                    double const qv = fma(yv[u], yv[u], a2);
                    if(qv <= res2) {
                        ar[u] += Rr;
                        ai[u] += Ri;
                    } else {
                        double const s = rcp_nr(qv);
                        ar[u] = fma(s, fma(Ki, yv[u], Kra), ar[u]);
                        ai[u] = fma(s, fma(-Kr, yv[u], Kia), ai[u]);
                    }
                }
            }
        }
    }
#pragma unroll
    for(int u = 0; u < kFPT; ++u) {
        int f = fbase + u * fstep;
        if(live && f < nfreq) {
            int64_t o = (int64_t(blockIdx.y) * nfreq + f) * 2;
            partial[o] += ar[u];
            partial[o + 1] += ai[u];
        }
    }
}

// ---- evaluation on separable grids (matrix cores) --------------------------------------------
// Grids of the reference's programs: triplet f = (b NR + r) NC + c with z1 depending on (b, r),
// z2 on c and z3 on r only (quantum_model.cpp:220-231: (W + w3, w2, w3), W outermost), all purely
// imaginary.  Then F1 = z1 is a (b, r) quantity, F2 = z2 a c quantity and F3 = -z3 an r quantity,
// and every non-resonant term factorizes per bosonic index b into a row factor (its r-dependent
// denominators) times a column factor (its c-dependent ones, times C):
//   kind 0 (D1 D2 D3): the slot holding F2 is the column factor, the other two the row factor;
//   kind 1 (D1 D2' D3, D2' = z1 + z2 + z3 - P1 - P2 - P3, Im(z1 + z2 + z3) = Y_b + Im F2(c) with
//          Y_b = Im(F1 + F3) the same for every r): D2' and the slot holding F2 (if any) form the
//          column factor.
// So V_b[r][c] = sum_t Row_t(b, r) Col_t(b, c): a complex GEMM over the terms, on the FP64 matrix
// cores (3 real v_mfma_f64_16x16x4f64 per 16 x 16 tile and 4 terms, the complex product formed
// from three real ones), 6 FLOP per term x triplet against ~16 on the vector pipe.  Each wave owns one K-slice of the terms and the whole
// 32 x 32 (r, c) tile of one b; its lane (i, k) = (lane & 15, lane >> 4) forms the row and
// column elements of term k of the step for rows/columns i and i + 16 (one shared reciprocal per
// element).  Partial tiles per wave are summed in a fixed order afterwards.
constexpr int kMT = 32; // tile rows / columns (NR, NC <= 32)

// Row / column factors of a non-resonant kind under a permutation (slot s holds F[perm[s]]):
// pole 0, 1, 2 = P1, P2, P3, 3 = P1 + P2 + P3; y source 0 = F1(b, r), 1 = F2(c), 2 = F3(r),
// 3 = Y_b + F2(c).
struct SepPlan {
    int nrow, ncol;
    int rp[2], rs[2], cp[2], cs[2];
};
__host__ __device__ constexpr SepPlan sep_plan_of(int kind, int perm) {
    int const p3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    SepPlan pl{0, 0, {0, 0}, {2, 2}, {0, 0}, {1, 1}};
    if(kind == 0) {
        for(int sl = 0; sl < 3; ++sl) {
            if(p3[perm][sl] == 1) {
                pl.cp[0] = sl;
                pl.cs[0] = 1;
            } else {
                pl.rp[pl.nrow] = sl;
                pl.rs[pl.nrow] = p3[perm][sl];
                ++pl.nrow;
            }
        }
        pl.ncol = 1;
    } else if(kind >= 2) {
        // E = z2 + z3 (kind 2) or z1 + z2 (kind 3) is F1 + F3 here, a function of b alone; D1 and
        // D3 are one row and one column factor
        int const sl_c = p3[perm][0] == 1 ? 0 : 2, sl_r = 2 - sl_c;
        pl.cp[0] = sl_c;
        pl.cs[0] = 1;
        pl.rp[0] = sl_r;
        pl.rs[0] = p3[perm][sl_r];
        pl.nrow = 1;
        pl.ncol = 1;
    } else {
        pl.cp[0] = 3; // z1 + z2 + z3 - P1 - P2 - P3
        pl.cs[0] = 3;
        if(p3[perm][1] == 1) { // F2 inside the sum only
            pl.rp[0] = 0;
            pl.rs[0] = p3[perm][0];
            pl.rp[1] = 2;
            pl.rs[1] = p3[perm][2];
            pl.nrow = 2;
            pl.ncol = 1;
        } else {
            int const sc = p3[perm][0] == 1 ? 0 : 2, sr = 2 - sc;
            pl.cp[1] = sc;
            pl.cs[1] = 1;
            pl.rp[0] = sr;
            pl.rs[0] = p3[perm][sr];
            pl.nrow = 1;
            pl.ncol = 2;
        }
    }
    return pl;
}

// ---- three-factor resonant terms split into a separable and a diagonal part -----------------
// The resonant kinds under the permutations where one of D1, D3 and E depend on F2 both, and E on
// F2 plus F1 or F3 (kind 2 under (F1, F3, F2) and (F3, F1, F2), kind 3 under (F2, F1, F3) and
// (F2, F3, F1)), have three frequency-dependent factors.  With x = F2 and the two F2 factors
// (x - a)(x - b), 1 / ((x - a)(x - b)) = (1 / (x - a) - 1 / (x - b)) / (a - b), where a - b
// = F1 - P or F3 - P (never zero: Matsubara F1, F3 are odd multiples of i pi / beta).  Then
//   Nc / (D1 D3 E) = Nc / ((s0 - P1)(s1 - P2)(s2 - P3))    [a kind-0 term, same permutation]
//                  - Nc / ((F1 - Pa)(F3 - Pb) E)           [a diagonal term, E unchanged],
//   kind 2, perm 1:  a - b = F3 - P2, Pa = P1, Pb = P2, E = F2 + F3 - P2 - P3 (d grid)
//   kind 2, perm 4:  a - b = F1 - P2, Pa = P2, Pb = P1, E = F1 + F2 - P2 - P3 (s grid)
//   kind 3, perm 2:  a - b = F1 - P2, Pa = P2, Pb = P3, E = F1 + F2 - P1 - P2 (s grid)
//   kind 3, perm 3:  a - b = F3 - P2, Pa = P3, Pb = P2, E = F2 + F3 - P1 - P2 (d grid)
// (checked against the direct formula on the CPU).  The diagonal kernel (dkind, dperm) forms
// Pa = Q1, Pb = Q3 and E's real part -(Q2 + Q3) (dkind 2) or -(Q1 + Q2) (dkind 3), so the split
// feeds it Q = (P[q0], P[q1], P[q2]); E's real part is then the same rounded sum of two poles.
// Only terms that can never resonate on a purely imaginary grid ((P sum)^2 > res^2, so that
// |E|^2 > res^2 at every frequency) are split; the others keep the grouped vector kernel.
struct SplitPlan {
    int dkind, dperm, q[3];
};
__host__ __device__ constexpr SplitPlan split_plan_of(int kind, int perm) {
    return (kind == 2 && perm == 1)   ? SplitPlan{2, 0, {0, 2, 1}}
           : (kind == 2 && perm == 4) ? SplitPlan{3, 0, {1, 2, 0}}
           : (kind == 3 && perm == 2) ? SplitPlan{3, 0, {1, 0, 2}}
           : (kind == 3 && perm == 3) ? SplitPlan{2, 0, {2, 0, 1}}
                                      : SplitPlan{0, 0, {0, 1, 2}};
}
__host__ __device__ constexpr bool split_kind_perm(int kind, int perm) {
    return (kind == 2 && (perm == 1 || perm == 4)) || (kind == 3 && (perm == 2 || perm == 3));
}
// E's real part is -(P2 + P3) (kind 2) or -(P1 + P2) (kind 3); its square above res^2 keeps |E|
// above res at every purely imaginary frequency
template <int KIND>
__host__ __device__ inline bool never_resonant(double P1, double P2, double P3, double res2) {
    double const e = (KIND == 2) ? P2 + P3 : P1 + P2;
    return e * e > res2;
}

// Term staging of the matrix-core kernels.  A wave works through its slice of terms in chunks of
// kStage; each lane loads one term of the next chunk while the current one is evaluated, and at the
// chunk boundary the wave stores its loads into its own LDS slots, from which every step reads its
// 4 terms (16 lanes per address).  The global loads thus have a whole chunk to arrive.  Staged
// slots: 0 = C (Rc), 1 = Nc (resonant lists only), 2..4 = P1, P2, P3, and for tangent lists
// (NR: C, P1, P2, P3, dC, dP1, dP2, dP3; R: Rc, Nc, P1, P2, P3, dRc, dNc, dP1, dP2, dP3) 5 = dC
// (dRc), 6 = dNc, 7..9 = dP1, dP2, dP3; terms past the slice get zero coefficients.
constexpr int kStage = 64;

template <bool RES, bool TAN = false>
__device__ inline void stage_load(const double* __restrict__ T, int64_t cap, int64_t t, int64_t t0, int64_t t1,
                                  double f[]) {
    constexpr int F0 = RES ? 1 : 0;
    constexpr int NV = RES ? 5 : 4; // value fields
    bool const tv = t < t1;
    int64_t const tt = tv ? t : t0;
    f[0] = tv ? T[tt] : 0.0;
    f[1] = (RES && tv) ? T[cap + tt] : 0.0;
    f[2] = T[(F0 + 1) * cap + tt];
    f[3] = T[(F0 + 2) * cap + tt];
    f[4] = T[(F0 + 3) * cap + tt];
    if constexpr(TAN) {
        f[5] = tv ? T[NV * cap + tt] : 0.0;
        f[6] = (RES && tv) ? T[(NV + 1) * cap + tt] : 0.0;
        f[7] = T[(NV + F0 + 1) * cap + tt];
        f[8] = T[(NV + F0 + 2) * cap + tt];
        f[9] = T[(NV + F0 + 3) * cap + tt];
    }
}

// The LDS operations of one wave are carried out in issue order, so the previous chunk's reads are
// served before these writes and the writes before the next chunk's reads; the fences keep the
// compiler from moving accesses across the boundary.
template <int NS = 5>
__device__ inline void stage_store(double (*W)[kStage], int lane, double const f[]) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for(int x = 0; x < NS; ++x)
        W[x][lane] = f[x];
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int W>
__device__ inline double sep_pole(double P1, double P2, double P3) {
    if constexpr(W == 0)
        return P1;
    else if constexpr(W == 1)
        return P2;
    else if constexpr(W == 2)
        return P3;
    else
        return (P1 + P2) + P3;
}

// element = scale / prod_f (i y_f - P_f) = scale prod_f (-P_f - i y_f) / prod_f (P_f^2 + y_f^2);
// y2 = y^2 and, for two factors, yy = y_a y_b are the lane's precomputed constants
template <int NF>
__device__ inline void sep_element(double Pa, double Pb, double ya, double yb, double ya2, double yb2, double yy,
                                   double PaPb, double scale, double& er, double& ei) {
    if constexpr(NF == 1) {
        double const r = scale * rcp_nr(fma(Pa, Pa, ya2));
        er = -Pa * r;
        ei = -ya * r;
    } else {
        double const r = scale * rcp_nr(fma(Pa, Pa, ya2) * fma(Pb, Pb, yb2));
        er = (PaPb - yy) * r;
        ei = fma(Pa, yb, Pb * ya) * r;
    }
}

// The element and its derivative along the poles' derivatives dPa, dPb:
//   E = scale / prod_f (i y_f - P_f),  E' = E sum_f dP_f / (i y_f - P_f),
// with 1 / (i y_f - P_f) = (-P_f - i y_f) / |.|^2 from the same shared reciprocal.
template <int NF>
__device__ inline void sep_element_tan(double Pa, double Pb, double dPa, double dPb, double ya, double yb, double ya2,
                                       double yb2, double yy, double PaPb, double scale, double& er, double& ei,
                                       double& dr, double& di) {
    if constexpr(NF == 1) {
        double const rr = rcp_nr(fma(Pa, Pa, ya2));
        double const ur = -Pa * rr, ui = -ya * rr; // 1 / (i y - P)
        er = scale * ur;
        ei = scale * ui;
        double const sr = dPa * ur, si = dPa * ui; // sigma
        dr = fma(er, sr, -ei * si);
        di = fma(er, si, ei * sr);
    } else {
        double const ma = fma(Pa, Pa, ya2), mb = fma(Pb, Pb, yb2);
        double const rr = rcp_nr(ma * mb);
        double const r = scale * rr;
        er = (PaPb - yy) * r;
        ei = fma(Pa, yb, Pb * ya) * r;
        double const wa = dPa * mb, wb = dPb * ma;
        double const sr = -rr * fma(wa, Pa, wb * Pb), si = -rr * fma(wa, ya, wb * yb);
        dr = fma(er, sr, -ei * si);
        di = fma(er, si, ei * sr);
    }
}

// Resonant kinds (2: E = z2 + z3 - P2 - P3, 3: E = z1 + z2 - P1 - P2) when E = i Y_b - (P sum):
// the column element carries the coefficient K = Rc where |E| <= res, else Nc / E.
// Occupancy: left to itself the compiler puts the accumulators in 96 AGPRs on top of ~180 VGPRs,
// one wave per SIMD, so the element arithmetic and the term loads of a wave stall its matrix
// instructions.  Asking for two waves per SIMD keeps everything in 184-196 VGPRs without spills;
// with the terms staged a chunk ahead, two waves keep the matrix cores as busy as three did
// (three fit in 168 VGPRs with the lane constants' squares recomputed, and measured no faster).
// SRC = KIND: the list holds terms of kind KIND.  SRC = 2 or 3 with KIND = 0: the list holds
// resonant terms of that kind, read as the kind-0 term Nc / ((s0 - P1)(s1 - P2)(s2 - P3)), the
// separable half of their partial-fraction split (split_applies); terms that may resonate
// (never_resonant false) get coefficient 0 here and are evaluated elsewhere.
// TAN: the list is a tangent list and the kernel forms the derivative instead of the value: with
// Row' and Col' the elements' derivatives (sep_element_tan; the column's coefficient C, or K = Rc
// or Nc / E with dK = (dNc + K dPe) / E, carries its own), the tile is sum_t Row'_t Col_t +
// Row_t Col'_t, two complex products per term.
template <int KIND, int PERM, int SRC = KIND, bool TAN = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) g2_eval_mfma_kernel(int64_t nterms, int64_t slice_len, const double* __restrict__ T,
                                                           int64_t cap, int NR, int NC, const double* __restrict__ yF1,
                                                           const double* __restrict__ yF2, const double* __restrict__ yF3,
                                                           const double* __restrict__ yB, double res2,
                                                           double* __restrict__ partial) {
    constexpr SepPlan PL = sep_plan_of(KIND, PERM);
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int const b = blockIdx.y;
    int64_t const slice = int64_t(blockIdx.x) * 4 + wave;
    int64_t const t0 = slice * slice_len, t1 = min(nterms, t0 + slice_len);
    int const i = lane & 15, k = lane >> 4;
    // this lane's factor arguments for rows / columns i and i + 16
    double ya[2], yb[2], ya2[2], yb2[2], yy[2], xa[2], xb[2], xa2[2], xb2[2], xx[2];
    double rs[2], cs[2]; // 1 or 0 (rows / columns beyond NR / NC)
#pragma unroll
    for(int m = 0; m < 2; ++m) {
        int const r = min(i + 16 * m, NR - 1), c = min(i + 16 * m, NC - 1);
        rs[m] = i + 16 * m < NR ? 1.0 : 0.0;
        cs[m] = i + 16 * m < NC ? 1.0 : 0.0;
        ya[m] = PL.rs[0] == 0 ? yF1[b * NR + r] : yF3[r];
        yb[m] = PL.rs[1] == 0 ? yF1[b * NR + r] : yF3[r];
        xa[m] = PL.cs[0] == 1 ? yF2[c] : yB[b] + yF2[c];
        xb[m] = PL.cs[1] == 1 ? yF2[c] : yB[b] + yF2[c];
        ya2[m] = ya[m] * ya[m];
        yb2[m] = yb[m] * yb[m];
        yy[m] = ya[m] * yb[m];
        xa2[m] = xa[m] * xa[m];
        xb2[m] = xb[m] * xb[m];
        xx[m] = xa[m] * xb[m];
    }
    typedef double dbl4 __attribute__((ext_vector_type(4)));
    // complex products from three real ones: with A1 = ar + ai, B2 = bi - br, B3 = br + bi,
    // Re = sum A1 br - sum ai B3, Im = sum A1 br + sum ar B2
    dbl4 acc[2][2][3];
#pragma unroll
    for(int m = 0; m < 2; ++m)
#pragma unroll
        for(int n = 0; n < 2; ++n)
#pragma unroll
            for(int x = 0; x < 3; ++x)
                acc[m][n][x] = dbl4{0.0, 0.0, 0.0, 0.0};
    // term k of each step of 4, read from the wave's staged chunk
    double const Yb = yB[b], Yb2 = Yb * Yb;
    constexpr int NS = TAN ? 10 : 5; // staged slots per term
    __shared__ double sT[4][NS][kStage];
    double(*W)[kStage] = sT[wave];
    double nf[NS];
    if(t0 < t1)
        stage_load<(SRC >= 2), TAN>(T, cap, t0 + lane, t0, t1, nf);
    for(int64_t c0 = t0; c0 < t1; c0 += kStage) {
        stage_store<NS>(W, lane, nf);
        if(c0 + kStage < t1)
            stage_load<(SRC >= 2), TAN>(T, cap, c0 + kStage + lane, t0, t1, nf);
        int const nsteps = (int(min<int64_t>(kStage, t1 - c0)) + 3) / 4;
        for(int j = 0; j < nsteps; ++j) {
            int const q = 4 * j + k;
            double const Cc = W[0][q], Nc = W[1][q], Q1 = W[2][q], Q2 = W[3][q], Q3 = W[4][q];
            double dKr = 0.0, dKi = 0.0, dQ1 = 0.0, dQ2 = 0.0, dQ3 = 0.0;
            if constexpr(TAN) {
                dKr = W[5][q];
                dQ1 = W[7][q];
                dQ2 = W[8][q];
                dQ3 = W[9][q];
            }
            // coefficient of the column element: C (non-resonant) or K = Rc / Nc / E (resonant)
            double Kr = Cc, Ki = 0.0;
            if constexpr(KIND == 0 && SRC >= 2) {
                bool const nrs = never_resonant<SRC>(Q1, Q2, Q3, res2);
                Kr = nrs ? Nc : 0.0;
                if constexpr(TAN)
                    dKr = nrs ? W[6][q] : 0.0;
            }
            if constexpr(KIND >= 2) {
                double const Er = -(KIND == 2 ? Q2 + Q3 : Q1 + Q2); // E = Er + i Yb
                double const e2 = fma(Er, Er, Yb2);
                if(!(e2 <= res2)) {
                    double const ie = Nc * rcp_nr(e2);
                    Kr = Er * ie;
                    Ki = -Yb * ie;
                    if constexpr(TAN) { // dK = (dNc + K dPe) / E
                        double const rr = rcp_nr(e2), dPe = KIND == 2 ? dQ2 + dQ3 : dQ1 + dQ2;
                        double const nr = fma(Kr, dPe, W[6][q]), ni = Ki * dPe;
                        dKr = fma(nr, Er, ni * Yb) * rr;
                        dKi = fma(ni, Er, -nr * Yb) * rr;
                    }
                }
            }
            double const Ra = sep_pole<PL.rp[0]>(Q1, Q2, Q3), Rb = sep_pole<PL.rp[1]>(Q1, Q2, Q3);
            double const Ka = sep_pole<PL.cp[0]>(Q1, Q2, Q3), Kb = sep_pole<PL.cp[1]>(Q1, Q2, Q3);
            double const RaRb = Ra * Rb, KaKb = Ka * Kb;
            double ar[2], ai[2], br[2], bi[2];
            double dar[2], dai[2], dbr[2], dbi[2]; // derivatives (TAN)
    #pragma unroll
            for(int m = 0; m < 2; ++m) {
                if constexpr(TAN) {
                    double const dRa = sep_pole<PL.rp[0]>(dQ1, dQ2, dQ3), dRb = sep_pole<PL.rp[1]>(dQ1, dQ2, dQ3);
                    double const dKa = sep_pole<PL.cp[0]>(dQ1, dQ2, dQ3), dKb = sep_pole<PL.cp[1]>(dQ1, dQ2, dQ3);
                    sep_element_tan<PL.nrow>(Ra, Rb, dRa, dRb, ya[m], yb[m], ya2[m], yb2[m], yy[m], RaRb, rs[m], ar[m],
                                             ai[m], dar[m], dai[m]);
                    double pr, pi, dpr, dpi; // column element without its coefficient, and its derivative
                    sep_element_tan<(KIND >= 2 ? 1 : PL.ncol)>(Ka, Kb, dKa, dKb, xa[m], xb[m], xa2[m], xb2[m], xx[m],
                                                               KaKb, cs[m], pr, pi, dpr, dpi);
                    br[m] = fma(Kr, pr, -Ki * pi);
                    bi[m] = fma(Kr, pi, Ki * pr);
                    dbr[m] = fma(dKr, pr, -dKi * pi) + fma(Kr, dpr, -Ki * dpi);
                    dbi[m] = fma(dKr, pi, dKi * pr) + fma(Kr, dpi, Ki * dpr);
                } else {
                    sep_element<PL.nrow>(Ra, Rb, ya[m], yb[m], ya2[m], yb2[m], yy[m], RaRb, rs[m], ar[m], ai[m]);
                    if constexpr(KIND >= 2) {
                        double er, ei;
                        sep_element<1>(Ka, Kb, xa[m], xb[m], xa2[m], xb2[m], xx[m], KaKb, cs[m], er, ei);
                        br[m] = fma(Kr, er, -Ki * ei);
                        bi[m] = fma(Kr, ei, Ki * er);
                    } else {
                        sep_element<PL.ncol>(Ka, Kb, xa[m], xb[m], xa2[m], xb2[m], xx[m], KaKb, cs[m] * Kr, br[m],
                                             bi[m]);
                    }
                }
            }
            double a1[2], b2[2], b3[2];
    #pragma unroll
            for(int m = 0; m < 2; ++m) {
                a1[m] = ar[m] + ai[m];
                b2[m] = bi[m] - br[m];
                b3[m] = br[m] + bi[m];
            }
            if constexpr(TAN) { // Row' Col + Row Col'
                double da1[2], db2[2], db3[2];
    #pragma unroll
                for(int m = 0; m < 2; ++m) {
                    da1[m] = dar[m] + dai[m];
                    db2[m] = dbi[m] - dbr[m];
                    db3[m] = dbr[m] + dbi[m];
                }
    #pragma unroll
                for(int m = 0; m < 2; ++m)
    #pragma unroll
                    for(int n = 0; n < 2; ++n) {
                        acc[m][n][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(da1[m], br[n], acc[m][n][0], 0, 0, 0);
                        acc[m][n][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(dai[m], b3[n], acc[m][n][1], 0, 0, 0);
                        acc[m][n][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(dar[m], b2[n], acc[m][n][2], 0, 0, 0);
                        acc[m][n][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[m], dbr[n], acc[m][n][0], 0, 0, 0);
                        acc[m][n][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[m], db3[n], acc[m][n][1], 0, 0, 0);
                        acc[m][n][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[m], db2[n], acc[m][n][2], 0, 0, 0);
                    }
            } else {
    #pragma unroll
                for(int m = 0; m < 2; ++m)
    #pragma unroll
                    for(int n = 0; n < 2; ++n) {
                        acc[m][n][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[m], br[n], acc[m][n][0], 0, 0, 0);
                        acc[m][n][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[m], b3[n], acc[m][n][1], 0, 0, 0);
                        acc[m][n][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[m], b2[n], acc[m][n][2], 0, 0, 0);
                    }
            }
        }
    }
    // D[i][j] of a 16 x 16 tile sits in lane j + 16 (i % 4), register i / 4
    double* out = partial + (slice * gridDim.y + b) * (kMT * kMT * 2);
#pragma unroll
    for(int m = 0; m < 2; ++m)
#pragma unroll
        for(int n = 0; n < 2; ++n)
#pragma unroll
            for(int x = 0; x < 4; ++x) {
                int const row = 16 * m + (lane >> 4) + 4 * x, col = 16 * n + (lane & 15);
                out[(row * kMT + col) * 2] = acc[m][n][0][x] - acc[m][n][1][x];
                out[(row * kMT + col) * 2 + 1] = acc[m][n][0][x] + acc[m][n][2][x];
            }
}

// acc[(b NR + r) NC + c] += sum over slices of the (b, r, c) entries, in slice order
__global__ void g2_mfma_reduce_kernel(int nslices, int NB, int NR, int NC, const double* __restrict__ partial,
                                      double* __restrict__ acc) {
    int const f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= NB * NR * NC)
        return;
    int const c = f % NC, r = (f / NC) % NR, b = f / (NC * NR);
    double sr = 0.0, si = 0.0;
    for(int sl = 0; sl < nslices; ++sl) {
        const double* p = partial + ((int64_t(sl) * NB + b) * kMT * kMT + r * kMT + c) * 2;
        sr += p[0];
        si += p[1];
    }
    acc[2 * f] += sr;
    acc[2 * f + 1] += si;
}

// ---- resonant terms with E along the diagonals of the (r, c) plane (matrix cores) ----------
// On the same grids, the resonant kinds under the permutations (F1, F2, F3) and (F3, F2, F1)
// have one factor of D1 D3 on F1(b, r) and the other on F3(r), and E on F2 + F3 (kind 2 under
// the first, kind 3 under the second: Im E = Im F2(c) + Im F3(r), a function of d = c - r on a
// Matsubara grid) or on F1 + F2 (the other two: Im E a function of s = b + r + c).  Then
//   term(b, r, c) = Row_t(b, r) H_t(e),  Row_t = 1 / ((F1 - Pa)(F3 - Pb)),
//   H_t(e) = Rc where |E| <= res, else Nc / E,
// with e = d or s, and for the 16 rows r0 .. r0 + 15 of one b the needed e form 16 + NC - 1 <= 47
// consecutive values.  A wave forms the 16 x 48 product of the rows and the 48 diagonal values
// H_t (3 column tiles; about 1.5 times the needed entries) over its terms; entry (i, q) holds
// output column c = q + i - 15 (d grids) or c = q - i (s grids) of row r0 + i.  Im E for index e
// comes from the table yE (checked on the host against every (b, r, c) of the grid).
constexpr int kDQ = 48; // diagonal tile columns

// SK = 0: the list holds resonant terms of kind KIND under permutation PERM.  SK = 2 or 3: it holds
// terms of kind SK under permutation SP, read as the diagonal half of their partial-fraction split
// (split_applies): Row = -Nc / ((F1 - Pa)(F3 - Pb)), E never resonant, fields (Q1, Q2, Q3)
// arranged so that this kernel's Pa, Pb and E are the split's; terms that may resonate get
// coefficient 0.  TAN: a tangent list, and the tile is sum_t Row'_t H_t + Row_t H'_t (the
// derivative), with H' = dRc (|E| <= res) or (dNc + H dPe) / E.
template <int KIND, int PERM, int SK = 0, int SP = 0, bool TAN = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) g2_eval_diag_kernel(
    int64_t nterms, int64_t slice_len, const double* __restrict__ T, int64_t cap, int NR, int nrt,
    const double* __restrict__ yF1, const double* __restrict__ yF3, const double* __restrict__ yE, int nE, double res2,
    double* __restrict__ partial) {
    static_assert((KIND == 2 || KIND == 3) && (PERM == 0 || PERM == 5), "no diagonal plan");
    constexpr bool TOEP = (KIND == 2) == (PERM == 0); // E on F2 + F3 (index d), else on F1 + F2 (index s)
    // pole of the F1 factor and of the F3 factor: D1 = F[perm 0] - P1, D3 = F[perm 2] - P3
    constexpr bool F1_IN_D1 = PERM == 0;
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int const b = int(blockIdx.y) / nrt, r0 = 16 * (int(blockIdx.y) % nrt);
    int64_t const slice = int64_t(blockIdx.x) * 4 + wave;
    int64_t const t0 = slice * slice_len, t1 = min(nterms, t0 + slice_len);
    int const i = lane & 15, k = lane >> 4;
    // row element of row r0 + i (zero beyond NR) and diagonal values of columns q = i + 16 n
    int const rr = min(r0 + i, NR - 1);
    double const rsc = r0 + i < NR ? 1.0 : 0.0;
    double const ya = yF1[b * NR + rr], yb = yF3[rr];
    double const ya2 = ya * ya, yb2 = yb * yb, yy = ya * yb;
    double ye[3];
#pragma unroll
    for(int n = 0; n < 3; ++n) {
        int const q = i + 16 * n;
        int const e = TOEP ? q - 15 - r0 + (NR - 1) : b + r0 + q;
        ye[n] = yE[min(max(e, 0), nE - 1)];
    }
    typedef double dbl4 __attribute__((ext_vector_type(4)));
    dbl4 acc[3][3];
#pragma unroll
    for(int n = 0; n < 3; ++n)
#pragma unroll
        for(int x = 0; x < 3; ++x)
            acc[n][x] = dbl4{0.0, 0.0, 0.0, 0.0};
    // resonant SoA (Rc, Nc, P1, P2, P3; tangent lists also dRc, dNc, dP1, dP2, dP3); term k of each
    // step of 4, read from the staged chunk
    constexpr int NS = TAN ? 10 : 5;
    __shared__ double sT[4][NS][kStage];
    double(*W)[kStage] = sT[wave];
    double nf[NS];
    if(t0 < t1)
        stage_load<true, TAN>(T, cap, t0 + lane, t0, t1, nf);
    for(int64_t c0 = t0; c0 < t1; c0 += kStage) {
        stage_store<NS>(W, lane, nf);
        if(c0 + kStage < t1)
            stage_load<true, TAN>(T, cap, c0 + kStage + lane, t0, t1, nf);
        int const nsteps = (int(min<int64_t>(kStage, t1 - c0)) + 3) / 4;
        for(int j = 0; j < nsteps; ++j) {
            int const q = 4 * j + k;
            double Rc = W[0][q], Nc = W[1][q], Q1 = W[2][q], Q2 = W[3][q], Q3 = W[4][q];
            double dRc = 0.0, dNc = 0.0, dQ1 = 0.0, dQ2 = 0.0, dQ3 = 0.0;
            if constexpr(TAN) {
                dRc = W[5][q];
                dNc = W[6][q];
                dQ1 = W[7][q];
                dQ2 = W[8][q];
                dQ3 = W[9][q];
            }
            if constexpr(SK != 0) {
                constexpr SplitPlan SPL = split_plan_of(SK, SP);
                static_assert(SPL.dkind == KIND && SPL.dperm == PERM, "split routed to the wrong diagonal kernel");
                double const P[3] = {Q1, Q2, Q3};
                bool const nrs = never_resonant<SK>(Q1, Q2, Q3, res2);
                Nc = nrs ? -Nc : 0.0;
                Rc = 0.0;
                Q1 = P[SPL.q[0]];
                Q2 = P[SPL.q[1]];
                Q3 = P[SPL.q[2]];
                if constexpr(TAN) {
                    double const dP[3] = {dQ1, dQ2, dQ3};
                    dNc = nrs ? -dNc : 0.0;
                    dRc = 0.0;
                    dQ1 = dP[SPL.q[0]];
                    dQ2 = dP[SPL.q[1]];
                    dQ3 = dP[SPL.q[2]];
                }
            }
            double const Pa = F1_IN_D1 ? Q1 : Q3, Pb = F1_IN_D1 ? Q3 : Q1;
            double ar, ai, dar = 0.0, dai = 0.0;
            if constexpr(TAN)
                sep_element_tan<2>(Pa, Pb, F1_IN_D1 ? dQ1 : dQ3, F1_IN_D1 ? dQ3 : dQ1, ya, yb, ya2, yb2, yy, Pa * Pb,
                                   rsc, ar, ai, dar, dai);
            else
                sep_element<2>(Pa, Pb, ya, yb, ya2, yb2, yy, Pa * Pb, rsc, ar, ai);
            double const Er = -(KIND == 2 ? Q2 + Q3 : Q1 + Q2);
            double const dPe = KIND == 2 ? dQ2 + dQ3 : dQ1 + dQ2;
            double br[3], bi[3], dbr[3], dbi[3];
    #pragma unroll
            for(int n = 0; n < 3; ++n) {
                double const e2 = fma(Er, Er, ye[n] * ye[n]);
                if(e2 <= res2) {
                    br[n] = Rc;
                    bi[n] = 0.0;
                    dbr[n] = dRc;
                    dbi[n] = 0.0;
                } else {
                    double const rr = rcp_nr(e2);
                    double const ie = Nc * rr;
                    br[n] = Er * ie;
                    bi[n] = -ye[n] * ie;
                    if constexpr(TAN) { // H' = (dNc + H dPe) / E
                        double const nr = fma(br[n], dPe, dNc), ni = bi[n] * dPe;
                        dbr[n] = fma(nr, Er, ni * ye[n]) * rr;
                        dbi[n] = fma(ni, Er, -nr * ye[n]) * rr;
                    }
                }
            }
            double const a1 = ar + ai;
            if constexpr(TAN) { // Row' H + Row H'
                double const da1 = dar + dai;
    #pragma unroll
                for(int n = 0; n < 3; ++n) {
                    double const b2 = bi[n] - br[n], b3 = br[n] + bi[n];
                    double const db2 = dbi[n] - dbr[n], db3 = dbr[n] + dbi[n];
                    acc[n][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(da1, br[n], acc[n][0], 0, 0, 0);
                    acc[n][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(dai, b3, acc[n][1], 0, 0, 0);
                    acc[n][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(dar, b2, acc[n][2], 0, 0, 0);
                    acc[n][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, dbr[n], acc[n][0], 0, 0, 0);
                    acc[n][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, db3, acc[n][1], 0, 0, 0);
                    acc[n][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, db2, acc[n][2], 0, 0, 0);
                }
            } else {
    #pragma unroll
                for(int n = 0; n < 3; ++n) {
                    double const b2 = bi[n] - br[n], b3 = br[n] + bi[n];
                    acc[n][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, br[n], acc[n][0], 0, 0, 0);
                    acc[n][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, b3, acc[n][1], 0, 0, 0);
                    acc[n][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, b2, acc[n][2], 0, 0, 0);
                }
            }
        }
    }
    // D[i][j] of a 16 x 16 tile sits in lane j + 16 (i % 4), register i / 4
    double* out = partial + (slice * gridDim.y + blockIdx.y) * (16 * kDQ * 2);
#pragma unroll
    for(int n = 0; n < 3; ++n)
#pragma unroll
        for(int x = 0; x < 4; ++x) {
            int const row = (lane >> 4) + 4 * x, col = 16 * n + (lane & 15);
            out[(row * kDQ + col) * 2] = acc[n][0][x] - acc[n][1][x];
            out[(row * kDQ + col) * 2 + 1] = acc[n][0][x] + acc[n][2][x];
        }
}

// acc[(b NR + r) NC + c] += sum over slices of entry (r - r0, q) of tile (b, r / 16), in slice order
__global__ void g2_diag_reduce_kernel(int nslices, int NB, int NR, int NC, int nrt, int toep,
                                      const double* __restrict__ partial, double* __restrict__ acc) {
    int const f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= NB * NR * NC)
        return;
    int const c = f % NC, r = (f / NC) % NR, b = f / (NC * NR);
    int const i = r % 16, q = toep ? c + 15 - i : c + i;
    int64_t const tile = int64_t(b) * nrt + r / 16;
    double sr = 0.0, si = 0.0;
    for(int sl = 0; sl < nslices; ++sl) {
        const double* p = partial + (((int64_t(sl) * NB * nrt + tile) * 16 + i) * kDQ + q) * 2;
        sr += p[0];
        si += p[1];
    }
    acc[2 * f] += sr;
    acc[2 * f + 1] += si;
}

// 1 in flag[0] if some triplet f does not share z2 and z3 with triplet f mod stride
__global__ void stride_mismatch_kernel(int nfreq, int stride, const double* __restrict__ fz, int* __restrict__ flag) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= nfreq)
        return;
    int const g = f % stride;
    bool same = true;
    for(int o = 2; o < 6; ++o) // z2 (doubles 2, 3) and z3 (doubles 4, 5)
        same = same && fz[6 * f + o] == fz[6 * g + o];
    if(!same)
        flag[0] = 1;
}

// 1 in flag[0] if some aligned group of kFPT triplets does not share z1 and z3
__global__ void group_mismatch_kernel(int nfreq, const double* __restrict__ fz, int* __restrict__ flag) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= nfreq)
        return;
    int const g = f - f % kFPT;
    bool same = true;
    for(int c = 0; c < 2; ++c) { // z1 (doubles 0, 1) and z3 (doubles 4, 5)
        int const o = (c == 0) ? 0 : 4;
        same = same && fz[6 * f + o] == fz[6 * g + o] && fz[6 * f + o + 1] == fz[6 * g + o + 1];
    }
    if(!same)
        flag[0] = 1;
}

__global__ void g2_eval_reduce_kernel(int nfreq, int nslices, const double* __restrict__ partial, double* __restrict__ acc) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= nfreq)
        return;
    double sr = 0, si = 0;
    for(int s = 0; s < nslices; ++s) {
        sr += partial[(int64_t(s) * nfreq + f) * 2];
        si += partial[(int64_t(s) * nfreq + f) * 2 + 1];
    }
    acc[2 * f] += sr;
    acc[2 * f + 1] += si;
}

// ---- tangent evaluation ---------------------------------------------------------------------
// Value f and derivative df of every merged term at every triplet, with the poles' derivatives
// entering through d(1/(z - P)) = dP / (z - P)^2:
//   NR:  f = c g, g = 1/(D1 D2 D3),  df = g (dc + c (dP1/D1 + dP2'/D2 + dP3/D3)),
//        D2 = z2 - P2 (dP2' = dP2) or z1 + z2 + z3 - P1 - P2 - P3 (dP2' = dP1 + dP2 + dP3);
//   R:   D = z2 + z3 - P2 - P3 (or z1 + z2 - P1 - P2), q = 1/(D1 D3); if |D| <= res
//        f = Rc q, df = q (dRc + Rc (dP1/D1 + dP3/D3)), else
//        f = Nc q/D, df = (q/D) (dNc + Nc (dP1/D1 + dP3/D3 + dPD/D)).
// One thread per kJF triplets, terms staged through LDS, per-slice partials (value, derivative)
// reduced in a fixed order.
constexpr int kJF = 4;

struct Cx {
    double r, i;
};
__device__ inline Cx cmul(Cx a, Cx b) { return Cx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ inline Cx crcp(double r, double i) {
    double const s = 1.0 / (r * r + i * i);
    return Cx{r * s, -i * s};
}

template <int KIND>
__global__ void __launch_bounds__(256) g2_tangent_eval_kernel(int64_t nterms, int64_t slice_len,
                                                              const double* __restrict__ T, int64_t cap, int nfreq,
                                                              const double* __restrict__ fz, int perm, double res,
                                                              double* __restrict__ partial) {
    constexpr int NF = KIND < 2 ? 8 : 10;
    __shared__ double sT[NF][kChunk];
    int const fbase = (blockIdx.x * blockDim.x + threadIdx.x) * kJF;
    int64_t const t0 = int64_t(blockIdx.y) * slice_len;
    int64_t const t1 = min(nterms, t0 + slice_len);
    double const res2 = res * res;
    Freq q[kJF];
    double acc[kJF][4];
#pragma unroll
    for(int u = 0; u < kJF; ++u) {
        q[u] = load_freq(fz, min(fbase + u, nfreq - 1), perm);
#pragma unroll
        for(int h = 0; h < 4; ++h)
            acc[u][h] = 0.0;
    }
    for(int64_t c0 = t0; c0 < t1; c0 += kChunk) {
        int const cnt = int(min<int64_t>(kChunk, t1 - c0));
        __syncthreads();
        if(int(threadIdx.x) < cnt)
#pragma unroll
            for(int f = 0; f < NF; ++f)
                sT[f][threadIdx.x] = T[f * cap + c0 + threadIdx.x];
        __syncthreads();
        for(int k = 0; k < cnt; ++k) {
            if(KIND < 2) {
                double const C = sT[0][k], P1 = sT[1][k], P2 = sT[2][k], P3 = sT[3][k];
                double const dC = sT[4][k], dP1 = sT[5][k], dP2 = sT[6][k], dP3 = sT[7][k];
                double const dQ2 = (KIND == 0) ? dP2 : (dP1 + dP2) + dP3;
#pragma unroll
                for(int u = 0; u < kJF; ++u) {
                    Freq const& z = q[u];
                    Cx const i1 = crcp(z.z1r - P1, z.z1i);
                    Cx const i3 = crcp(z.z3r - P3, z.z3i);
                    Cx const i2 = (KIND == 0) ? crcp(z.z2r - P2, z.z2i)
                                              : crcp(z.z1r + z.z2r + z.z3r - P1 - P2 - P3, z.z1i + z.z2i + z.z3i);
                    Cx const gg = cmul(cmul(i1, i2), i3);
                    // s = dc + c (dP1 i1 + dQ2 i2 + dP3 i3)
                    double const sr = fma(C, fma(dP1, i1.r, fma(dQ2, i2.r, dP3 * i3.r)), dC);
                    double const si = C * fma(dP1, i1.i, fma(dQ2, i2.i, dP3 * i3.i));
                    acc[u][0] = fma(C, gg.r, acc[u][0]);
                    acc[u][1] = fma(C, gg.i, acc[u][1]);
                    acc[u][2] += gg.r * sr - gg.i * si;
                    acc[u][3] += gg.r * si + gg.i * sr;
                }
            } else {
                double const Rc = sT[0][k], Nc = sT[1][k], P1 = sT[2][k], P2 = sT[3][k], P3 = sT[4][k];
                double const dRc = sT[5][k], dNc = sT[6][k], dP1 = sT[7][k], dP2 = sT[8][k], dP3 = sT[9][k];
                double const dPD = (KIND == 3) ? dP1 + dP2 : dP2 + dP3;
#pragma unroll
                for(int u = 0; u < kJF; ++u) {
                    Freq const& z = q[u];
                    Cx const i1 = crcp(z.z1r - P1, z.z1i);
                    Cx const i3 = crcp(z.z3r - P3, z.z3i);
                    double const Er = (KIND == 3) ? z.z1r + z.z2r - P1 - P2 : z.z2r + z.z3r - P2 - P3;
                    double const Ei = (KIND == 3) ? z.z1i + z.z2i : z.z2i + z.z3i;
                    Cx gg = cmul(i1, i3);
                    double K, dK, xr = fma(dP1, i1.r, dP3 * i3.r), xi = fma(dP1, i1.i, dP3 * i3.i);
                    if(fma(Er, Er, Ei * Ei) <= res2) {
                        K = Rc;
                        dK = dRc;
                    } else {
                        Cx const iE = crcp(Er, Ei);
                        gg = cmul(gg, iE);
                        K = Nc;
                        dK = dNc;
                        xr = fma(dPD, iE.r, xr);
                        xi = fma(dPD, iE.i, xi);
                    }
                    double const sr = fma(K, xr, dK), si = K * xi;
                    acc[u][0] = fma(K, gg.r, acc[u][0]);
                    acc[u][1] = fma(K, gg.i, acc[u][1]);
                    acc[u][2] += gg.r * sr - gg.i * si;
                    acc[u][3] += gg.r * si + gg.i * sr;
                }
            }
        }
    }
#pragma unroll
    for(int u = 0; u < kJF; ++u) {
        int f = fbase + u;
        if(f < nfreq) {
            int64_t o = (int64_t(blockIdx.y) * nfreq + f) * 4;
#pragma unroll
            for(int h = 0; h < 4; ++h)
                partial[o + h] += acc[u][h];
        }
    }
}

__global__ void g2_tangent_reduce_kernel(int nfreq, int nslices, const double* __restrict__ partial,
                                         double* __restrict__ acc, double* __restrict__ dacc) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= nfreq)
        return;
    double s[4] = {0, 0, 0, 0};
    for(int sl = 0; sl < nslices; ++sl)
        for(int h = 0; h < 4; ++h)
            s[h] += partial[(int64_t(sl) * nfreq + f) * 4 + h];
    acc[2 * f] += s[0];
    acc[2 * f + 1] += s[1];
    dacc[2 * f] += s[2];
    dacc[2 * f + 1] += s[3];
}

// 1 if some frequency has a non-zero real part
__global__ void any_real_kernel(int n, const double* __restrict__ fz, int* __restrict__ flag) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k < n && fz[2 * k] != 0.0)
        flag[0] = 1;
}

struct TermSpan {
    const double* base; // SoA base (NF arrays of stride cap)
    int64_t cap, begin, count;
};

// The terms of a split resonant list (kind KIND) that may resonate: flag[t] = 1 (never_resonant
// false), gathered in list order into out (capacity ocap) by gather_terms_kernel.
template <int KIND>
__global__ void maybe_resonant_kernel(int64_t n, const double* __restrict__ T, int64_t cap, double res2,
                                      int64_t* __restrict__ flag) {
    int64_t const t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(t < n)
        flag[t] = never_resonant<KIND>(T[2 * cap + t], T[3 * cap + t], T[4 * cap + t], res2) ? 0 : 1;
}
__global__ void gather_terms_kernel(int64_t n, int nf, const double* __restrict__ T, int64_t cap,
                                    const int64_t* __restrict__ flag, const int64_t* __restrict__ pos,
                                    double* __restrict__ out, int64_t ocap) {
    int64_t const t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(t >= n || !flag[t])
        return;
    int64_t const o = pos[t];
    for(int f = 0; f < nf; ++f)
        out[f * ocap + o] = T[f * cap + t];
}

template <bool IMAG>
void launch_kind(pomx_ctx* ctx, int k, dim3 grid, TermSpan const& s, int64_t slice, int nfreq, const double* d_freq,
                 int perm, double res, double* partial) {
    const double* T = s.base + s.begin;
    switch(k) {
    case 0: g2_eval_kernel<0, IMAG><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, nfreq, d_freq, perm, res, partial); break;
    case 1: g2_eval_kernel<1, IMAG><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, nfreq, d_freq, perm, res, partial); break;
    case 2: g2_eval_kernel<2, IMAG><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, nfreq, d_freq, perm, res, partial); break;
    default: g2_eval_kernel<3, IMAG><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, nfreq, d_freq, perm, res, partial); break;
    }
}

template <bool STRIDED>
void launch_kind_s(pomx_ctx* ctx, int k, int slot, dim3 grid, TermSpan const& s, int64_t slice, int nfreq, int stride,
                   const double* d_freq, int perm, double res, double* partial) {
    const double* T = s.base + s.begin;
#define POMX_G2S(K, S)                                                                                                    \
    g2_eval_s_kernel<K, S, STRIDED><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, nfreq, stride, d_freq, perm, \
                                                                     res, partial)
    switch(k * 3 + slot) {
    case 0: POMX_G2S(0, 0); break;
    case 1: POMX_G2S(0, 1); break;
    case 2: POMX_G2S(0, 2); break;
    case 4: POMX_G2S(1, 1); break;
    case 6: POMX_G2S(2, 0); break;
    case 7: POMX_G2S(2, 1); break;
    case 10: POMX_G2S(3, 1); break;
    case 11: POMX_G2S(3, 2); break;
    default: break; // not a structured (kind, slot): the caller launches the general kernel
    }
#undef POMX_G2S
}

// Separable grid structure of a triplet list (see g2_eval_mfma_kernel), found on the host.
struct SepGrid {
    bool ok = false, sum_ok = false, toep_ok = false, hank_ok = false;
    int NB = 0, NR = 0, NC = 0;
    std::vector<double> yF1, yF2, yF3, yB; // Im F1(b, r), Im F2(c), Im F3(r), Y_b = Im(F1 + F3)
    std::vector<double> yT;                // Im(F2 + F3) at d = c - r, index d + NR - 1
    std::vector<double> yH;                // Im(F1 + F2) at s = b + r + c
};

SepGrid detect_sep_grid(std::vector<double> const& h, int nfreq) {
    SepGrid g;
    auto same = [&](int f, int e, int slot) {
        return h[6 * f + 2 * slot] == h[6 * e + 2 * slot] && h[6 * f + 2 * slot + 1] == h[6 * e + 2 * slot + 1];
    };
    int NC = 0;
    for(int nc = 1; nc <= std::min(nfreq, kMT) && !NC; ++nc) {
        if(nfreq % nc)
            continue;
        bool okc = true;
        for(int f = 0; f < nfreq && okc; ++f)
            okc = same(f, f - f % nc, 0) && same(f, f - f % nc, 2) && same(f, f % nc, 1);
        if(okc)
            NC = nc;
    }
    if(!NC)
        return g;
    int const R = nfreq / NC;
    int NR = 0;
    for(int nr = 1; nr <= std::min(R, kMT) && !NR; ++nr) {
        if(R % nr)
            continue;
        bool okr = true;
        for(int run = 0; run < R && okr; ++run)
            okr = same(run * NC, (run % nr) * NC, 2);
        if(okr)
            NR = nr;
    }
    if(!NR)
        return g;
    g.NB = R / NR;
    g.NR = NR;
    g.NC = NC;
    g.yF1.resize(std::size_t(g.NB) * NR);
    g.yF2.resize(NC);
    g.yF3.resize(NR);
    g.yB.resize(g.NB);
    for(int c = 0; c < NC; ++c)
        g.yF2[c] = h[6 * c + 3];
    for(int r = 0; r < NR; ++r)
        g.yF3[r] = -h[6 * (r * NC) + 5];
    for(int b = 0; b < g.NB; ++b)
        for(int r = 0; r < NR; ++r)
            g.yF1[std::size_t(b) * NR + r] = h[6 * ((b * NR + r) * NC) + 1];
    g.ok = true;
    // Im(F1 + F3) must not depend on r for the z1 + z2 + z3 factor to be a column factor
    g.sum_ok = true;
    for(int b = 0; b < g.NB; ++b) {
        g.yB[b] = g.yF1[std::size_t(b) * NR] + g.yF3[0];
        for(int r = 0; r < NR && g.sum_ok; ++r) {
            double const v = g.yF1[std::size_t(b) * NR + r] + g.yF3[r];
            double const scale = std::fabs(g.yF1[std::size_t(b) * NR + r]) + std::fabs(g.yF3[r]);
            g.sum_ok = std::fabs(v - g.yB[b]) <= 8.0 * 2.220446049250313e-16 * scale;
        }
    }
    // diagonal tables (g2_eval_diag_kernel), each entry taken from one (b, r, c) of its diagonal
    // and checked against all others to within a few roundings of the grid's largest frequency
    // (z1 = W + w3 is itself rounded at the magnitude of W and w3, so the sums along a diagonal
    // agree to that scale, not to their own)
    auto amax = [](std::vector<double> const& v) {
        double m = 0.0;
        for(double x : v)
            m = std::max(m, std::fabs(x));
        return m;
    };
    double const tolT = 4.0 * 2.220446049250313e-16 * (amax(g.yF2) + amax(g.yF3));
    double const tolH = 4.0 * 2.220446049250313e-16 * (amax(g.yF1) + amax(g.yF2));
    g.yT.resize(std::size_t(NR + NC - 1));
    for(int e = 0; e < NR + NC - 1; ++e) {
        int const d = e - (NR - 1), c = d >= 0 ? d : 0, r = d >= 0 ? 0 : -d;
        g.yT[e] = g.yF2[c] + g.yF3[r];
    }
    g.toep_ok = true;
    for(int r = 0; r < NR && g.toep_ok; ++r)
        for(int c = 0; c < NC && g.toep_ok; ++c)
            g.toep_ok = std::fabs(g.yF2[c] + g.yF3[r] - g.yT[c - r + NR - 1]) <= tolT;
    g.yH.resize(std::size_t(g.NB + NR + NC - 2));
    for(int e = 0; e < g.NB + NR + NC - 2; ++e) {
        int const b = std::min(e, g.NB - 1), r = std::min(e - b, NR - 1), c = e - b - r;
        g.yH[e] = g.yF1[std::size_t(b) * NR + r] + g.yF2[c];
    }
    g.hank_ok = true;
    for(int b = 0; b < g.NB && g.hank_ok; ++b)
        for(int r = 0; r < NR && g.hank_ok; ++r)
            for(int c = 0; c < NC && g.hank_ok; ++c)
                g.hank_ok = std::fabs(g.yF1[std::size_t(b) * NR + r] + g.yF2[c] - g.yH[b + r + c]) <= tolH;
    return g;
}

// One non-resonant kind on a separable grid: matrix-core partial tiles, then their sum into acc.
template <int KIND, int PERM, int SRC, bool TAN>
void launch_sep_k(pomx_ctx* ctx, dim3 grid, TermSpan const& s, int64_t slice, const double* T, SepGrid const& g,
                  const double* yF1, const double* yF2, const double* yF3, const double* yB, double res,
                  double* partial) {
    g2_eval_mfma_kernel<KIND, PERM, SRC, TAN><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, g.NR, g.NC, yF1, yF2, yF3,
                                                                   yB, res * res, partial);
}

// kinds and permutations the matrix-core path evaluates on a separable grid: every non-resonant
// one (kind 1 needs Im(F1 + F3) independent of r), and the resonant ones whose E is F1 + F3
bool sep_applies(int kind, int perm, bool sum_ok) {
    static constexpr int p3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    if(kind == 0)
        return true;
    if(!sum_ok)
        return false;
    if(kind == 1)
        return true;
    return kind == 2 ? p3[perm][0] == 1 : p3[perm][2] == 1;
}

// The separable matrix-core evaluation of span s (kind, perm) into d_acc; src >= 0: a resonant list
// of kind src read as kind-0 terms (the separable half of the split).  TAN: a tangent list, whose
// derivative goes to d_acc.
template <bool TAN>
int launch_sep(pomx_ctx* ctx, TermSpan const& s, int kind, int perm, SepGrid const& g, const double* d_y, double res,
               double* d_acc, int src = -1) {
    int64_t const want = std::max<int64_t>(4, 4096 / g.NB);
    int64_t slice = std::max<int64_t>(4, (s.count + want - 1) / want);
    slice = (slice + 3) / 4 * 4;
    int64_t nsl = (s.count + slice - 1) / slice;
    nsl = (nsl + 3) / 4 * 4;
    WsBuf<double> partial(ctx, size_t(nsl) * g.NB * kMT * kMT * 2);
    if(!partial.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    const double* yF1 = d_y;
    const double* yF2 = yF1 + std::size_t(g.NB) * g.NR;
    const double* yF3 = yF2 + g.NC;
    const double* yB = yF3 + g.NR;
    const double* T = s.base + s.begin;
    dim3 const grid(unsigned(nsl / 4), unsigned(g.NB));
#define POMX_SEP_CASE(K, P)                                                                        \
    case K * 6 + P: launch_sep_k<K, P, K, TAN>(ctx, grid, s, slice, T, g, yF1, yF2, yF3, yB, res, partial); break
#define POMX_SEP_SPLIT(S, P)                                                                       \
    case 100 + S * 6 + P: launch_sep_k<0, P, S, TAN>(ctx, grid, s, slice, T, g, yF1, yF2, yF3, yB, res, partial); break
    switch(src >= 0 ? 100 + src * 6 + perm : kind * 6 + perm) {
        POMX_SEP_SPLIT(2, 1); POMX_SEP_SPLIT(2, 4); POMX_SEP_SPLIT(3, 2); POMX_SEP_SPLIT(3, 3);
        POMX_SEP_CASE(0, 0); POMX_SEP_CASE(0, 1); POMX_SEP_CASE(0, 2); POMX_SEP_CASE(0, 3); POMX_SEP_CASE(0, 4);
        POMX_SEP_CASE(0, 5); POMX_SEP_CASE(1, 0); POMX_SEP_CASE(1, 1); POMX_SEP_CASE(1, 2); POMX_SEP_CASE(1, 3);
        POMX_SEP_CASE(1, 4); POMX_SEP_CASE(1, 5); POMX_SEP_CASE(2, 2); POMX_SEP_CASE(2, 3); POMX_SEP_CASE(3, 1);
        POMX_SEP_CASE(3, 4);
    default: return pomx_fail(ctx, POMX_ERR_ARG, "pomx_g2_part_compute: no separable plan for this kind");
    }
#undef POMX_SEP_CASE
#undef POMX_SEP_SPLIT
    POMX_CHECK_LAUNCH(ctx);
    int const nout = g.NB * g.NR * g.NC;
    g2_mfma_reduce_kernel<<<grid_for(nout), kBlock, 0, ctx->stream>>>(int(nsl), g.NB, g.NR, g.NC, partial, d_acc);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

// resonant kinds and permutations the diagonal matrix-core path takes (g2_eval_diag_kernel)
bool diag_applies(int kind, int perm, SepGrid const& g) {
    if(!g.ok || (kind != 2 && kind != 3) || (perm != 0 && perm != 5))
        return false;
    bool const toep = (kind == 2) == (perm == 0);
    return toep ? g.toep_ok : g.hank_ok;
}

// resonant kinds and permutations whose three-factor terms split into a separable and a diagonal
// part (split_plan_of) on this grid
bool split_applies(int kind, int perm, SepGrid const& g) {
    if(!g.ok || !split_kind_perm(kind, perm))
        return false;
    SplitPlan const pl = split_plan_of(kind, perm);
    bool const toep = (pl.dkind == 2) == (pl.dperm == 0);
    return toep ? g.toep_ok : g.hank_ok;
}

// The terms of span s (resonant, kind 2 or 3; nf fields per term) that may resonate, copied into
// store: rem (count 0 when there are none).
int split_remainder(pomx_ctx* ctx, TermSpan const& s, int kind, double res, int nf, WsBuf<double>& store,
                    TermSpan& rem) {
    int64_t const n = s.count;
    rem = TermSpan{nullptr, 0, 0, 0};
    WsBuf<int64_t> flag(ctx, n), pos(ctx, n);
    if(!flag.p || !pos.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    const double* T = s.base + s.begin;
    if(kind == 2)
        maybe_resonant_kernel<2><<<grid_for(n), kBlock, 0, ctx->stream>>>(n, T, s.cap, res * res, flag);
    else
        maybe_resonant_kernel<3><<<grid_for(n), kBlock, 0, ctx->stream>>>(n, T, s.cap, res * res, flag);
    POMX_CHECK_LAUNCH(ctx);
    int rc = pomx_exclusive_scan(ctx, n, flag, pos);
    if(rc)
        return rc;
    int64_t last_pos = 0, last_flag = 0;
    if((rc = read_i64(ctx, pos.p + n - 1, &last_pos)) || (rc = read_i64(ctx, flag.p + n - 1, &last_flag)))
        return rc;
    int64_t const m = last_pos + last_flag;
    if(m == 0)
        return POMX_OK;
    store = WsBuf<double>(ctx, size_t(nf) * m);
    if(!store.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    gather_terms_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, nf, T, s.cap, flag, pos, store, m);
    POMX_CHECK_LAUNCH(ctx);
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream)); // flag and pos outlive the gather
    rem = TermSpan{store.p, m, 0, m};
    return POMX_OK;
}

// The diagonal matrix-core evaluation of span s (kind, perm) into d_acc; src != 0: terms of kind src
// under src_perm, the diagonal half of their split ((kind, perm) being split_plan_of(src,
// src_perm)'s diagonal kernel).  TAN: a tangent list, whose derivative goes to d_acc.
template <bool TAN>
int launch_diag(pomx_ctx* ctx, TermSpan const& s, int kind, int perm, SepGrid const& g, const double* d_y,
                double res, double* d_acc, int src = 0, int src_perm = 0) {
    bool const toep = (kind == 2) == (perm == 0);
    int const nrt = (g.NR + 15) / 16;
    int const ny = g.NB * nrt;
    int64_t const want = std::max<int64_t>(4, 4096 / ny);
    int64_t slice = std::max<int64_t>(4, (s.count + want - 1) / want);
    slice = (slice + 3) / 4 * 4;
    int64_t nsl = (s.count + slice - 1) / slice;
    nsl = (nsl + 3) / 4 * 4;
    WsBuf<double> partial(ctx, size_t(nsl) * ny * 16 * kDQ * 2);
    if(!partial.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    // d_y holds yF1, yF2, yF3, yB, yT, yH
    const double* yF1 = d_y;
    const double* yF3 = yF1 + std::size_t(g.NB) * g.NR + g.NC;
    const double* yT = yF3 + g.NR + g.NB;
    const double* yH = yT + g.yT.size();
    const double* yE = toep ? yT : yH;
    int const nE = int(toep ? g.yT.size() : g.yH.size());
    const double* T = s.base + s.begin;
    dim3 const grid(unsigned(nsl / 4), unsigned(ny));
    double const res2 = res * res;
#define POMX_DIAG(K, P, S, SP)                                                                                         \
    g2_eval_diag_kernel<K, P, S, SP, TAN><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, g.NR, nrt, yF1, yF3, \
                                                                         yE, nE, res2, partial)
    switch(src ? 1000 + src * 6 + src_perm : kind * 6 + perm) {
    case 1000 + 2 * 6 + 1: POMX_DIAG(2, 0, 2, 1); break;
    case 1000 + 2 * 6 + 4: POMX_DIAG(3, 0, 2, 4); break;
    case 1000 + 3 * 6 + 2: POMX_DIAG(3, 0, 3, 2); break;
    case 1000 + 3 * 6 + 3: POMX_DIAG(2, 0, 3, 3); break;
    case 2 * 6 + 0: POMX_DIAG(2, 0, 0, 0); break;
    case 2 * 6 + 5: POMX_DIAG(2, 5, 0, 0); break;
    case 3 * 6 + 0: POMX_DIAG(3, 0, 0, 0); break;
    case 3 * 6 + 5: POMX_DIAG(3, 5, 0, 0); break;
    default: return pomx_fail(ctx, POMX_ERR_ARG, "pomx_g2_part_compute: no diagonal plan for this kind");
    }
#undef POMX_DIAG
    POMX_CHECK_LAUNCH(ctx);
    int const nout = g.NB * g.NR * g.NC;
    g2_diag_reduce_kernel<<<grid_for(nout), kBlock, 0, ctx->stream>>>(int(nsl), g.NB, g.NR, g.NC, nrt, toep ? 1 : 0,
                                                                      partial, d_acc);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

// ---- work model of the evaluation kernels (FLOP: FMA 2, add / mul 1, v_rcp_f64 1; rcp_nr 5) ----
// Separable grids (g2_eval_mfma_kernel), per term and bosonic frequency: the complex GEMM over the
// NR x NC tile from three real products (6 NR NC, matrix cores), one row element per row (two
// factors 17, one 10, plus the a1 sum), one column element per column (one factor 11, two 18,
// resonant 16 with K = Rc or Nc / E folded in, plus the b2, b3 sums), and per term the pole sums
// and products (2) and, resonant, K's reciprocal (11).
double sep_flop_per_term_b(int kind, int perm, int NR, int NC) {
    SepPlan const pl = sep_plan_of(kind, perm);
    double const r_el = pl.nrow == 2 ? 17.0 : 10.0;
    double const c_el = kind >= 2 ? 16.0 : (pl.ncol == 2 ? 18.0 : 11.0);
    double const t_el = 2.0 + (kind >= 2 ? 11.0 : 0.0);
    return 6.0 * NR * NC + NR * (r_el + 1.0) + NC * (c_el + 2.0) + t_el;
}
// Diagonal resonant terms (g2_eval_diag_kernel), per term, bosonic frequency and 16-row tile: the
// 16 x (16 + NC - 1) band product (6 per entry), 16 row elements (17 + 1), 16 + NC - 1 diagonal
// values (|E|^2, Nc / E: 11, plus 2 sums) and 3 per term.
double diag_flop_per_term_tile(int NC) {
    double const ne = 16.0 + NC - 1.0;
    return 6.0 * 16.0 * ne + 16.0 * 18.0 + ne * 13.0 + 3.0;
}
// Grouped grids (g2_eval_s_kernel), per term and frequency: 54 per four frequencies (|V|^2, a
// quarter of one reciprocal, K conj(V) / |V|^2 accumulated), plus K = Kc / G once per term and
// thread of kFPT frequencies (20 kind 0, 22 kinds 1 and 2/3 with E varying, 27 with E fixed).
double grouped_flop_per_term_freq(int kind, int slot) {
    bool const evar = (kind == 2 || kind == 3) && slot == 1;
    double const per_thread = kind == 0 ? 20.0 : (kind == 1 || evar) ? 22.0 : 27.0;
    return 54.0 / 4.0 + per_thread / kFPT;
}
// Any triplets (g2_eval_kernel): purely imaginary triplets 83 per four frequencies (one shared
// reciprocal); otherwise the direct formula as written, 28 / 31 (non-resonant z2 / z4) and 32
// (resonant) per term and frequency.
double general_flop_per_term_freq(int kind, bool imag) {
    if(imag)
        return 83.0 / 4.0;
    return kind == 0 ? 28.0 : kind == 1 ? 31.0 : 32.0;
}

int launch_g2_tangent_eval(pomx_ctx* ctx, TermSpan const spans[4], int perm, double res, int nfreq,
                           const double* d_freq, double* d_acc, double* d_dacc);

// The Lehmann sum of a part's kind lists at the triplets d_freq into d_acc.  With d_dacc the lists
// are tangent lists (value fields, then the derivatives) and the derivative goes to d_dacc: the
// matrix-core classes run once for the value and once in their tangent form, and the kinds left
// to the vector kernels take the tangent vector kernel (value and derivative together).
int launch_g2_eval(pomx_ctx* ctx, TermSpan const spans[4], int perm, double res, int nfreq, const double* d_freq,
                   double* d_acc, double* d_dacc) {
    bool const tan = d_dacc != nullptr;
    int64_t total = 0;
    for(int k = 0; k < 4; ++k)
        total += spans[k].count;
    if(total == 0 || nfreq == 0)
        return POMX_OK;
    int const stride = nfreq / kFPT;
    bool const divisible = nfreq % kFPT == 0;
    WsBuf<int> flag(ctx, 3);
    int rc = pomx_fill_zero(ctx, flag.p, 3 * sizeof(int));
    if(rc)
        return rc;
    any_real_kernel<<<grid_for(3 * nfreq), kBlock, 0, ctx->stream>>>(3 * nfreq, d_freq, flag);
    if(divisible) {
        group_mismatch_kernel<<<grid_for(nfreq), kBlock, 0, ctx->stream>>>(nfreq, d_freq, flag.p + 1);
        stride_mismatch_kernel<<<grid_for(nfreq), kBlock, 0, ctx->stream>>>(nfreq, stride, d_freq, flag.p + 2);
    }
    int hflag[3] = {0, 0, 0};
    POMX_HIP(ctx, hipMemcpyAsync(hflag, flag.p, 3 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int const has_real = hflag[0];
    // separable grids take the matrix cores for the non-resonant kinds
    SepGrid sep;
    WsBuf<double> d_y;
    if(!has_real) {
        std::vector<double> h(std::size_t(6) * nfreq);
        POMX_HIP(ctx, hipMemcpyAsync(h.data(), d_freq, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        sep = detect_sep_grid(h, nfreq);
        if(sep.ok) {
            std::vector<double> y(sep.yF1);
            y.insert(y.end(), sep.yF2.begin(), sep.yF2.end());
            y.insert(y.end(), sep.yF3.begin(), sep.yF3.end());
            y.insert(y.end(), sep.yB.begin(), sep.yB.end());
            y.insert(y.end(), sep.yT.begin(), sep.yT.end());
            y.insert(y.end(), sep.yH.begin(), sep.yH.end());
            d_y = WsBuf<double>(ctx, y.size());
            if(!d_y.p)
                return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
            POMX_HIP(ctx, h2d_sync(ctx, d_y.p, y.data(), sizeof(double) * y.size()));
        }
    }
    // purely imaginary triplets whose aligned groups share z1 and z3 (F2 = z2 varies inside a
    // group), or whose stride-apart triplets share z2 and z3 (F1 = z1 varies): the slot of the
    // varying F under this part's permutation decides which kinds take the grouped kernels
    bool const grouped = !has_real && divisible && !hflag[1];
    bool const strided = !has_real && divisible && !hflag[2];
    static constexpr int perm3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int const slot2 = (perm3[perm][0] == 1) ? 0 : (perm3[perm][1] == 1) ? 1 : 2; // slot of F2
    int const slot1 = (perm3[perm][0] == 0) ? 0 : (perm3[perm][1] == 0) ? 1 : 2; // slot of F1
    int ftiles = (nfreq + 256 * kFPT - 1) / (256 * kFPT);
    int64_t want = std::max<int64_t>(1, 4096 / ftiles);
    int64_t slice = std::max<int64_t>(kChunk, (total + want - 1) / want);
    slice = (slice + kChunk - 1) / kChunk * kChunk;
    int nslices = 1;
    for(int k = 0; k < 4; ++k)
        nslices = std::max<int64_t>(nslices, (spans[k].count + slice - 1) / slice);
    bool any_vector = false; // some kind takes the vector kernels, which need zeroed partial sums
    for(int k = 0; k < 4; ++k)
        if(spans[k].count && !(sep.ok && sep_applies(k, perm, sep.sum_ok)) && !diag_applies(k, perm, sep))
            any_vector = !tan; // (split kinds: the terms that may resonate)
    WsBuf<double> partial;
    if(any_vector) {
        partial = WsBuf<double>(ctx, size_t(nslices) * nfreq * 2);
        if(!partial.p)
            return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
        if((rc = pomx_fill_zero(ctx, partial.p, sizeof(double) * size_t(nslices) * nfreq * 2)))
            return rc;
    }
    // tangent: the spans left to the tangent vector kernel (a kind's whole span, or the terms of a
    // split kind that may resonate, held in vstore)
    TermSpan vspans[4] = {};
    WsBuf<double> vstore[4];
    // the vector kernels (grouped, strided or general) on span v of kind k
    auto vector_eval = [&](TermSpan const& v, int k) {
        if(tan) {
            vspans[k] = v;
            return;
        }
        double const cnt = double(v.count), tbytes = 8.0 * (k < 2 ? 4 : 5);
        // a span smaller than the part (the terms of a split kind that may resonate) is spread over
        // up to nslices slices of at least kChunk terms
        int64_t const vsl = std::min<int64_t>(nslices, (v.count + kChunk - 1) / kChunk);
        int64_t vslice = (v.count + vsl - 1) / vsl;
        vslice = (vslice + kChunk - 1) / kChunk * kChunk;
        dim3 grid(ftiles, unsigned((v.count + vslice - 1) / vslice));
        bool const g = grouped && g2_structured(k, slot2), st = !g && strided && g2_structured(k, slot1);
        double const fpt = (g || st) ? grouped_flop_per_term_freq(k, g ? slot2 : slot1)
                                     : general_flop_per_term_freq(k, !has_real);
        int const h = stat_begin(ctx, (g || st) ? POMX_STAT_G2_GROUPED : POMX_STAT_G2_GENERAL);
        if(g)
            launch_kind_s<false>(ctx, k, slot2, grid, v, vslice, nfreq, stride, d_freq, perm, res, partial);
        else if(st)
            launch_kind_s<true>(ctx, k, slot1, grid, v, vslice, nfreq, stride, d_freq, perm, res, partial);
        else if(has_real)
            launch_kind<false>(ctx, k, grid, v, vslice, nfreq, d_freq, perm, res, partial);
        else
            launch_kind<true>(ctx, k, grid, v, vslice, nfreq, d_freq, perm, res, partial);
        stat_end(ctx, h, cnt * nfreq * fpt, cnt * nfreq, cnt * tbytes);
    };
    for(int k = 0; k < 4; ++k) {
        TermSpan const& s = spans[k];
        if(!s.count)
            continue;
        double const cnt = double(s.count), tbytes = 8.0 * (k < 2 ? 4 : 5); // a term's fields, read once
        if(sep.ok && sep_applies(k, perm, sep.sum_ok)) {
            double const nout = double(sep.NB) * sep.NR * sep.NC;
            int const h = stat_begin(ctx, POMX_STAT_G2_SEP);
            if((rc = launch_sep<false>(ctx, s, k, perm, sep, d_y, res, d_acc)) ||
               (tan && (rc = launch_sep<true>(ctx, s, k, perm, sep, d_y, res, d_dacc))))
                return rc;
            stat_end(ctx, h, cnt * sep.NB * sep_flop_per_term_b(k, perm, sep.NR, sep.NC), cnt * nout,
                     cnt * tbytes + 32.0 * nout);
        } else if(diag_applies(k, perm, sep)) {
            double const nout = double(sep.NB) * sep.NR * sep.NC;
            int const h = stat_begin(ctx, POMX_STAT_G2_DIAG);
            if((rc = launch_diag<false>(ctx, s, k, perm, sep, d_y, res, d_acc)) ||
               (tan && (rc = launch_diag<true>(ctx, s, k, perm, sep, d_y, res, d_dacc))))
                return rc;
            stat_end(ctx, h, cnt * sep.NB * ((sep.NR + 15) / 16) * diag_flop_per_term_tile(sep.NC), cnt * nout,
                     cnt * tbytes + 32.0 * nout);
        } else if(split_applies(k, perm, sep)) {
            // the separable and the diagonal half of the never-resonant terms, then the vector
            // kernels on the terms that may resonate
            double const nout = double(sep.NB) * sep.NR * sep.NC;
            int h = stat_begin(ctx, POMX_STAT_G2_SEP);
            if((rc = launch_sep<false>(ctx, s, 0, perm, sep, d_y, res, d_acc, k)) ||
               (tan && (rc = launch_sep<true>(ctx, s, 0, perm, sep, d_y, res, d_dacc, k))))
                return rc;
            stat_end(ctx, h, cnt * sep.NB * sep_flop_per_term_b(0, perm, sep.NR, sep.NC), cnt * nout,
                     cnt * tbytes + 32.0 * nout);
            SplitPlan const pl = split_plan_of(k, perm);
            h = stat_begin(ctx, POMX_STAT_G2_DIAG);
            if((rc = launch_diag<false>(ctx, s, pl.dkind, pl.dperm, sep, d_y, res, d_acc, k, perm)) ||
               (tan && (rc = launch_diag<true>(ctx, s, pl.dkind, pl.dperm, sep, d_y, res, d_dacc, k, perm))))
                return rc;
            stat_end(ctx, h, cnt * sep.NB * ((sep.NR + 15) / 16) * diag_flop_per_term_tile(sep.NC), cnt * nout,
                     cnt * tbytes + 32.0 * nout);
            TermSpan rem;
            if((rc = split_remainder(ctx, s, k, res, tan ? 10 : 5, vstore[k], rem)))
                return rc;
            if(rem.count) {
                vector_eval(rem, k);
                if(!tan)
                    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream)); // the store outlives the kernels
            }
        } else
            vector_eval(s, k);
    }
    if(tan) { // value and derivative of the rest by the tangent vector kernel
        if((rc = launch_g2_tangent_eval(ctx, vspans, perm, res, nfreq, d_freq, d_acc, d_dacc)))
            return rc;
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream)); // vstore outlives the kernels
    }
    if(any_vector) { // the vector kernels' per-slice partial sums into the grid
        int const h = stat_begin(ctx, POMX_STAT_G2_REDUCE);
        g2_eval_reduce_kernel<<<grid_for(nfreq), kBlock, 0, ctx->stream>>>(nfreq, nslices, partial, d_acc);
        POMX_CHECK_LAUNCH(ctx);
        stat_end(ctx, h, 2.0 * nslices * nfreq, 0.0, 16.0 * nslices * nfreq + 32.0 * nfreq);
    }
    return POMX_OK;
}

int launch_g2_tangent_eval(pomx_ctx* ctx, TermSpan const spans[4], int perm, double res, int nfreq,
                           const double* d_freq, double* d_acc, double* d_dacc) {
    int64_t total = 0;
    for(int k = 0; k < 4; ++k)
        total += spans[k].count;
    if(total == 0 || nfreq == 0)
        return POMX_OK;
    int ftiles = (nfreq + 256 * kJF - 1) / (256 * kJF);
    int64_t want = std::max<int64_t>(1, 4096 / ftiles);
    int64_t slice = std::max<int64_t>(kChunk, (total + want - 1) / want);
    slice = (slice + kChunk - 1) / kChunk * kChunk;
    int nslices = 1;
    for(int k = 0; k < 4; ++k)
        nslices = std::max<int64_t>(nslices, (spans[k].count + slice - 1) / slice);
    WsBuf<double> partial(ctx, size_t(nslices) * nfreq * 4);
    int rc = pomx_fill_zero(ctx, partial.p, sizeof(double) * size_t(nslices) * nfreq * 4);
    if(rc)
        return rc;
    for(int k = 0; k < 4; ++k) {
        TermSpan const& s = spans[k];
        if(!s.count)
            continue;
        dim3 grid(ftiles, unsigned((s.count + slice - 1) / slice));
        const double* T = s.base + s.begin;
        switch(k) {
        case 0: g2_tangent_eval_kernel<0><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, nfreq, d_freq, perm, res, partial); break;
        case 1: g2_tangent_eval_kernel<1><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, nfreq, d_freq, perm, res, partial); break;
        case 2: g2_tangent_eval_kernel<2><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, nfreq, d_freq, perm, res, partial); break;
        default: g2_tangent_eval_kernel<3><<<grid, 256, 0, ctx->stream>>>(s.count, slice, T, s.cap, nfreq, d_freq, perm, res, partial); break;
        }
    }
    g2_tangent_reduce_kernel<<<grid_for(nfreq), kBlock, 0, ctx->stream>>>(nfreq, nslices, partial, d_acc, d_dacc);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

} // namespace

// Shared with the 3-point susceptibility pipeline (declared in pomx_internal.hpp).
int read_i64(pomx_ctx* ctx, const int64_t* d, int64_t* h) {
    POMX_HIP(ctx, hipMemcpyAsync(h, d, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

// dense ranks of round((Ey - Ex)/res) over all (x, y); 'sorted' receives the (x*ny+y) order
int rank_table(pomx_ctx* ctx, int nx, int ny, const double* Ex, const double* Ey, double res, WsBuf<uint32_t>& rank,
               WsBuf<uint64_t>* sorted, int64_t* nranks, int* d_err, WsBuf<uint8_t>* uniq) {
    int64_t n = int64_t(nx) * ny;
    WsBuf<uint64_t> k0(ctx, n), k1(ctx, n), v0(ctx, n), v1(ctx, n);
    WsBuf<int64_t> head(ctx, n), excl(ctx, n);
    bins_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(nx, ny, Ex, Ey, res, k0, v0, d_err);
    POMX_CHECK_LAUNCH(ctx);
    int rc = pomx_sort_pairs(ctx, n, k0, k1, v0, v1, 0, 64);
    if(rc)
        return rc;
    rank_heads_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, k1, head);
    if((rc = pomx_exclusive_scan(ctx, n, head, excl)))
        return rc;
    rank = WsBuf<uint32_t>(ctx, n);
    rank_scatter_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, head, excl, v1, rank);
    POMX_CHECK_LAUNCH(ctx);
    int64_t lh = 0, le = 0;
    if((rc = read_i64(ctx, head.p + n - 1, &lh)) || (rc = read_i64(ctx, excl.p + n - 1, &le)))
        return rc;
    *nranks = lh + le;
    if(sorted) {
        *sorted = WsBuf<uint64_t>(ctx, n);
        POMX_HIP(ctx, hipMemcpyAsync(sorted->p, v1.p, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, ctx->stream));
    }
    if(uniq) {
        *uniq = WsBuf<uint8_t>(ctx, n);
        if(!uniq->p)
            return pomx_fail(ctx, POMX_ERR_HIP, "rank_table: out of device memory");
        unique_bin_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, k1, v1, uniq->p);
        POMX_CHECK_LAUNCH(ctx);
    }
    return POMX_OK;
}

namespace {

// Tangent reduction, compaction and evaluation of one chunk's sorted groups.
// Chunk budget in raw terms.  A chunk's generation holds keys, values, sort buffers and group
// arrays (about 110 bytes per term while every group is a singleton): 2^29 terms keep every part
// of the 4096-state Anderson workload (largest 4.5e8 terms) in one chunk at about 60 GB.
constexpr int64_t kChunkTerms = int64_t(1) << 29;

// The chunk budget in force: a positive integer in the environment variable POMX_G2_CHUNK_TERMS
// replaces kChunkTerms, so that small parts can be generated in several chunks (a test hook).
int64_t chunk_terms() {
    if(const char* e = std::getenv("POMX_G2_CHUNK_TERMS")) {
        long long const v = std::atoll(e);
        if(v > 0)
            return int64_t(v);
    }
    return kChunkTerms;
}

// Direct terms in force: the environment variable POMX_G2_DIRECT=0 sends every term through the sort
// (a test hook, so that the sorted path and its chunking are tested on every part).
bool direct_terms_enabled() {
    if(const char* e = std::getenv("POMX_G2_DIRECT"))
        return std::atoi(e) != 0;
    return true;
}

} // namespace

extern "C" int pomx_g2_terms_eval(pomx_ctx* ctx, const pomx_g2_terms* t, int perm, double res, int nfreq,
                                  const double* d_freq, double* d_acc) {
    POMX_REQUIRE(ctx, t && perm >= 0 && perm < 6 && nfreq >= 0, "pomx_g2_terms_eval: bad arguments");
    int64_t nnr = t->count[0] + t->count[1], nr = t->count[2] + t->count[3];
    const double* NR = t->data;
    const double* R = t->data + 4 * nnr;
    TermSpan spans[4] = {{NR, nnr, 0, t->count[0]}, {NR, nnr, t->count[0], t->count[1]},
                         {R, nr, 0, t->count[2]},   {R, nr, t->count[2], t->count[3]}};
    int rc = launch_g2_eval(ctx, spans, perm, res, nfreq, d_freq, d_acc, nullptr);
    if(rc)
        return rc;
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

namespace {

// One chunk's terms, sorted by key (the members of every group in visiting order), and, once
// reduced, its group results.
struct G2Chunk {
    WsBuf<uint64_t> keys, vals;
    WsBuf<int64_t> start; // first member of every group
    int64_t n = 0, ng = 0;
    WsBuf<double> g;      // 5 doubles per group (g2_reduce_kernel)
    WsBuf<int64_t> live;
    bool reduced = false; // g and live hold the final decisions
};

// Tables shared by a part's chunks.
struct G2Gen {
    const double *O3c, *C4t;
    const int* rowcnt;     // sorted terms per (r12-sorted pair, i3)
    const uint32_t *r12, *r23, *r34;
    const uint8_t *u12, *u34; // unique-bin flags (direct terms), or null (every term sorted)
};

// Generate, sort and group the terms of the pairs [pb, pe) of the r12-sorted pair list.
int g2_generate_chunk(pomx_ctx* ctx, G2Dev const& D, G2Gen const& T, std::vector<uint64_t> const& hpairs,
                      std::vector<int64_t> const& hpair, int64_t pb, int64_t pe, int64_t n, G2Chunk& out) {
    // the chunk's pairs with sorted terms in ascending (i1, i2) order, in i1 segments (dropping the
    // pairs without any keeps the visiting order of the others)
    std::vector<std::pair<uint64_t, uint32_t>> ps;
    ps.reserve(size_t(pe - pb));
    for(int64_t p = pb; p < pe; ++p)
        if(hpair[size_t(p)] > 0)
            ps.emplace_back(hpairs[size_t(p)], uint32_t(p));
    std::sort(ps.begin(), ps.end());
    std::vector<uint4> rmap(ps.size());
    for(size_t a = 0; a < ps.size();) {
        size_t b = a;
        uint64_t const i1 = ps[a].first / uint64_t(D.n2);
        while(b < ps.size() && ps[b].first / uint64_t(D.n2) == i1)
            ++b;
        for(size_t k = a; k < b; ++k)
            rmap[k] = make_uint4(uint32_t(ps[k].first), ps[k].second, uint32_t(a), uint32_t(b - a));
        a = b;
    }
    int64_t const nrow = int64_t(rmap.size()) * D.n3;
    WsBuf<uint4> d_rmap(ctx, rmap.size());
    WsBuf<int64_t> wide(ctx, nrow), off(ctx, nrow);
    if(!d_rmap.p || !wide.p || !off.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    POMX_HIP(ctx, hipMemcpyAsync(d_rmap.p, rmap.data(), sizeof(uint4) * rmap.size(), hipMemcpyHostToDevice, ctx->stream));
    chunk_rowcnt_kernel<<<grid_for(nrow), kBlock, 0, ctx->stream>>>(nrow, D.n3, d_rmap, T.rowcnt, wide);
    POMX_CHECK_LAUNCH(ctx);
    int rc = pomx_exclusive_scan(ctx, nrow, wide, off);
    if(rc)
        return rc;
    WsBuf<uint64_t> k0(ctx, n), v0(ctx, n);
    out.keys = WsBuf<uint64_t>(ctx, n);
    out.vals = WsBuf<uint64_t>(ctx, n);
    if(!k0.p || !v0.p || !out.keys.p || !out.vals.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    unsigned const eblocks = unsigned(std::min<int64_t>((nrow + 3) / 4, 1 << 16));
    g2_emit_kernel<<<eblocks, kBlock, 0, ctx->stream>>>(D, T.O3c, T.C4t, d_rmap, nrow, n, off, T.r12, T.r23, T.r34, T.u12,
                                                         T.u34, k0, v0);
    POMX_CHECK_LAUNCH(ctx);
    if((rc = pomx_sort_pairs(ctx, n, k0, out.keys, v0, out.vals, 0, 64)))
        return rc;
    k0 = WsBuf<uint64_t>();
    v0 = WsBuf<uint64_t>();
    WsBuf<int64_t> head(ctx, n), gid(ctx, n);
    if(!head.p || !gid.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    heads_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, out.keys, head);
    if((rc = pomx_exclusive_scan(ctx, n, head, gid)))
        return rc;
    int64_t lh = 0, lg = 0; // (read_i64 synchronises: rmap's copy has finished before it is freed)
    if((rc = read_i64(ctx, head.p + n - 1, &lh)) || (rc = read_i64(ctx, gid.p + n - 1, &lg)))
        return rc;
    out.n = n;
    out.ng = lh + lg;
    out.start = WsBuf<int64_t>(ctx, out.ng);
    if(!out.start.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    group_start_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, head, gid, out.start);
    POMX_CHECK_LAUNCH(ctx);
    return POMX_OK;
}

int g2_alloc_groups(pomx_ctx* ctx, G2Chunk& c) {
    if(!c.g.p)
        c.g = WsBuf<double>(ctx, size_t(c.ng) * 5);
    if(!c.live.p)
        c.live = WsBuf<int64_t>(ctx, c.ng);
    if(!c.g.p || !c.live.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    return POMX_OK;
}

// A chunk's groups, reduced (with the threshold thr, or the settled rule), appended to the part's
// kind lists at fill[k]; with T the tangent reduce, whose groups carry their derivatives, into
// the tangent kind lists.
int g2_finish_chunk(pomx_ctx* ctx, G2Dev const& D, G2Chunk& c, double thr, EraseRule const& rule, const G2Tan* T,
                    KindLists const& out, int64_t fill[4], int64_t counts[4]) {
    int rc;
    int64_t const ng = c.ng;
    WsBuf<double> gt;
    WsBuf<int64_t> lt;
    const double* g;
    const int64_t* live;
    if(T) {
        gt = WsBuf<double>(ctx, size_t(ng) * 10);
        lt = WsBuf<int64_t>(ctx, ng);
        if(!gt.p || !lt.p)
            return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_tangent: out of device memory");
        g2_reduce_tangent_kernel<<<grid_for(ng), kBlock, 0, ctx->stream>>>(D, *T, c.n, ng, c.start, c.keys, c.vals, thr,
                                                                           rule, gt, lt);
        POMX_CHECK_LAUNCH(ctx);
        g = gt.p;
        live = lt.p;
    } else {
        if(!c.reduced) {
            if((rc = g2_alloc_groups(ctx, c)))
                return rc;
            g2_reduce_kernel<false><<<grid_for(ng), kBlock, 0, ctx->stream>>>(D, c.n, ng, c.start, c.keys, c.vals, thr,
                                                                                rule, c.g, c.live, nullptr, nullptr,
                                                                                nullptr, nullptr);
            POMX_CHECK_LAUNCH(ctx);
        }
        g = c.g.p;
        live = c.live.p;
    }
    WsBuf<int64_t> lpos(ctx, ng), kc(ctx, 4);
    if(!lpos.p || !kc.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    if((rc = pomx_exclusive_scan(ctx, ng, live, lpos)))
        return rc;
    kind_count_kernel<<<1, 64, 0, ctx->stream>>>(ng, live, lpos, c.start, c.keys, kc);
    int64_t k[4];
    POMX_HIP(ctx, hipMemcpyAsync(k, kc.p, sizeof(k), hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    KindSlots slot;
    int64_t first = 0;
    for(int q = 0; q < 4; ++q) {
        slot.first[q] = first; // live groups of kind q start at this live index
        slot.at[q] = fill[q];
        first += k[q];
        if(fill[q] + k[q] > out.cap[q] || (k[q] > 0 && !out.base[q]))
            return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_g2_part_compute: term list capacity exceeded");
    }
    if(T)
        g2_compact_tangent_kernel<<<grid_for(ng), kBlock, 0, ctx->stream>>>(ng, live, lpos, g, c.start, c.keys, out,
                                                                            slot);
    else
        g2_compact_kernel<<<grid_for(ng), kBlock, 0, ctx->stream>>>(ng, live, lpos, g, c.start, c.keys, out, slot);
    POMX_CHECK_LAUNCH(ctx);
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream)); // slot and the chunk's buffers outlive the kernel
    c = G2Chunk();
    for(int q = 0; q < 4; ++q) {
        fill[q] += k[q];
        counts[q] += k[q];
    }
    return POMX_OK;
}

// The reference's erase decisions (TermList.hpp add_term: |C| <= tol/(size+1)) for the held
// chunks, whose groups include every merge of the part.  The threshold of a merge needs the
// number of groups of its list started before it, over the whole part: the merge positions of the
// held chunks are sorted (with a sentinel between the two lists, g2_merge_positions) and every
// group's first position is counted into the gap it falls in (g2_first_hist_kernel over the
// sorted chunks' groups, g2_direct_emit_kernel over the direct terms, into hist); g2_settle then
// settles the erase and restart events by the fixed point of erase_fixed_point over all held
// chunks.  On return every held chunk is reduced and 'rule' is the settled rule.
constexpr uint64_t kListGap = (uint64_t(1) << 63) - 1; // above every first-list, below every second-list position

int g2_merge_positions(pomx_ctx* ctx, G2Dev const& D, std::vector<G2Chunk*> const& held, WsBuf<uint64_t>& qs,
                       int64_t& m, WsBuf<int64_t>& hist) {
    m = 1;
    for(G2Chunk const* c : held)
        m += c->n - c->ng;
    WsBuf<uint64_t> q0(ctx, m), v0(ctx, m), v1(ctx, m);
    qs = WsBuf<uint64_t>(ctx, m);
    hist = WsBuf<int64_t>(ctx, m + 1);
    if(!q0.p || !v0.p || !v1.p || !qs.p || !hist.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    int64_t at = 0;
    for(G2Chunk const* c : held) {
        g2_merge_pos_kernel<<<grid_for(c->ng), kBlock, 0, ctx->stream>>>(D, c->n, c->ng, c->start, c->keys, c->vals,
                                                                          q0.p + at);
        POMX_CHECK_LAUNCH(ctx);
        at += c->n - c->ng;
    }
    POMX_HIP(ctx, hipMemcpyAsync(q0.p + at, &kListGap, sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    int rc = pomx_sort_pairs(ctx, m, q0, qs, v0, v1, 0, 64);
    if(rc)
        return rc;
    if((rc = pomx_fill_zero(ctx, hist.p, sizeof(int64_t) * size_t(m + 1))))
        return rc;
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream)); // kListGap's copy has finished
    return POMX_OK;
}

int g2_settle(pomx_ctx* ctx, G2Dev const& D, double tol, std::vector<G2Chunk*> const& held,
              std::vector<std::pair<const uint64_t*, int64_t>> const& fks, WsBuf<uint64_t> const& qs, int64_t m,
              WsBuf<int64_t> const& hist, EraseRule& rule, EraseEvents& ev, WsBuf<int64_t>& started) {
    int64_t ng = 0;
    for(G2Chunk const* c : held)
        ng += c->ng;
    started = WsBuf<int64_t>(ctx, m + 1);
    WsBuf<int64_t> base(ctx, 1);
    if(!started.p || !base.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    int rc;
    for(auto const& f : fks) {
        if(f.second == 0)
            continue;
        unsigned const blocks = unsigned(std::min<int64_t>((f.second + kBlock - 1) / kBlock, 4096));
        g2_first_hist_kernel<<<blocks, kBlock, 0, ctx->stream>>>(f.second, f.first, qs, m,
                                                                  reinterpret_cast<unsigned long long*>(hist.p));
        POMX_CHECK_LAUNCH(ctx);
    }
    if((rc = pomx_exclusive_scan(ctx, m + 1, hist, started)))
        return rc;
    add_own_kernel<<<grid_for(m + 1), kBlock, 0, ctx->stream>>>(m + 1, hist, started);
    lookup_started_kernel<<<1, 64, 0, ctx->stream>>>(qs, m, started, kListGap, base);
    POMX_CHECK_LAUNCH(ctx);
    int64_t base_hi = 0;
    if((rc = read_i64(ctx, base.p, &base_hi)))
        return rc;
    rule.fk = qs.p;
    rule.nfk = m;
    rule.qstarted = started.p;
    rule.base_hi = base_hi;
    rule.ev_key = nullptr;
    rule.ev_cum = nullptr;
    rule.nev = 0;
    for(G2Chunk* c : held)
        if((rc = g2_alloc_groups(ctx, *c)))
            return rc;
    auto pass = [&](EraseRule const& R, bool emit, int64_t* evn, uint64_t* ko, uint64_t* wo) -> int {
        int64_t goff = 0;
        for(G2Chunk* c : held) {
            if(emit)
                g2_reduce_kernel<true><<<grid_for(c->ng), kBlock, 0, ctx->stream>>>(D, c->n, c->ng, c->start, c->keys,
                                                                                    c->vals, 0.0, R, c->g, c->live,
                                                                                    evn + goff, ko, wo, nullptr);
            else
                g2_reduce_kernel<false><<<grid_for(c->ng), kBlock, 0, ctx->stream>>>(D, c->n, c->ng, c->start, c->keys,
                                                                                     c->vals, 0.0, R, c->g, c->live,
                                                                                     evn + goff, ko, wo, nullptr);
            POMX_CHECK_LAUNCH(ctx);
            goff += c->ng;
        }
        return POMX_OK;
    };
    if((rc = erase_fixed_point(ctx, ng, rule, pass, ev, "pomx_g2_part_compute")))
        return rc;
    rule.ev_key = ev.key.p;
    rule.ev_cum = ev.cum.p;
    rule.nev = ev.n;
    for(G2Chunk* c : held)
        c->reduced = true;
    return POMX_OK;
}

// Shared by the value path (tan == NULL) and the tangent path (keep must be NULL).  The value path
// writes the direct terms (unique bins) straight into the part's kind lists and sends the others
// through the sort; the tangent path sorts every term.
int g2_part_run(pomx_ctx* ctx, const pomx_g2_part* part, const pomx_g2_tangent* tan, double beta, double res, double tol,
                int nfreq, const double* d_freq, double* d_acc, double* d_dacc, pomx_g2_terms* keep, int64_t* h_counts,
                int64_t* h_raw_terms) {
    POMX_REQUIRE(ctx, part && res > 0.0 && nfreq >= 0 && (nfreq == 0 || (d_freq && d_acc)),
                 "pomx_g2_part_compute: bad arguments");
    POMX_REQUIRE(ctx, !tan || (!keep && (nfreq == 0 || d_dacc)), "pomx_g2_part_tangent: bad arguments");
    POMX_REQUIRE(ctx, part->perm >= 0 && part->perm < 6, "pomx_g2_part_compute: bad permutation");
    pomx_g2_part const& P = *part;
    POMX_REQUIRE(ctx, P.n1 >= 1 && P.n2 >= 1 && P.n3 >= 1 && P.n4 >= 1, "pomx_g2_part_compute: empty block");
    if(P.n1 > 65535 || P.n2 > 65535 || P.n3 > 65535 || P.n4 > 65535)
        return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_g2_part_compute: block larger than 65535");
    G2Dev D{P.n1, P.n2, P.n3, P.n4, P.O1, P.O2, P.O3, P.CX4, P.E1, P.E2, P.E3, P.E4,
            P.w1, P.w2, P.w3, P.w4, P.sign, beta, tol};
    bool const direct = direct_terms_enabled();
    G2Tan TT{};
    if(tan)
        TT = G2Tan{tan->dO1, tan->dO2, tan->dO3, tan->dCX4, tan->dE1, tan->dE2, tan->dE3, tan->dE4,
                   tan->dw1, tan->dw2, tan->dw3, tan->dw4};
    const G2Tan* Tp = tan ? &TT : nullptr;
    int64_t counts[4] = {0, 0, 0, 0};
    if(keep) {
        keep->data = nullptr;
        for(int k = 0; k < 4; ++k)
            keep->count[k] = 0;
    }
    // statistics: term generation and merge up to the evaluation; raw terms in the term_freqs field;
    // bytes = the terms written once (the direct ones in their final form, 32-40 B; the sorted ones
    // as keys and values, sorted and grouped once: 48 B)
    int const hg = stat_begin(ctx, POMX_STAT_G2_GEN);
    WsBuf<int> err(ctx, 1);
    int rc = pomx_fill_zero(ctx, err.p, sizeof(int));
    if(rc)
        return rc;
    // 1. rank tables (and, for direct terms, which bins of (i1, i2) and (i3, i4) are unique)
    WsBuf<uint32_t> r12, r23, r34;
    WsBuf<uint64_t> pairs;
    WsBuf<uint8_t> u12, u34;
    int64_t nr12 = 0, nr23 = 0, nr34 = 0;
    if((rc = rank_table(ctx, P.n1, P.n2, P.E1, P.E2, res, r12, &pairs, &nr12, err, direct ? &u12 : nullptr)) ||
       (rc = rank_table(ctx, P.n2, P.n3, P.E2, P.E3, res, r23, nullptr, &nr23, err)) ||
       (rc = rank_table(ctx, P.n3, P.n4, P.E3, P.E4, res, r34, nullptr, &nr34, err, direct ? &u34 : nullptr)))
        return rc;
    if(nr12 >= (int64_t(1) << 20) || nr23 >= (int64_t(1) << 21) || nr34 >= (int64_t(1) << 21))
        return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_g2_part_compute: too many distinct pole bins for the 64-bit key");
    // 2. count terms: sorted ones per row and pair, direct ones per pair and kind
    int64_t const npairs = int64_t(P.n1) * P.n2;
    int64_t const nrows = npairs * P.n3;
    WsBuf<double> O3c(ctx, int64_t(P.n3) * P.n4), C4t(ctx, int64_t(P.n1) * P.n4);
    g2_stage_kernel<<<grid_for(int64_t(P.n3) * P.n4 + int64_t(P.n1) * P.n4), kBlock, 0, ctx->stream>>>(D, O3c, C4t);
    // bounds that prune the sweep over i4 (row_bound, i4_may_pass)
    WsBuf<double> wsm(ctx, P.n4), cdm(ctx, int64_t(P.n1) * P.n3);
    if(!O3c.p || !C4t.p || !wsm.p || !cdm.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    suffix_max_kernel<<<1, 256, 0, ctx->stream>>>(P.n4, P.w4, wsm);
    cd_max_kernel<<<unsigned(std::min<int64_t>((int64_t(P.n1) * P.n3 + 3) / 4, 1 << 16)), kBlock, 0, ctx->stream>>>(
        D, O3c, C4t, cdm);
    POMX_CHECK_LAUNCH(ctx);
    D.wsm = wsm.p;
    D.cdm = cdm.p;
    D.bf = std::max(beta, 1.0) * (1.0 + 1e-9);
    WsBuf<int> rowcnt(ctx, nrows);
    WsBuf<int64_t> paircnt(ctx, npairs), dpc(ctx, 4 * npairs), doff(ctx, 4 * npairs);
    if(!rowcnt.p || !paircnt.p || !dpc.p || !doff.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
    unsigned const pair_blocks = unsigned(std::min<int64_t>((npairs + 3) / 4, 1 << 16)); // 4 waves per block
    g2_count_kernel<<<pair_blocks, kBlock, 0, ctx->stream>>>(D, O3c, C4t, pairs, npairs, u12.p, u34.p, rowcnt, paircnt,
                                                             dpc);
    POMX_CHECK_LAUNCH(ctx);
    for(int k = 0; k < 4; ++k)
        if((rc = pomx_exclusive_scan(ctx, npairs, dpc.p + k * npairs, doff.p + k * npairs)))
            return rc;
    std::vector<int64_t> hpair(npairs);
    std::vector<uint32_t> hr12(npairs);
    std::vector<uint64_t> hpairs(npairs);
    POMX_HIP(ctx, hipMemcpyAsync(hpair.data(), paircnt.p, sizeof(int64_t) * npairs, hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipMemcpyAsync(hr12.data(), r12.p, sizeof(uint32_t) * npairs, hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipMemcpyAsync(hpairs.data(), pairs.p, sizeof(uint64_t) * npairs, hipMemcpyDeviceToHost, ctx->stream));
    int64_t dlast[8];
    for(int k = 0; k < 4; ++k) {
        POMX_HIP(ctx, hipMemcpyAsync(dlast + 2 * k, doff.p + k * npairs + npairs - 1, sizeof(int64_t),
                                     hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipMemcpyAsync(dlast + 2 * k + 1, dpc.p + k * npairs + npairs - 1, sizeof(int64_t),
                                     hipMemcpyDeviceToHost, ctx->stream));
    }
    int herr = 0;
    POMX_HIP(ctx, hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if(herr)
        return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_g2_part_compute: |pole/pole_res| out of range");
    int64_t ndirect[4], direct_total = 0;
    for(int k = 0; k < 4; ++k) {
        ndirect[k] = dlast[2 * k] + dlast[2 * k + 1];
        direct_total += ndirect[k];
    }
    int64_t sorted_total = 0;
    struct Span {
        int64_t pb, pe, n;
    };
    std::vector<Span> chunks; // [pair_begin, pair_end) of the r12-sorted pairs, n sorted terms
    {
        int64_t const chunk_cap = chunk_terms();
        int64_t b = 0, acc = 0;
        for(int64_t p = 0; p < npairs; ++p) {
            bool boundary = (p == 0) || hr12[hpairs[p]] != hr12[hpairs[p - 1]];
            if(boundary && acc > 0 && acc + hpair[p] > chunk_cap) {
                chunks.push_back(Span{b, p, acc});
                b = p;
                acc = 0;
            }
            acc += hpair[p];
            sorted_total += hpair[p];
        }
        if(npairs > b && acc > 0)
            chunks.push_back(Span{b, npairs, acc});
    }
    int64_t const quads_total = direct_total + sorted_total;
    // 3. the part's kind lists (tangent path: with the derivative fields): direct terms first, then
    // the sorted groups
    KindLists lists{};
    std::vector<WsBuf<double>> list_store;
    int64_t fill[4];
    for(int k = 0; k < 4; ++k) {
        lists.cap[k] = std::max<int64_t>(ndirect[k] + sorted_total, 1);
        fill[k] = ndirect[k];
        int const nf = tan ? (k < 2 ? 8 : 10) : (k < 2 ? 4 : 5);
        list_store.emplace_back(ctx, size_t(lists.cap[k]) * nf);
        if(!list_store.back().p)
            return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
        lists.base[k] = list_store.back().p;
        counts[k] = ndirect[k];
    }
    // visiting positions (list bit, ((i1 n3 + i3) n2 + i2) n4 + i4, slot) must fit below the list gap
    bool const positions_ok = double(P.n1) * P.n2 * P.n3 * P.n4 * 2.0 < 9.2e18;
    bool const multi = chunks.size() > 1;
    G2Gen const T{O3c, C4t, rowcnt, r12, r23, r34, u12.p, u34.p};
    EraseRule rule{tol, nullptr, 0, nullptr, nullptr, 0};
    // 4. sorted terms, per chunk: generate; a chunk without merges (or, alone in its part, without a
    // merge the reference's threshold could decide otherwise) is final at once; the others are held
    // until the part's erase decisions are settled
    std::vector<G2Chunk> held_store;
    held_store.reserve(chunks.size());
    std::vector<WsBuf<uint64_t>> fk_store;
    std::vector<std::pair<const uint64_t*, int64_t>> fks;
    int64_t ng_sorted = 0;
    for(auto const& ch : chunks) {
        G2Chunk c;
        int const hs = stat_begin(ctx, POMX_STAT_G2_SORT);
        if((rc = g2_generate_chunk(ctx, D, T, hpairs, hpair, ch.pb, ch.pe, ch.n, c)))
            return rc;
        stat_end(ctx, hs, 0.0, double(ch.n), 48.0 * double(ch.n));
        ng_sorted += c.ng;
        int64_t const merges = c.n - c.ng;
        if(multi && positions_ok) { // every group's first position, for the thresholds of other chunks' merges
            fk_store.emplace_back(ctx, c.ng);
            if(!fk_store.back().p)
                return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
            g2_first_pos_kernel<<<grid_for(c.ng), kBlock, 0, ctx->stream>>>(D, c.ng, c.start, c.keys, c.vals,
                                                                             fk_store.back());
            POMX_CHECK_LAUNCH(ctx);
            fks.emplace_back(fk_store.back().p, c.ng);
        }
        // a pass with a threshold below every one the reference can use here, tol/(groups + 1),
        // tells whether any decision could differ: only a merge whose max(|c0|, |c1|) lies in
        // (that, tol] can
        bool final_now = merges == 0;
        double const thr = tol / double(c.ng + direct_total + 1);
        if(!final_now && !multi) {
            WsBuf<int> ambig(ctx, 1);
            if(!ambig.p)
                return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
            if((rc = pomx_fill_zero(ctx, ambig.p, sizeof(int))) || (rc = g2_alloc_groups(ctx, c)))
                return rc;
            g2_reduce_kernel<false><<<grid_for(c.ng), kBlock, 0, ctx->stream>>>(D, c.n, c.ng, c.start, c.keys, c.vals,
                                                                                  thr, rule, c.g, c.live, nullptr,
                                                                                  nullptr, nullptr, ambig);
            POMX_CHECK_LAUNCH(ctx);
            int ha = 0;
            POMX_HIP(ctx, hipMemcpyAsync(&ha, ambig.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
            c.reduced = ha == 0;
            final_now = c.reduced;
        }
        if(final_now) { // (the tangent path re-reduces with the same threshold as that pass)
            if((rc = g2_finish_chunk(ctx, D, c, thr, rule, Tp, lists, fill, counts)))
                return rc;
            continue;
        }
        if(!positions_ok)
            return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_g2_part_compute: blocks too large for the exact term merge");
        if(!multi) { // the only chunk: its own first positions
            fk_store.emplace_back(ctx, c.ng);
            if(!fk_store.back().p)
                return pomx_fail(ctx, POMX_ERR_HIP, "pomx_g2_part_compute: out of device memory");
            g2_first_pos_kernel<<<grid_for(c.ng), kBlock, 0, ctx->stream>>>(D, c.ng, c.start, c.keys, c.vals,
                                                                             fk_store.back());
            POMX_CHECK_LAUNCH(ctx);
            fks.emplace_back(fk_store.back().p, c.ng);
        }
        held_store.push_back(std::move(c));
    }
    // 5. direct terms (their first positions counted against the merges when some chunk is held),
    // then the held chunks' erase decisions over the whole part
    std::vector<G2Chunk*> held;
    for(auto& c : held_store)
        held.push_back(&c);
    WsBuf<uint64_t> qs;
    WsBuf<int64_t> hist, started;
    int64_t m = 0;
    if(!held.empty() && (rc = g2_merge_positions(ctx, D, held, qs, m, hist)))
        return rc;
    if(direct && direct_total > 0) {
        const uint64_t* dq = held.empty() ? nullptr : qs.p;
        unsigned long long* dh = held.empty() ? nullptr : reinterpret_cast<unsigned long long*>(hist.p);
        if(tan)
            g2_direct_emit_kernel<true><<<pair_blocks, kBlock, 0, ctx->stream>>>(D, TT, O3c, C4t, pairs, npairs, u12.p,
                                                                                u34.p, doff, lists, dq, m, dh);
        else
            g2_direct_emit_kernel<false><<<pair_blocks, kBlock, 0, ctx->stream>>>(D, TT, O3c, C4t, pairs, npairs, u12.p,
                                                                                 u34.p, doff, lists, dq, m, dh);
        POMX_CHECK_LAUNCH(ctx);
    }
    if(!held.empty()) {
        EraseEvents ev;
        if((rc = g2_settle(ctx, D, tol, held, fks, qs, m, hist, rule, ev, started)))
            return rc;
        fk_store.clear();
        for(G2Chunk* c : held)
            if((rc = g2_finish_chunk(ctx, D, *c, 0.0, rule, Tp, lists, fill, counts)))
                return rc;
    }
    stat_end(ctx, hg, 0.0, double(quads_total), 36.0 * double(direct_total) + 48.0 * double(sorted_total));
    // 6. the Lehmann sum over the kind lists (value path)
    if(nfreq > 0) {
        TermSpan spans[4];
        for(int k = 0; k < 4; ++k)
            spans[k] = TermSpan{lists.base[k], lists.cap[k], 0, fill[k]};
        if((rc = launch_g2_eval(ctx, spans, P.perm, res, nfreq, d_freq, d_acc, tan ? d_dacc : nullptr)))
            return rc;
    }
    if(keep) {
        int64_t nnr = counts[0] + counts[1], nr = counts[2] + counts[3];
        size_t bytes = sizeof(double) * size_t(4 * nnr + 5 * nr + 1);
        POMX_HIP(ctx, device_malloc(ctx, reinterpret_cast<void**>(&keep->data), bytes));
        double* NR = keep->data;
        double* R = keep->data + 4 * nnr;
        // kind-major packing: NR z2 | NR z4 in the non-resonant fields, R z2z3 | R z1z2 in the resonant ones
        int64_t const at[4] = {0, counts[0], 0, counts[2]};
        for(int k = 0; k < 4; ++k) {
            if(!fill[k])
                continue;
            int const nf = k < 2 ? 4 : 5;
            double* dst = k < 2 ? NR : R;
            int64_t const len = k < 2 ? nnr : nr;
            for(int f = 0; f < nf; ++f)
                POMX_HIP(ctx, hipMemcpyAsync(dst + f * len + at[k], lists.base[k] + f * lists.cap[k],
                                             sizeof(double) * fill[k], hipMemcpyDeviceToDevice, ctx->stream));
        }
        for(int k = 0; k < 4; ++k)
            keep->count[k] = counts[k];
    }
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if(h_counts)
        for(int k = 0; k < 4; ++k)
            h_counts[k] = counts[k];
    if(h_raw_terms)
        *h_raw_terms = quads_total;
    return POMX_OK;
}

} // namespace

extern "C" int pomx_g2_part_compute(pomx_ctx* ctx, const pomx_g2_part* part, double beta, double res, double tol,
                                    int nfreq, const double* d_freq, double* d_acc, pomx_g2_terms* keep,
                                    int64_t* h_counts, int64_t* h_raw_terms) {
    return g2_part_run(ctx, part, nullptr, beta, res, tol, nfreq, d_freq, d_acc, nullptr, keep, h_counts, h_raw_terms);
}

extern "C" int pomx_g2_part_tangent(pomx_ctx* ctx, const pomx_g2_part* part, const pomx_g2_tangent* tangent, double beta,
                                    double res, double tol, int nfreq, const double* d_freq, double* d_acc,
                                    double* d_dacc, int64_t* h_counts, int64_t* h_raw_terms) {
    POMX_REQUIRE(ctx, tangent != nullptr, "pomx_g2_part_tangent: no tangent");
    return g2_part_run(ctx, part, tangent, beta, res, tol, nfreq, d_freq, d_acc, d_dacc, nullptr, h_counts, h_raw_terms);
}
