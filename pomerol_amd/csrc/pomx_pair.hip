// K6/K10 (two-operator Lehmann terms), K8 (pole merge) and K9 (evaluation) for the
// single-particle Green's function and the dynamical susceptibility.
//
// Term generation follows GreensFunctionPart::computeImpl (GreensFunctionPart.cpp:65-110) and
// SusceptibilityPart::computeImpl (SusceptibilityPart.cpp:76-127): candidates are enumerated
// in the reference's emission order (part, outer index, inner index), flagged, compacted by a
// prefix sum, keyed by (part, round(pole/res)) and stably radix-sorted, so each group of a
// TermList (same hash bin, TermList.hpp:65-76) is contiguous and in emission order.  A
// group's residues are summed in that order and its first pole is kept (GF/chi terms do not
// average poles, GreensFunctionPart.cpp:34-37).  A running sum that becomes negligible
// (|R| <= tol/(size+1), size+1 the number of terms of the part's list at that moment) is
// dropped and the group restarts, as TermList::add_term does (EraseRule).
#include "pomx_internal.hpp"

#include <algorithm>
#include <mutex>

namespace {

struct PairPartDev {
    int n_out, n_in;
    pomx_view A, B;
    const double *E_out, *E_in, *w_out, *w_in;
    int64_t cand_off; // prefix of n_out * n_in
    int64_t tile_off; // prefix of the parts' kPairTile x kPairTile candidate tiles
    int tiles_in;     // tiles along the inner index
    int outer_fast;   // 1: the views are contiguous along the outer index (load the tile along it)
};

constexpr int kPairTile = 64; // candidate tile (outer x inner) of one workgroup in mark / emit

constexpr int kPartShift = 40;
constexpr uint64_t kBinMask = (uint64_t(1) << kPartShift) - 1;
constexpr int64_t kBinOffset = int64_t(1) << 39;
constexpr uint64_t kZeroPoleBin = kBinMask; // chi |P| < res entries

__device__ inline int find_part(const PairPartDev* parts, int nparts, int64_t g) {
    int lo = 0, hi = nparts - 1;
    while(lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if(parts[mid].cand_off <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

struct Cand {
    int part, i1, i2;
    double a, b, P;
};

__device__ inline Cand load_cand(const PairPartDev* parts, int nparts, int64_t g) {
    Cand c;
    c.part = find_part(parts, nparts, g);
    PairPartDev const& p = parts[c.part];
    int64_t loc = g - p.cand_off;
    c.i1 = int(loc / p.n_in);
    c.i2 = int(loc % p.n_in);
    c.a = p.A.base[c.i1 * p.A.rs + c.i2 * p.A.cs];
    c.b = p.B.base[c.i2 * p.B.rs + c.i1 * p.B.cs];
    c.P = p.E_in[c.i2] - p.E_out[c.i1];
    return c;
}

// 0: not emitted, 1: term, 2: chi zero-pole contribution
__device__ inline int classify(int kind, const PairPartDev* parts, Cand const& c, double res, double tol, double& R) {
    if(c.a == 0.0 || c.b == 0.0)
        return 0;
    PairPartDev const& p = parts[c.part];
    if(kind == POMX_KIND_GF) {
        R = c.a * c.b * (p.w_out[c.i1] + p.w_in[c.i2]);
        return fabs(R) > tol ? 1 : 0;
    }
    if(fabs(c.P) < res) {
        R = c.a * c.b * p.w_out[c.i1];
        return 2;
    }
    R = c.a * c.b * (p.w_out[c.i1] - p.w_in[c.i2]);
    return fabs(R) > tol ? 1 : 0;
}

__device__ inline int find_tile_part(const PairPartDev* parts, int nparts, int64_t t) {
    int lo = 0, hi = nparts - 1;
    while(lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if(parts[mid].tile_off <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Classes (classify) of the candidates of one kPairTile x kPairTile tile, cl[outer][inner] (0
// outside the part).  The views are read along the index they are contiguous in (outer_fast),
// so the strided operand of the emission order (inner fastest) is not read with a stride.
__device__ inline void tile_classes(int kind, const PairPartDev* parts, int part, int o0, int i0, double res, double tol,
                                    uint8_t (*cl)[kPairTile + 4]) {
    PairPartDev const& p = parts[part];
    int const lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
#pragma unroll 4
    for(int r = 0; r < kPairTile / 4; ++r) {
        int const u = lane, v = grp + 4 * r;
        int const x1 = p.outer_fast ? u : v, x2 = p.outer_fast ? v : u;
        int const i1 = o0 + x1, i2 = i0 + x2;
        int c = 0;
        if(i1 < p.n_out && i2 < p.n_in) {
            Cand q;
            q.part = part;
            q.i1 = i1;
            q.i2 = i2;
            q.a = p.A.base[i1 * p.A.rs + i2 * p.A.cs];
            q.b = p.B.base[i2 * p.B.rs + i1 * p.B.cs];
            q.P = p.E_in[i2] - p.E_out[i1];
            double R;
            c = classify(kind, parts, q, res, tol, R);
        }
        cl[x1][x2] = uint8_t(c);
    }
}

// One workgroup per candidate tile: flag[g] = 1 for an emitted candidate, g its position in the
// emission order (part, outer, inner); written along the inner index.
__global__ void __launch_bounds__(256) pair_mark_kernel(int kind, const PairPartDev* __restrict__ parts, int nparts,
                                                        double res, double tol, int64_t* __restrict__ flag) {
    __shared__ uint8_t cl[kPairTile][kPairTile + 4];
    int64_t const t = blockIdx.x;
    int const part = find_tile_part(parts, nparts, t);
    PairPartDev const& p = parts[part];
    int64_t const loc = t - p.tile_off;
    int const o0 = int(loc / p.tiles_in) * kPairTile, i0 = int(loc % p.tiles_in) * kPairTile;
    tile_classes(kind, parts, part, o0, i0, res, tol, cl);
    __syncthreads();
    int const lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    int const i2 = i0 + lane;
    if(i2 >= p.n_in)
        return;
    for(int r = 0; r < kPairTile / 4; ++r) {
        int const x1 = grp + 4 * r, i1 = o0 + x1;
        if(i1 < p.n_out)
            flag[p.cand_off + int64_t(i1) * p.n_in + i2] = cl[x1][lane] ? 1 : 0;
    }
}

// Same tiles: every emitted candidate g writes (part << kPartShift | pole bin, g) at pos[g].
__global__ void __launch_bounds__(256) pair_emit_kernel(int kind, const PairPartDev* __restrict__ parts, int nparts,
                                                        double res, double tol, const int64_t* __restrict__ pos,
                                                        int64_t nout, uint64_t* __restrict__ keys,
                                                        uint64_t* __restrict__ vals, int* __restrict__ err) {
    __shared__ uint8_t cl[kPairTile][kPairTile + 4];
    int64_t const t = blockIdx.x;
    int const part = find_tile_part(parts, nparts, t);
    PairPartDev const& p = parts[part];
    int64_t const loc = t - p.tile_off;
    int const o0 = int(loc / p.tiles_in) * kPairTile, i0 = int(loc % p.tiles_in) * kPairTile;
    tile_classes(kind, parts, part, o0, i0, res, tol, cl);
    __syncthreads();
    int const lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    int const i2 = i0 + lane;
    if(i2 >= p.n_in)
        return;
    double const e_in = p.E_in[i2];
    for(int r = 0; r < kPairTile / 4; ++r) {
        int const x1 = grp + 4 * r, i1 = o0 + x1;
        if(i1 >= p.n_out)
            break;
        int const c = cl[x1][lane];
        if(!c)
            continue;
        uint64_t bin;
        if(c == 2)
            bin = kZeroPoleBin;
        else {
            double q = round((e_in - p.E_out[i1]) / res);
            if(!(fabs(q) < double(kBinOffset - 1))) {
                err[0] = 1;
                q = 0;
            }
            bin = uint64_t(int64_t(q) + kBinOffset);
        }
        int64_t const g = p.cand_off + int64_t(i1) * p.n_in + i2;
        int64_t const at = pos[g];
        if(at >= nout)
            continue;
        keys[at] = (uint64_t(part) << kPartShift) | bin;
        vals[at] = uint64_t(g);
    }
}

// ---- G candidates with a weight bound (POMX_PAIR_UNIT_ELEMENTS) --------------------------
// With |A|, |B| <= 1 a candidate can pass |R| > tol only if |w_out| + |w_in| > tol, so only the
// rows of hot outer states (|w_out| > tol / 2, every inner index) and the hot inner columns of the
// other rows need to be read.  Compressed candidate space: row i1 holds n_in positions (hot row)
// or nhc positions (the hot columns, ascending); roff[i1] is its first position, rows in
// ascending i1, so the emission order (part, outer, inner) is kept.
struct PairHot {
    const int* hr;       // hot rows, ascending
    const int* cr;       // cold rows, ascending
    const int* hc;       // hot columns, ascending
    const int64_t* roff; // compressed position of (i1, first column of its row)
    int nhr, ncr, nhc, pad;
};

struct HotTile {
    int part, region, r0, c0; // region 0: hot rows x all columns, 1: cold rows x hot columns
};

__device__ inline void hot_classes(const PairPartDev* parts, const PairHot* hot, HotTile const& T, double res, double tol,
                                   uint8_t (*cl)[kPairTile + 4]) {
    PairPartDev const& p = parts[T.part];
    PairHot const& h = hot[T.part];
    const int* __restrict__ rows = T.region ? h.cr : h.hr;
    int const nrows = T.region ? h.ncr : h.nhr, ncols = T.region ? h.nhc : p.n_in;
    int const lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
#pragma unroll 4
    for(int r = 0; r < kPairTile / 4; ++r) {
        int const u = lane, v = grp + 4 * r;
        int const x1 = p.outer_fast ? u : v, x2 = p.outer_fast ? v : u;
        int const rr = T.r0 + x1, cc = T.c0 + x2;
        int c = 0;
        if(rr < nrows && cc < ncols) {
            Cand q;
            q.part = T.part;
            q.i1 = rows[rr];
            q.i2 = T.region ? h.hc[cc] : cc;
            q.a = p.A.base[q.i1 * p.A.rs + q.i2 * p.A.cs];
            q.b = p.B.base[q.i2 * p.B.rs + q.i1 * p.B.cs];
            q.P = p.E_in[q.i2] - p.E_out[q.i1];
            double R;
            c = classify(POMX_KIND_GF, parts, q, res, tol, R);
        }
        cl[x1][x2] = uint8_t(c);
    }
}

__global__ void __launch_bounds__(256) pair_hot_mark_kernel(const PairPartDev* __restrict__ parts, const PairHot* __restrict__ hot,
                                                            const HotTile* __restrict__ tiles, double res, double tol,
                                                            int64_t* __restrict__ flag) {
    __shared__ uint8_t cl[kPairTile][kPairTile + 4];
    HotTile const T = tiles[blockIdx.x];
    hot_classes(parts, hot, T, res, tol, cl);
    __syncthreads();
    PairHot const& h = hot[T.part];
    const int* __restrict__ rows = T.region ? h.cr : h.hr;
    int const nrows = T.region ? h.ncr : h.nhr, ncols = T.region ? h.nhc : parts[T.part].n_in;
    int const lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    int const cc = T.c0 + lane;
    if(cc >= ncols)
        return;
    for(int r = 0; r < kPairTile / 4; ++r) {
        int const x1 = grp + 4 * r, rr = T.r0 + x1;
        if(rr < nrows)
            flag[h.roff[rows[rr]] + cc] = cl[x1][lane] ? 1 : 0;
    }
}

__global__ void __launch_bounds__(256) pair_hot_emit_kernel(const PairPartDev* __restrict__ parts, const PairHot* __restrict__ hot,
                                                            const HotTile* __restrict__ tiles, double res, double tol,
                                                            const int64_t* __restrict__ pos, int64_t nout,
                                                            uint64_t* __restrict__ keys, uint64_t* __restrict__ vals,
                                                            int* __restrict__ err) {
    __shared__ uint8_t cl[kPairTile][kPairTile + 4];
    HotTile const T = tiles[blockIdx.x];
    hot_classes(parts, hot, T, res, tol, cl);
    __syncthreads();
    PairPartDev const& p = parts[T.part];
    PairHot const& h = hot[T.part];
    const int* __restrict__ rows = T.region ? h.cr : h.hr;
    int const nrows = T.region ? h.ncr : h.nhr, ncols = T.region ? h.nhc : p.n_in;
    int const lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    int const cc = T.c0 + lane;
    if(cc >= ncols)
        return;
    int const i2 = T.region ? h.hc[cc] : cc;
    double const e_in = p.E_in[i2];
    for(int r = 0; r < kPairTile / 4; ++r) {
        int const x1 = grp + 4 * r, rr = T.r0 + x1;
        if(rr >= nrows)
            break;
        if(!cl[x1][lane])
            continue;
        int const i1 = rows[rr];
        double q = round((e_in - p.E_out[i1]) / res);
        if(!(fabs(q) < double(kBinOffset - 1))) {
            err[0] = 1;
            q = 0;
        }
        int64_t const at = pos[h.roff[i1] + cc];
        if(at >= nout)
            continue;
        keys[at] = (uint64_t(T.part) << kPartShift) | uint64_t(int64_t(q) + kBinOffset);
        vals[at] = uint64_t(p.cand_off + int64_t(i1) * p.n_in + i2);
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

// One thread per group: walk members in emission order, which is the reference's visiting order
// (part, outer, inner), the candidate index being the visiting position.  The erase threshold is
// tol/(groups of the part + 1), or with R.fk set the reference's tol/(size+1) (EraseRule, each
// part its own TermList starting at its first candidate index); ambig and EMIT as in the 2PGF
// merge (pomx_g2.hip).
template <bool EMIT>
__global__ void pair_reduce_kernel(int kind, const PairPartDev* __restrict__ parts, int nparts, int64_t n,
                                   int64_t ngroups, const int64_t* __restrict__ start, const uint64_t* __restrict__ keys,
                                   const uint64_t* __restrict__ vals, const int64_t* __restrict__ part_groups,
                                   double res, double tol, EraseRule R, double* __restrict__ gR, double* __restrict__ gP,
                                   int64_t* __restrict__ live, double* __restrict__ zero_pole, int64_t* __restrict__ evn,
                                   uint64_t* __restrict__ ev_key_out, uint64_t* __restrict__ ev_w_out,
                                   int* __restrict__ ambig) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups)
        return;
    int64_t s = start[gi], e = (gi + 1 < ngroups) ? start[gi + 1] : n;
    uint64_t key = keys[s];
    int part = int(key >> kPartShift);
    if((key & kBinMask) == kZeroPoleBin) {
        if(EMIT)
            return;
        double z = 0.0;
        for(int64_t k = s; k < e; ++k) {
            Cand c = load_cand(parts, nparts, int64_t(vals[k]));
            double R;
            classify(kind, parts, c, res, tol, R);
            z += R;
        }
        zero_pole[part] = z;
        live[gi] = 0;
        if(evn)
            evn[gi] = 0;
        return;
    }
    double const thr = tol / double(part_groups[part] + 1);
    EraseCursor const cur = R.fk ? erase_cursor(R, uint64_t(parts[part].cand_off)) : EraseCursor{0, 0, 0};
    int64_t nevents = 0, at = EMIT ? evn[gi] : 0;
    bool alive = false;
    double sum = 0.0, pole = 0.0;
    for(int64_t k = s; k < e; ++k) {
        uint64_t const q = vals[k];
        Cand c = load_cand(parts, nparts, int64_t(q));
        double Rv;
        classify(kind, parts, c, res, tol, Rv);
        if(!alive) {
            if(k > s && R.fk) { // restart of an erased group
                if(EMIT) {
                    ev_key_out[at] = q;
                    ev_w_out[at++] = uint64_t(int64_t(-1));
                }
                ++nevents;
            }
            alive = true;
            sum = Rv;
            pole = c.P;
        } else {
            sum += Rv;
            double const t = R.fk ? erase_threshold(R, cur, q) : thr;
            if(ambig && fabs(sum) > thr && fabs(sum) <= tol)
                ambig[0] = 1;
            if(fabs(sum) <= t) {
                alive = false;
                if(R.fk) {
                    if(EMIT) {
                        ev_key_out[at] = q;
                        ev_w_out[at++] = 1;
                    }
                    ++nevents;
                }
            }
        }
    }
    if(!EMIT) {
        gR[gi] = sum;
        gP[gi] = pole;
        live[gi] = alive ? 1 : 0;
        if(evn)
            evn[gi] = nevents;
    }
}

// visiting position of each group's first member (zero-pole groups are no TermList terms and
// sort last)
__global__ void pair_first_kernel(int64_t ngroups, const int64_t* __restrict__ start, const uint64_t* __restrict__ keys,
                                  const uint64_t* __restrict__ vals, uint64_t* __restrict__ fk,
                                  uint64_t* __restrict__ gi_out) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi < ngroups) {
        uint64_t const key = keys[start[gi]];
        fk[gi] = (key & kBinMask) == kZeroPoleBin ? ~uint64_t(0) : vals[start[gi]];
        gi_out[gi] = uint64_t(gi);
    }
}

// Group ranges per part: groups are sorted by key and the part is the top key field, so
// first_group[p] (first group of part >= p) follows from a binary search; no atomics.
__device__ inline int64_t first_group_of(int64_t ngroups, const int64_t* start, const uint64_t* keys, uint64_t part) {
    int64_t lo = 0, hi = ngroups;
    while(lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if((keys[start[mid]] >> kPartShift) >= part)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// groups of each part that are pole terms (the zero-pole group of chi is excluded)
__global__ void part_group_count_kernel(int nparts, int64_t ngroups, const int64_t* __restrict__ start,
                                        const uint64_t* __restrict__ keys, int64_t* __restrict__ first,
                                        int64_t* __restrict__ cnt) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p > nparts)
        return;
    int64_t f = first_group_of(ngroups, start, keys, uint64_t(p));
    first[p] = f;
    if(p < nparts) {
        int64_t g = first_group_of(ngroups, start, keys, uint64_t(p) + 1) - f;
        if(g > 0 && (keys[start[f + g - 1]] & kBinMask) == kZeroPoleBin)
            --g;
        cnt[p] = g;
    }
}

// live terms of each part (zero-pole groups are never live)
__global__ void part_live_count_kernel(int nparts, int64_t ngroups, const int64_t* __restrict__ first,
                                       const int64_t* __restrict__ live, const int64_t* __restrict__ lpos,
                                       int64_t* __restrict__ cnt) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= nparts)
        return;
    auto before = [&](int64_t g) { return g < ngroups ? lpos[g] : lpos[ngroups - 1] + live[ngroups - 1]; };
    cnt[p] = before(first[p + 1]) - before(first[p]);
}

__global__ void compact_kernel(int64_t ngroups, const int64_t* __restrict__ live, const int64_t* __restrict__ pos,
                               const double* __restrict__ gR, const double* __restrict__ gP,
                               double* __restrict__ outR, double* __restrict__ outP) {
    int64_t gi = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(gi >= ngroups || !live[gi])
        return;
    int64_t at = pos[gi];
    outR[at] = gR[gi];
    outP[at] = gP[gi];
}

// ---- K9: acc[w] += sgn * sum_t R_t / (z_w - P_t) --------------------------------------------
constexpr int kEvalChunk = 256;

__device__ inline double rcp_nr(double x) {
    // reciprocal: hardware estimate refined by two Newton-Raphson steps
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__global__ void __launch_bounds__(256) pair_eval_kernel(int64_t nterms, int64_t slice_len,
                                                        const double* __restrict__ R, const double* __restrict__ P,
                                                        int nz, const double* __restrict__ z,
                                                        double* __restrict__ partial) {
    __shared__ double sR[kEvalChunk], sP[kEvalChunk];
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t t0 = int64_t(blockIdx.y) * slice_len;
    int64_t t1 = min(nterms, t0 + slice_len);
    double zr = 0, zi = 0;
    if(f < nz) {
        zr = z[2 * f];
        zi = z[2 * f + 1];
    }
    double ar = 0, ai = 0;
    for(int64_t c0 = t0; c0 < t1; c0 += kEvalChunk) {
        int64_t cnt = min<int64_t>(kEvalChunk, t1 - c0);
        __syncthreads();
        if(threadIdx.x < cnt) {
            sR[threadIdx.x] = R[c0 + threadIdx.x];
            sP[threadIdx.x] = P[c0 + threadIdx.x];
        }
        __syncthreads();
        for(int k = 0; k < cnt; ++k) {
            double dr = zr - sP[k];
            double inv = rcp_nr(fma(dr, dr, zi * zi));
            double s = sR[k] * inv;
            ar = fma(s, dr, ar);
            ai = fma(-s, zi, ai);
        }
    }
    if(f < nz) {
        int64_t o = (int64_t(blockIdx.y) * nz + f) * 2;
        partial[o] = ar;
        partial[o + 1] = ai;
    }
}

__global__ void pair_eval_reduce_kernel(int nz, int nslices, double sgn, const double* __restrict__ partial,
                                        double* __restrict__ acc) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= nz)
        return;
    double sr = 0, si = 0;
    for(int s = 0; s < nslices; ++s) {
        sr += partial[(int64_t(s) * nz + f) * 2];
        si += partial[(int64_t(s) * nz + f) * 2 + 1];
    }
    acc[2 * f] += sgn * sr;
    acc[2 * f + 1] += sgn * si;
}

// Imaginary-time Lehmann sums, thread per tau, terms staged through LDS; per-slice partials.
//   GF  (GreensFunctionPart.cpp:28-31):  P > 0: -R e^{-tau P} / (1 + e^{-beta P}),
//                                        else   -R e^{(beta - tau) P} / (e^{beta P} + 1);
//   CHI (SusceptibilityPart.cpp:31-35):  P > 0: -R e^{-tau P} / expm1(-beta P),
//                                        else    R e^{(beta - tau) P} / expm1(beta P).
// Both branches are the same function; each keeps its exponents non-positive.
template <int KIND>
__global__ void __launch_bounds__(256) pair_tau_kernel(int64_t nterms, int64_t slice_len, const double* __restrict__ R,
                                                       const double* __restrict__ P, int ntau,
                                                       const double* __restrict__ tau, double beta,
                                                       double* __restrict__ partial) {
    __shared__ double sR[kEvalChunk], sP[kEvalChunk], sD[kEvalChunk];
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t t0 = int64_t(blockIdx.y) * slice_len;
    int64_t t1 = min(nterms, t0 + slice_len);
    double const tf = (f < ntau) ? tau[f] : 0.0;
    double acc = 0.0;
    for(int64_t c0 = t0; c0 < t1; c0 += kEvalChunk) {
        int64_t cnt = min<int64_t>(kEvalChunk, t1 - c0);
        __syncthreads();
        if(threadIdx.x < cnt) {
            double const r = R[c0 + threadIdx.x], p = P[c0 + threadIdx.x];
            sP[threadIdx.x] = p;
            // residue over the statistical denominator, with the branch's sign
            double d;
            if(KIND == POMX_KIND_GF)
                d = (p > 0) ? 1.0 + exp(-beta * p) : exp(beta * p) + 1.0;
            else
                d = (p > 0) ? expm1(-beta * p) : -expm1(beta * p);
            sR[threadIdx.x] = -r / d;
            sD[threadIdx.x] = (p > 0) ? 0.0 : beta; // exponent origin: -tau P or (beta - tau) P
        }
        __syncthreads();
        for(int k = 0; k < cnt; ++k)
            acc = fma(sR[k], exp((sD[k] - tf) * sP[k]), acc);
    }
    if(f < ntau)
        partial[int64_t(blockIdx.y) * ntau + f] = acc;
}

__global__ void pair_tau_reduce_kernel(int ntau, int nslices, const double* __restrict__ partial,
                                       double* __restrict__ acc) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= ntau)
        return;
    double s = 0;
    for(int sl = 0; sl < nslices; ++sl)
        s += partial[int64_t(sl) * ntau + f];
    acc[2 * f] += s;
}

} // namespace

extern "C" int pomx_pair_eval_tau(pomx_ctx* ctx, int kind, int64_t nterms, const double* d_residue,
                                  const double* d_pole, int ntau, const double* d_tau, double beta, double* d_acc) {
    POMX_REQUIRE(ctx, (kind == POMX_KIND_GF || kind == POMX_KIND_CHI) && nterms >= 0 && ntau >= 0 && beta > 0.0,
                 "pomx_pair_eval_tau: bad arguments");
    if(nterms == 0 || ntau == 0)
        return POMX_OK;
    int ftiles = (ntau + 255) / 256;
    int64_t want = std::max<int64_t>(1, 2048 / ftiles);
    int64_t slice = std::max<int64_t>(kEvalChunk, (nterms + want - 1) / want);
    slice = (slice + kEvalChunk - 1) / kEvalChunk * kEvalChunk;
    int nslices = int((nterms + slice - 1) / slice);
    WsBuf<double> partial(ctx, size_t(nslices) * ntau);
    if(!partial.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_pair_eval_tau: out of device memory");
    if(kind == POMX_KIND_GF)
        pair_tau_kernel<POMX_KIND_GF><<<dim3(ftiles, nslices), 256, 0, ctx->stream>>>(nterms, slice, d_residue, d_pole,
                                                                                    ntau, d_tau, beta, partial);
    else
        pair_tau_kernel<POMX_KIND_CHI><<<dim3(ftiles, nslices), 256, 0, ctx->stream>>>(nterms, slice, d_residue, d_pole,
                                                                                     ntau, d_tau, beta, partial);
    pair_tau_reduce_kernel<<<grid_for(ntau), kBlock, 0, ctx->stream>>>(ntau, nslices, partial, d_acc);
    POMX_CHECK_LAUNCH(ctx);
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

extern "C" int pomx_pair_terms(pomx_ctx* ctx, int kind_flags, int nparts, const pomx_pair_part* h_parts, double res,
                               double tol, double** d_residue, double** d_pole, int64_t* h_count, double* h_zero_pole) {
    int const kind = kind_flags & ~int(POMX_PAIR_UNIT_ELEMENTS);
    POMX_REQUIRE(ctx, (kind == POMX_KIND_GF || kind == POMX_KIND_CHI) && nparts >= 0 && d_residue && d_pole && h_count,
                 "pomx_pair_terms: bad arguments");
    POMX_REQUIRE(ctx, res > 0.0, "pomx_pair_terms: pole resolution must be positive");
    POMX_REQUIRE(ctx, nparts < (1 << 23), "pomx_pair_terms: too many parts");
    *d_residue = nullptr;
    *d_pole = nullptr;
    std::vector<PairPartDev> parts(nparts);
    int64_t ncand = 0, ntiles = 0;
    for(int p = 0; p < nparts; ++p) {
        pomx_pair_part const& q = h_parts[p];
        POMX_REQUIRE(ctx, q.n_out >= 1 && q.n_in >= 1, "pomx_pair_terms: empty part");
        int const to = (q.n_out + kPairTile - 1) / kPairTile, ti = (q.n_in + kPairTile - 1) / kPairTile;
        // load along the outer index when more of the two operands are contiguous along it
        int const ofast = int((q.A.rs == 1) + (q.B.cs == 1) > (q.A.cs == 1) + (q.B.rs == 1));
        parts[p] = PairPartDev{q.n_out, q.n_in, q.A, q.B, q.E_out, q.E_in, q.w_out, q.w_in, ncand, ntiles, ti, ofast};
        ncand += int64_t(q.n_out) * q.n_in;
        ntiles += int64_t(to) * ti;
        h_count[p] = 0;
        if(h_zero_pole)
            h_zero_pole[p] = 0.0;
    }
    if(ncand == 0) {
        void* a = nullptr;
        void* b = nullptr;
        POMX_HIP(ctx, device_malloc(ctx, &a, 8));
        POMX_HIP(ctx, device_malloc(ctx, &b, 8));
        *d_residue = static_cast<double*>(a);
        *d_pole = static_cast<double*>(b);
        return POMX_OK;
    }
    WsBuf<PairPartDev> d_parts(ctx, nparts);
    WsBuf<int> err(ctx, 1);
    POMX_HIP(ctx, h2d_sync(ctx, d_parts.p, parts.data(), sizeof(PairPartDev) * nparts));
    int rc = pomx_fill_zero(ctx, err.p, sizeof(int));
    if(rc)
        return rc;
    // G with unit-bounded elements: split every part into hot and cold states by the weights
    // (host side; the weight vectors are short) and enumerate the compressed candidate space
    bool hot = kind == POMX_KIND_GF && (kind_flags & POMX_PAIR_UNIT_ELEMENTS) && tol > 0.0;
    WsBuf<int> hot_rows, hot_cols;
    WsBuf<int64_t> hot_roff;
    std::vector<PairHot> hh;
    std::vector<int> hrows, hcols;
    std::vector<int64_t> hroff;
    std::vector<HotTile> htiles;
    int64_t nq = 0;
    if(hot) {
        // |a b (w_out + w_in)| <= (1 + 1e-9)^2 (|w_out| + |w_in|) (1 + 4u): a pair of states with
        // both |w| <= thr cannot pass |R| > tol
        double const thr = 0.5 * tol / (1.0 + 1e-8);
        struct Off {
            size_t hr, cr, hc, roff;
        };
        std::vector<Off> off(nparts);
        std::vector<double> wo, wi;
        for(int p = 0; p < nparts && hot; ++p) {
            pomx_pair_part const& q = h_parts[p];
            wo.resize(size_t(q.n_out));
            wi.resize(size_t(q.n_in));
            POMX_HIP(ctx, hipMemcpyAsync(wo.data(), q.w_out, sizeof(double) * q.n_out, hipMemcpyDeviceToHost, ctx->stream));
            POMX_HIP(ctx, hipMemcpyAsync(wi.data(), q.w_in, sizeof(double) * q.n_in, hipMemcpyDeviceToHost, ctx->stream));
            POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
            off[p].hr = hrows.size();
            for(int i = 0; i < q.n_out; ++i)
                if(!(fabs(wo[size_t(i)]) <= thr)) // (NaN counts as hot)
                    hrows.push_back(i);
            int const nhr = int(hrows.size() - off[p].hr);
            off[p].cr = hrows.size();
            for(int i = 0; i < q.n_out; ++i)
                if(fabs(wo[size_t(i)]) <= thr)
                    hrows.push_back(i);
            off[p].hc = hcols.size();
            for(int i = 0; i < q.n_in; ++i)
                if(!(fabs(wi[size_t(i)]) <= thr))
                    hcols.push_back(i);
            int const nhc = int(hcols.size() - off[p].hc);
            off[p].roff = hroff.size();
            for(int i = 0, at = 0; i < q.n_out; ++i) {
                bool const h = at < nhr && hrows[off[p].hr + size_t(at)] == i;
                hroff.push_back(nq);
                nq += h ? q.n_in : nhc;
                at += h;
            }
            PairHot x{nullptr, nullptr, nullptr, nullptr, nhr, q.n_out - nhr, nhc, 0};
            hh.push_back(x);
            for(int r0 = 0; r0 < x.nhr; r0 += kPairTile)
                for(int c0 = 0; c0 < q.n_in; c0 += kPairTile)
                    htiles.push_back(HotTile{p, 0, r0, c0});
            for(int r0 = 0; r0 < x.ncr; r0 += kPairTile)
                for(int c0 = 0; c0 < x.nhc; c0 += kPairTile)
                    htiles.push_back(HotTile{p, 1, r0, c0});
        }
        if(2 * nq > ncand) // little to skip: the plain enumeration reads fewer index lists
            hot = false;
        else {
            WsBuf<int> dr(ctx, std::max<size_t>(hrows.size(), 1)), dc(ctx, std::max<size_t>(hcols.size(), 1));
            WsBuf<int64_t> dro(ctx, std::max<size_t>(hroff.size(), 1));
            if(!dr.p || !dc.p || !dro.p)
                return pomx_fail(ctx, POMX_ERR_HIP, "pomx_pair_terms: out of device memory");
            for(int p = 0; p < nparts; ++p) {
                hh[size_t(p)].hr = dr.p + off[p].hr;
                hh[size_t(p)].cr = dr.p + off[p].cr;
                hh[size_t(p)].hc = dc.p + off[p].hc;
                hh[size_t(p)].roff = dro.p + off[p].roff;
            }
            if(!hrows.empty())
                POMX_HIP(ctx, h2d_sync(ctx, dr.p, hrows.data(), sizeof(int) * hrows.size()));
            if(!hcols.empty())
                POMX_HIP(ctx, h2d_sync(ctx, dc.p, hcols.data(), sizeof(int) * hcols.size()));
            POMX_HIP(ctx, h2d_sync(ctx, dro.p, hroff.data(), sizeof(int64_t) * hroff.size()));
            hot_rows = std::move(dr);
            hot_cols = std::move(dc);
            hot_roff = std::move(dro);
        }
    }
    int64_t const nflag = hot ? nq : ncand;
    WsBuf<PairHot> d_hot(ctx, hot ? hh.size() : 1);
    WsBuf<HotTile> d_htiles(ctx, hot ? std::max<size_t>(htiles.size(), 1) : 1);
    if(hot) {
        if(!d_hot.p || !d_htiles.p)
            return pomx_fail(ctx, POMX_ERR_HIP, "pomx_pair_terms: out of device memory");
        POMX_HIP(ctx, h2d_sync(ctx, d_hot.p, hh.data(), sizeof(PairHot) * hh.size()));
        if(!htiles.empty())
            POMX_HIP(ctx, h2d_sync(ctx, d_htiles.p, htiles.data(), sizeof(HotTile) * htiles.size()));
    }
    WsBuf<int64_t> flag(ctx, size_t(std::max<int64_t>(nflag, 1))), pos(ctx, size_t(std::max<int64_t>(nflag, 1)));
    if(!flag.p || !pos.p)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_pair_terms: out of device memory");
    int64_t n = 0;
    if(nflag > 0) {
        if(hot) {
            POMX_REQUIRE(ctx, htiles.size() < (size_t(1) << 31), "pomx_pair_terms: too many candidates");
            pair_hot_mark_kernel<<<unsigned(htiles.size()), 256, 0, ctx->stream>>>(d_parts, d_hot, d_htiles, res, tol, flag);
        } else {
            POMX_REQUIRE(ctx, ntiles < (int64_t(1) << 31), "pomx_pair_terms: too many candidates");
            pair_mark_kernel<<<unsigned(ntiles), 256, 0, ctx->stream>>>(kind, d_parts, nparts, res, tol, flag);
        }
        POMX_CHECK_LAUNCH(ctx);
        if((rc = pomx_exclusive_scan(ctx, nflag, flag, pos)))
            return rc;
        int64_t last_flag = 0, last_pos = 0;
        POMX_HIP(ctx, hipMemcpyAsync(&last_flag, flag.p + nflag - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipMemcpyAsync(&last_pos, pos.p + nflag - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        n = last_pos + last_flag;
    }
    std::vector<int64_t> cnt(nparts, 0);
    std::vector<double> zp(nparts, 0.0);
    double *outR = nullptr, *outP = nullptr;
    if(n > 0) {
        WsBuf<uint64_t> k0(ctx, n), k1(ctx, n), v0(ctx, n), v1(ctx, n);
        if(hot)
            pair_hot_emit_kernel<<<unsigned(htiles.size()), 256, 0, ctx->stream>>>(d_parts, d_hot, d_htiles, res, tol, pos, n,
                                                                                   k0, v0, err);
        else
            pair_emit_kernel<<<unsigned(ntiles), 256, 0, ctx->stream>>>(kind, d_parts, nparts, res, tol, pos, n, k0, v0, err);
        POMX_CHECK_LAUNCH(ctx);
        int part_bits = 1;
        while((1 << part_bits) < nparts)
            ++part_bits;
        if((rc = pomx_sort_pairs(ctx, n, k0, k1, v0, v1, 0, kPartShift + part_bits)))
            return rc;
        WsBuf<int64_t> head(ctx, n), gid(ctx, n);
        heads_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, k1, head);
        if((rc = pomx_exclusive_scan(ctx, n, head, gid)))
            return rc;
        int64_t lh = 0, lg = 0;
        POMX_HIP(ctx, hipMemcpyAsync(&lh, head.p + n - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipMemcpyAsync(&lg, gid.p + n - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        int64_t ngroups = lh + lg;
        WsBuf<int64_t> start(ctx, ngroups), live(ctx, ngroups), lpos(ctx, ngroups), pgroups(ctx, nparts);
        WsBuf<int64_t> first(ctx, nparts + 1), d_cnt(ctx, nparts);
        WsBuf<double> gR(ctx, ngroups), gP(ctx, ngroups), d_zp(ctx, nparts);
        group_start_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, head, gid, start);
        if((rc = pomx_fill_zero(ctx, d_zp.p, sizeof(double) * nparts)))
            return rc;
        part_group_count_kernel<<<grid_for(nparts + 1), kBlock, 0, ctx->stream>>>(nparts, ngroups, start, k1, first, pgroups);
        // TermList's erase threshold tol/(size+1): a pass with tol/(groups of the part + 1) is exact
        // unless some merge is ambiguous; then the reference's rule is settled (EraseRule)
        WsBuf<int> ambig(ctx, 1);
        if(!ambig.p)
            return pomx_fail(ctx, POMX_ERR_HIP, "pomx_pair_terms: out of device memory");
        if((rc = pomx_fill_zero(ctx, ambig.p, sizeof(int))))
            return rc;
        EraseRule rule{tol, nullptr, 0, nullptr, nullptr, 0};
        pair_reduce_kernel<false><<<grid_for(ngroups), kBlock, 0, ctx->stream>>>(
            kind, d_parts, nparts, n, ngroups, start, k1, v1, pgroups, res, tol, rule, gR, gP, live, d_zp, nullptr, nullptr,
            nullptr, ambig);
        POMX_CHECK_LAUNCH(ctx);
        int ha = 0;
        POMX_HIP(ctx, hipMemcpyAsync(&ha, ambig.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if(ha) {
            WsBuf<uint64_t> f0(ctx, ngroups), f1(ctx, ngroups), i0(ctx, ngroups), i1(ctx, ngroups);
            if(!f0.p || !f1.p || !i0.p || !i1.p)
                return pomx_fail(ctx, POMX_ERR_HIP, "pomx_pair_terms: out of device memory");
            pair_first_kernel<<<grid_for(ngroups), kBlock, 0, ctx->stream>>>(ngroups, start, k1, v1, f0, i0);
            POMX_CHECK_LAUNCH(ctx);
            if((rc = pomx_sort_pairs(ctx, ngroups, f0, f1, i0, i1, 0, 64)))
                return rc;
            rule.fk = f1.p;
            rule.nfk = ngroups;
            EraseEvents ev;
            auto pass = [&](EraseRule const& R, bool emit, int64_t* evn, uint64_t* ko, uint64_t* wo) -> int {
                if(emit)
                    pair_reduce_kernel<true><<<grid_for(ngroups), kBlock, 0, ctx->stream>>>(
                        kind, d_parts, nparts, n, ngroups, start, k1, v1, pgroups, res, tol, R, gR, gP, live, d_zp, evn,
                        ko, wo, nullptr);
                else
                    pair_reduce_kernel<false><<<grid_for(ngroups), kBlock, 0, ctx->stream>>>(
                        kind, d_parts, nparts, n, ngroups, start, k1, v1, pgroups, res, tol, R, gR, gP, live, d_zp, evn,
                        ko, wo, nullptr);
                POMX_CHECK_LAUNCH(ctx);
                return POMX_OK;
            };
            if((rc = erase_fixed_point(ctx, ngroups, rule, pass, ev, "pomx_pair_terms")))
                return rc;
        }
        if((rc = pomx_exclusive_scan(ctx, ngroups, live, lpos)))
            return rc;
        int64_t ll = 0, lp = 0;
        POMX_HIP(ctx, hipMemcpyAsync(&ll, live.p + ngroups - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipMemcpyAsync(&lp, lpos.p + ngroups - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        int64_t nlive = ll + lp;
        POMX_HIP(ctx, device_malloc(ctx, reinterpret_cast<void**>(&outR), sizeof(double) * std::max<int64_t>(nlive, 1)));
        POMX_HIP(ctx, device_malloc(ctx, reinterpret_cast<void**>(&outP), sizeof(double) * std::max<int64_t>(nlive, 1)));
        compact_kernel<<<grid_for(ngroups), kBlock, 0, ctx->stream>>>(ngroups, live, lpos, gR, gP, outR, outP);
        part_live_count_kernel<<<grid_for(nparts), kBlock, 0, ctx->stream>>>(nparts, ngroups, first, live, lpos, d_cnt);
        POMX_CHECK_LAUNCH(ctx);
        POMX_HIP(ctx, hipMemcpyAsync(cnt.data(), d_cnt.p, sizeof(int64_t) * nparts, hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipMemcpyAsync(zp.data(), d_zp.p, sizeof(double) * nparts, hipMemcpyDeviceToHost, ctx->stream));
        int herr = 0;
        POMX_HIP(ctx, hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if(herr) {
            hipFree(outR);
            hipFree(outP);
            return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_pair_terms: |pole/pole_res| exceeds 2^39");
        }
    } else {
        POMX_HIP(ctx, device_malloc(ctx, reinterpret_cast<void**>(&outR), 8));
        POMX_HIP(ctx, device_malloc(ctx, reinterpret_cast<void**>(&outP), 8));
    }
    for(int p = 0; p < nparts; ++p) {
        h_count[p] = int64_t(cnt[p]);
        if(h_zero_pole)
            h_zero_pole[p] = zp[p];
    }
    *d_residue = outR;
    *d_pole = outP;
    return POMX_OK;
}

extern "C" int pomx_pair_eval(pomx_ctx* ctx, int kind, int64_t nterms, const double* d_residue, const double* d_pole,
                              int nz, const double* d_z, double* d_acc) {
    POMX_REQUIRE(ctx, (kind == POMX_KIND_GF || kind == POMX_KIND_CHI) && nterms >= 0 && nz >= 0,
                 "pomx_pair_eval: bad arguments");
    if(nterms == 0 || nz == 0)
        return POMX_OK;
    int ftiles = (nz + 255) / 256;
    int64_t want = std::max<int64_t>(1, 2048 / ftiles);
    int64_t slice = std::max<int64_t>(kEvalChunk, (nterms + want - 1) / want);
    slice = (slice + kEvalChunk - 1) / kEvalChunk * kEvalChunk;
    int nslices = int((nterms + slice - 1) / slice);
    WsBuf<double> partial(ctx, size_t(nslices) * nz * 2);
    int const h = stat_begin(ctx, POMX_STAT_PAIR);
    pair_eval_kernel<<<dim3(ftiles, nslices), 256, 0, ctx->stream>>>(nterms, slice, d_residue, d_pole, nz, d_z, partial);
    pair_eval_reduce_kernel<<<grid_for(nz), kBlock, 0, ctx->stream>>>(nz, nslices, kind == POMX_KIND_GF ? 1.0 : -1.0,
                                                                       partial, d_acc);
    POMX_CHECK_LAUNCH(ctx);
    // 14 FLOP per term and frequency (R / (z - P)); residue and pole read once
    stat_end(ctx, h, 14.0 * double(nterms) * nz, double(nterms) * nz, 16.0 * double(nterms) + 32.0 * nz);
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}

// ---- the Lehmann sum without allocation ----------------------------------------------------
// pomx_pair_eval_noalloc keeps its per-slice partial sums in this file-scope device array (16 MiB
// on every device the library runs on, held until the process ends) instead of the context's
// pool.  Host code never passes the array's name to a kernel: it stands for a host-side shadow;
// the device address comes from hipGetSymbolAddress, once per device, at context set-up.
constexpr int64_t kPairScratchDoubles = int64_t(1) << 21;
__device__ __align__(256) double scratch[kPairScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
} // namespace

// Called by pomx_ctx_create with ctx->device current.
int pair_scratch_setup(pomx_ctx* ctx) {
    if(ctx->device < 0 || ctx->device >= kMaxDevices)
        return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_ctx_create: device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if(g_scratch[ctx->device])
        return POMX_OK;
    void* p = nullptr;
    hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if(e != hipSuccess || !p)
        return pomx_fail(ctx, POMX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[ctx->device] = static_cast<double*>(p);
    return POMX_OK;
}

extern "C" int pomx_pair_eval_noalloc(pomx_ctx* ctx, int kind, int64_t nterms, const double* d_residue,
                                      const double* d_pole, int nz, const double* d_z, double* d_acc) {
    POMX_REQUIRE(ctx, (kind == POMX_KIND_GF || kind == POMX_KIND_CHI) && nterms >= 0 && nz >= 0,
                 "pomx_pair_eval_noalloc: bad arguments");
    if(nterms == 0 || nz == 0)
        return POMX_OK;
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        part = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch[ctx->device] : nullptr;
    }
    if(!part)
        return pomx_fail(ctx, POMX_ERR_HIP, "pomx_pair_eval_noalloc: no scratch address for this device");
    int64_t const max_slices = kPairScratchDoubles / (2 * int64_t(nz));
    if(max_slices < 1)
        return pomx_fail(ctx, POMX_ERR_RANGE, "pomx_pair_eval_noalloc: more than 1048576 frequencies");
    // the slicing of pomx_pair_eval, with fewer and longer slices when its partial sums would not fit
    int ftiles = (nz + 255) / 256;
    int64_t want = std::max<int64_t>(1, 2048 / ftiles);
    int64_t slice = std::max<int64_t>(kEvalChunk, (nterms + want - 1) / want);
    slice = (slice + kEvalChunk - 1) / kEvalChunk * kEvalChunk;
    if((nterms + slice - 1) / slice > max_slices) {
        slice = (nterms + max_slices - 1) / max_slices;
        slice = (slice + kEvalChunk - 1) / kEvalChunk * kEvalChunk;
    }
    int nslices = int((nterms + slice - 1) / slice);
    // every (slice, frequency) entry of the partial sums is written before it is read: no clearing
    pair_eval_kernel<<<dim3(ftiles, nslices), 256, 0, ctx->stream>>>(nterms, slice, d_residue, d_pole, nz, d_z, part);
    pair_eval_reduce_kernel<<<grid_for(nz), kBlock, 0, ctx->stream>>>(nz, nslices, kind == POMX_KIND_GF ? 1.0 : -1.0,
                                                                       part, d_acc);
    POMX_CHECK_LAUNCH(ctx);
    POMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return POMX_OK;
}
