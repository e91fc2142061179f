// Internal definitions shared by the gfx950 kernels behind include/pomx_c.h.
#pragma once

#include "../../include/pomx_c.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <vector>

struct pomx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // caching pool for temporaries: size -> free blocks
    std::multimap<size_t, void*> free_blocks;
    std::map<void*, size_t> live_blocks;
    // timing of the Lehmann evaluation kernels (the dominant kernel) and the term generation:
    // event pairs recorded per launch group, resolved when read (or when many are pending)
    bool timing = false;
    struct StatRec {
        int cls;
        hipEvent_t a, b;
        double flop, term_freqs, bytes;
    };
    std::vector<StatRec> stat_recs;
    std::vector<hipEvent_t> stat_pool;
    pomx_kernel_stat stats[POMX_STAT_NCLASS] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

int pomx_fail(pomx_ctx* ctx, int code, std::string const& msg);

#define POMX_HIP(ctx, expr)                                                                          \
    do {                                                                                             \
        hipError_t pomx_e_ = (expr);                                                                 \
        if(pomx_e_ != hipSuccess)                                                                    \
            return pomx_fail(ctx, POMX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(pomx_e_)); \
    } while(0)

#define POMX_CHECK_LAUNCH(ctx) POMX_HIP(ctx, hipGetLastError())

#define POMX_REQUIRE(ctx, cond, msg)                                                                 \
    do {                                                                                             \
        if(!(cond))                                                                                  \
            return pomx_fail(ctx, POMX_ERR_ARG, msg);                                                \
    } while(0)

// Temporary device memory from the context's caching pool (returned with ws_put).
void* ws_get(pomx_ctx* ctx, size_t bytes);
void ws_put(pomx_ctx* ctx, void* p);

// RAII holder for a pooled temporary.
template <class T> struct WsBuf {
    pomx_ctx* ctx = nullptr;
    T* p = nullptr;
    WsBuf() = default;
    WsBuf(pomx_ctx* c, size_t n) : ctx(c), p(static_cast<T*>(ws_get(c, (n ? n : 1) * sizeof(T)))) {}
    WsBuf(WsBuf const&) = delete;
    WsBuf& operator=(WsBuf const&) = delete;
    WsBuf(WsBuf&& o) noexcept : ctx(o.ctx), p(o.p) { o.p = nullptr; }
    WsBuf& operator=(WsBuf&& o) noexcept {
        if(this != &o) {
            if(p)
                ws_put(ctx, p);
            ctx = o.ctx;
            p = o.p;
            o.p = nullptr;
        }
        return *this;
    }
    ~WsBuf() {
        if(p)
            ws_put(ctx, p);
    }
    operator T*() const { return p; }
};

// hipMalloc, retried once after the cached workspace blocks are released.  A failed attempt
// leaves its error as the runtime's last error, which a later hipGetLastError() launch check
// would report: it is consumed here.  Every device allocation of the library goes through it.
hipError_t device_malloc(pomx_ctx* ctx, void** p, size_t bytes);

// Stable LSD radix sort of (uint64 key, uint64 value) pairs on the context stream.
// keys_in/vals_in are consumed; sorted output lands in keys_out/vals_out.
int pomx_sort_pairs(pomx_ctx* ctx, int64_t n, uint64_t* keys_in, uint64_t* keys_out, uint64_t* vals_in,
                    uint64_t* vals_out, int begin_bit, int end_bit);
// Exclusive prefix sum of int64 (out may alias in).
int pomx_exclusive_scan(pomx_ctx* ctx, int64_t n, const int64_t* in, int64_t* out);
// Host-to-device copy from a host buffer the caller may release as soon as this returns: the
// copy has finished (context stream synchronised) on return.
inline hipError_t h2d_sync(pomx_ctx* ctx, void* dst, const void* src, size_t bytes) {
    hipError_t const e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    return e == hipSuccess ? hipStreamSynchronize(ctx->stream) : e;
}

// Copy one int64 from the device (synchronises the context stream).
int read_i64(pomx_ctx* ctx, const int64_t* d, int64_t* h);
// Dense ranks of round((Ey[y] - Ex[x]) / res) over all (x, y), stored at rank[x * ny + y]; equal
// bins get equal ranks, ascending with the bin.  'sorted' (optional) receives the x * ny + y
// indices in rank order; *nranks the number of distinct bins; d_err[0] is set to 1 where
// |bin| >= 1e18; 'uniq' (optional) receives 1 at x * ny + y where no other entry shares its bin.
int rank_table(pomx_ctx* ctx, int nx, int ny, const double* Ex, const double* Ey, double res, WsBuf<uint32_t>& rank,
               WsBuf<uint64_t>* sorted, int64_t* nranks, int* d_err, WsBuf<uint8_t>* uniq = nullptr);

// ---- TermList erase rule (the 3-point susceptibility and 2PGF merges) ---------------------
// A TermList erases a merged term when |C| <= tol/(size+1), size+1 being the number of terms in
// that list at that moment (TermList.hpp, add_term).  For a one-thread-per-group merge that
// number is the count of groups of the same list whose first member precedes the merge in the
// reference's visiting order, less those erased and not restarted before it.  Positions carry
// the list in their top bits, so one ascending array holds every list.

// lo + number of entries of the ascending a[lo, hi) that are <= q
__device__ inline int64_t upper_index(const uint64_t* __restrict__ a, int64_t lo, int64_t hi, uint64_t q) {
    while(lo < hi) {
        int64_t const mid = lo + (hi - lo) / 2;
        if(a[mid] <= q)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

struct EraseRule {
    double tol;
    const uint64_t* fk;     // (list, first position) of every group, ascending; with qstarted set:
                            // the (list, position) of every merge, ascending
    int64_t nfk;
    const uint64_t* ev_key; // assumed erase (+1) and restart (-1) events, (list, position) ascending
    const int64_t* ev_cum;  // inclusive prefix sums of their weights
    int64_t nev;
    // merge-position mode (two lists, the second with bit 63 set): qstarted[j] is the number of
    // groups of either list whose first member precedes the merge at fk[j]; base_hi the number of
    // groups of the first list
    const int64_t* qstarted = nullptr;
    int64_t base_hi = 0;
};

struct EraseCursor {
    int64_t fk_lo, ev_lo, ev_base;
};

// where the list with top bits 'lbits' starts in fk and in the events
__device__ inline EraseCursor erase_cursor(EraseRule const& R, uint64_t lbits) {
    EraseCursor c;
    c.fk_lo = (lbits && !R.qstarted) ? upper_index(R.fk, 0, R.nfk, lbits - 1) : 0;
    c.ev_lo = (R.nev > 0 && lbits) ? upper_index(R.ev_key, 0, R.nev, lbits - 1) : 0;
    c.ev_base = c.ev_lo > 0 ? R.ev_cum[c.ev_lo - 1] : 0;
    return c;
}

// threshold of a merge at position q (q > 0, list bits included)
__device__ inline double erase_threshold(EraseRule const& R, EraseCursor const& c, uint64_t q) {
    int64_t started;
    if(R.qstarted) // q is one of the merges listed in fk
        started = R.qstarted[upper_index(R.fk, 0, R.nfk, q) - 1] - ((q >> 63) ? R.base_hi : 0);
    else
        started = upper_index(R.fk, c.fk_lo, R.nfk, q) - c.fk_lo;
    int64_t dead = 0;
    if(R.nev > 0) { // events of this list strictly before q
        int64_t const hi = upper_index(R.ev_key, c.ev_lo, R.nev, q - 1);
        dead = (hi > 0 ? R.ev_cum[hi - 1] : 0) - c.ev_base;
    }
    int64_t const size1 = started - dead;
    return R.tol / double(size1 > 1 ? size1 : 1);
}

// Erase and restart events settled by fixed-point passes of a group reduce: a pass takes the
// previous pass's events (none in the first) into the thresholds, and the iteration stops when a
// pass reproduces them.  Every decision depends only on events at earlier positions, so by
// induction over the visiting order the settled pass is the sequential result.
// pass(rule, emit, evn, key_out, w_out) launches one reduce over all groups: emit == false writes
// each group's event count to evn (and the group results), emit == true writes the events at the
// offsets evn (key = list | position, weight +1 erase, -1 restart as uint64).
struct EraseEvents {
    WsBuf<uint64_t> key, w;
    WsBuf<int64_t> cum;
    int64_t n = 0;
};
using ErasePass = std::function<int(EraseRule const&, bool emit, int64_t* evn, uint64_t* key_out, uint64_t* w_out)>;
// 'rule' brings tol and fk; on return ev holds the settled events and the last pass's group
// results are in place.
int erase_fixed_point(pomx_ctx* ctx, int64_t ng, EraseRule rule, ErasePass const& pass, EraseEvents& ev,
                      const char* who);

// Device address of the no-allocation entry points' scratch array on ctx->device (pomx_pair.hip),
// looked up once per device; pomx_ctx_create calls it with the device current.
int pair_scratch_setup(pomx_ctx* ctx);

// Kernel statistics hooks (HIP events on the context stream, when timing is on): stat_begin
// returns a handle (-1 when off) that stat_end closes with the launch group's work model.
int stat_begin(pomx_ctx* ctx, int cls);
void stat_end(pomx_ctx* ctx, int handle, double flop, double term_freqs, double bytes);

constexpr int kBlock = 256;
inline unsigned grid_for(int64_t n, int per_block = kBlock) {
    int64_t g = (n + per_block - 1) / per_block;
    if(g < 1)
        g = 1;
    if(g > (int64_t(1) << 30))
        g = int64_t(1) << 30;
    return unsigned(g);
}

// ---- shared device routines (used by the plain and the batched paths) ---------------------
struct EighJob {
    int n;
    int pad;
    int64_t mat_off, val_off, ws_off; // doubles: block matrix, eigenvalues, workspace (n*n + n)
};
size_t eigh_lds_bytes(int nmax);
int64_t eigh_ws_doubles(int n);
int eigh_max_n();
// Launch the per-block eigensolver on device-resident jobs, without host synchronisation.
// persistent_grid > 0: that many workgroups pull jobs from *counter (zeroed by the caller).
int launch_eigh(pomx_ctx* ctx, const EighJob* d_jobs, int njobs, int nmax, int block_threads, int persistent_grid,
                int* counter, double* d_mat, double* d_eval, double* d_ws, int* d_status);

struct RotTile {
    int job, i0, j0, pad;
};
// Rotation tile list of the jobs (L2-friendly walk dealt to the XCDs; padding tiles have job -1).
void rot_tiles(const pomx_rot_job* jobs, int njobs, std::vector<RotTile>& tiles);
// Compaction of the gather maps (no host synchronisation): for every job, the pairs
// (t, src[t]) with src[t] != 0 in ascending t to d_kmap[src_off ..], their number to d_kcnt[job].
int launch_rot_compact(pomx_ctx* ctx, const pomx_rot_job* d_jobs, int njobs, const int* d_src, int2* d_kmap, int* d_kcnt);
// Launch the rotation GEMM over a device-resident tile list and compacted gather maps (no host
// synchronisation).
int launch_rotate(pomx_ctx* ctx, const RotTile* d_tiles, int ntiles, const pomx_rot_job* d_jobs, const double* d_U,
                  const int2* d_kmap, const int* d_kcnt, double* d_out, double cutoff);

struct GemmTile {
    int job, i0, j0, pad;
};
// tt x tt output tiles of every job, in job order.
void gemm_tiles(const pomx_gemm_job* jobs, int njobs, std::vector<GemmTile>& tiles, int tt = 64);
bool gemm_job_ok(pomx_gemm_job const& j);
// Launch the batched MFMA GEMM over device-resident tiles whose jobs share (trans_a, trans_b).
int launch_gemm(pomx_ctx* ctx, int trans_a, int trans_b, const GemmTile* d_tiles, int ntiles,
                const pomx_gemm_job* d_jobs, const double* d_A, const double* d_B, double* d_C);
// Launch one GEMM job (tile from the block index, no device descriptors).
int launch_gemm1(pomx_ctx* ctx, pomx_gemm_job const& J, const double* d_A, const double* d_B, double* d_C);

// Batched GEMM, job blockIdx.y, tiles along x (max_tiles = largest job's 64 x 64 tile count);
// job sizes may be written on the device before the launch.
// upper: only output tiles touching the upper triangle (for symmetric updates of square outputs
// that start on the diagonal; TN only).  big: 128 x 128 output tiles of eight waves (NN and TN,
// not upper), for jobs large enough to fill the device with them.
int launch_gemm_jobs(pomx_ctx* ctx, int trans_a, int trans_b, const pomx_gemm_job* d_jobs, int njobs, int64_t max_tiles,
                     const double* d_A, const double* d_B, double* d_C, bool upper = false, bool big = false);

// Device-driven symmetric eigensolver for a fixed set of blocks, all solved together: blocked
// Householder tridiagonalisation with MFMA trailing updates (blocks of at least 2048 rows: a
// two-stage reduction, dense to band by MFMA panel products, then a bulge chase to tridiagonal
// form), divide and conquer on the tridiagonal matrices (deflation on the device), blocked (WY)
// MFMA back-transformation.  The launch sequence depends only on the block sizes, so it can be
// captured in a hipGraph.
struct BigPlan;
// Device scratch of one block of size n in a plan (bytes, about).
size_t big_block_bytes(int n);
// Blocks: size n[b], row-major matrix at d_mat + mat_off[b] (overwritten by the eigenvectors as
// columns), eigenvalues ascending to d_eval + val_off[b].
int big_plan_create(pomx_ctx* ctx, int nblk, const int* n, const int64_t* mat_off, const int64_t* val_off, BigPlan** plan);
// Enqueue the whole solve on the context stream (no host synchronisation).  Non-convergence of a
// leaf eigensolve is flagged in the plan's status (big_plan_status).
int big_plan_enqueue(pomx_ctx* ctx, BigPlan* plan, double* d_mat, double* d_eval);
int big_plan_status(pomx_ctx* ctx, BigPlan* plan); // after the stream is synchronised
size_t big_plan_bytes(BigPlan const* plan);
void big_plan_destroy(BigPlan* plan);
// Blocks of at least this size go to the device-driven large-block solver in pomx_eigh_batched.
constexpr int kEighLargeMin = 129;
