// Internal state of a batched solve (pomx_batch_*), shared by the forward graph and the adjoint.
#pragma once

#include "pomx_internal.hpp"

namespace pomx_batch_detail {

constexpr int kBuckets = 4;
constexpr int kBucketMax[kBuckets] = {32, 128, 512, 1 << 30};
constexpr int kGThreads = 256;

struct GItem {
    int job, inst;
    int64_t t0, t1; // term range, term t = i_out * n_in + i_in
};

struct JobInfo {
    int n_out, n_in, inst, pad;
    int64_t ops_off, out_state, in_state;
};

} // namespace pomx_batch_detail

using pomx_batch_detail::GItem;
using pomx_batch_detail::JobInfo;
using pomx_batch_detail::kBucketMax;
using pomx_batch_detail::kBuckets;
using pomx_batch_detail::kGThreads;

struct AdjointState;

struct pomx_batch {
    int ninst = 0, nfreq = 0, nblocks = 0, njobs = 0, ntiles = 0, nitems = 0;
    int64_t mat_len = 0, nstates = 0, nnz = 0, ws_len = 0;
    double tol = 0;
    // device data
    double *mat = nullptr, *eval = nullptr, *w = nullptr, *ws = nullptr, *h_val = nullptr, *beta = nullptr;
    double *ops = nullptr, *partial = nullptr, *G = nullptr, *E0 = nullptr;
    int64_t *h_pos = nullptr, *inst_off = nullptr, *inst_cnt = nullptr, *item_off = nullptr;
    int *src = nullptr, *counters = nullptr, *status = nullptr;
    int2* kmap = nullptr; // gather maps compacted to their mapped rows (built at create)
    int* kcnt = nullptr;
    EighJob* eigh_jobs = nullptr;
    pomx_rot_job* jobs = nullptr;
    RotTile* tiles = nullptr;
    JobInfo* jinfo = nullptr;
    GItem* items = nullptr;
    // host copies needed to rebuild work lists
    std::vector<JobInfo> h_jinfo;
    std::vector<int64_t> inst_first_job; // jobs sorted by instance: first job of each instance
    std::vector<EighJob> h_eigh;          // all blocks (size-sorted, bucketed)
    std::vector<pomx_rot_job> h_jobs;
    std::vector<int> h_src;
    std::vector<int64_t> h_pos_host;      // Hamiltonian positions (matrix arena)
    AdjointState* adj = nullptr;          // adjoint buffers, allocated on first use
    BigPlan* big = nullptr;               // blocks of size >= kEighLargeMin (device-driven solver)
    int nbig = 0;
    int bucket_begin[kBuckets + 1] = {0, 0, 0, 0, 0}; // ranges of eigh_jobs (sorted by size, descending)
    int bucket_nmax[kBuckets] = {1, 1, 1, 1};
    // tuning
    int eigh_threads[kBuckets] = {64, 128, 256, 256};
    int eigh_grid = 0;
    int64_t terms_per_item = 4096;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};


// Release the adjoint buffers of a batch (no-op when the adjoint was never used).
void adjoint_free(pomx_batch* b);
