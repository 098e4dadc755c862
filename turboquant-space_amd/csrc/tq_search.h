// tq_search.h -- persistent top-k search over a queue of work items.
#pragma once
#include "../../include/tq_hip.h"
#include "tq_common.h"

namespace tq {

constexpr int kSearchMaxK = TQ_FUSED_MAX_K;  // fused form: per-lane list capacity, k <= this
constexpr int kTopkMaxK = TQ_SEARCH_MAX_K;  // row top-k (blocked and fp32 plans): k <= this
constexpr int kSearchNTile = 256;   // codes per tile

// One (collection, query batch) pair: prepared queries, their fp16 planes
// (launch_split_queries), the codes, and where the top-k goes.
struct SearchJobDev {
  SpaceDev s;
  const _Float16 *planes;  // split query planes (tq_mfma_dev.h)
  const float *prep;       // [m][prep_stride(D)]
  const uint8_t *codes;    // [n][cs]
  long m, n;
  float *out_d;            // [m][k]
  int64_t *out_i;          // [m][k]; -1 where fewer than k codes exist
  // [m] the fused form's running bound on each query's k-th best distance
  // (+inf before the launch, lowered by atomicMin on the bits of
  // non-negative floats), for the coarse form's pruning
  float *thr;
};

// Code tiles [t0, t1) of query tile `qtile` (of 32 * qr queries) of job `job`.
struct SearchItem {
  int job, qtile, t0, t1;
};

// The items [ibeg, iend) that cover query tile `qtile` (of 32 * qr queries)
// of job `job`; all items of a group share its qr.
struct SearchGroup {
  int job, qtile, ibeg, iend, qr;
};

int search_max_k();
// p[0 .. n) = v (a kernel, so graph captures hold no memset node).
hipError_t launch_fill_f32(float *p, long n, float v, hipStream_t st);
// Per query-tile kind (qr = 1: 32 queries, qr = 2: 64 queries): resets that
// kind's queue counter (d_counter[qr - 1]) and runs a persistent search over
// grid_qr[qr - 1] workgroups on the nitems_qr[qr - 1] items listed next in
// `order` (all qr = 1 items first); then merges each query's candidate lists
// into its top-k. Candidate buffers hold nitems * 4 * 64 * kSearchMaxK
// entries, indexed by item.
hipError_t launch_search(const SearchJobDev *d_jobs, const SearchItem *d_items,
                         const int *d_order, const int *nitems_qr, const SearchGroup *d_groups,
                         int ngroups, int k, int *d_counter, float *d_cand_d,
                         uint32_t *d_cand_i, const int *grid_qr, hipStream_t st);
// Row-wise top-k of a dense [m][n] distance matrix (ties by column index),
// k <= kTopkMaxK, m <= 65535, n < 2^32 - 1. Scratch: seg_d / seg_i of
// m * topk_rows_segments(m, n) * k entries each.
long topk_rows_segments(long m, long n);
hipError_t launch_topk_rows(const float *dist, long m, long n, int k, float *out_d,
                            int64_t *out_i, float *seg_d, uint32_t *seg_i, hipStream_t st);
// The same over one chunk of columns, kept as partial lists of k entries:
// the topk_rows_segments(m, n) sorted lists of each row of the [m][n] block
// go to lists seg0 .. of that row's `stride` lists, with ids = column + id0;
// a list keeps its segment's best k (the rest empty). With use_run
// (seg0 >= 1), list 0 of each row holds that row's best k over earlier
// chunks (smaller ids), and the chunk's lists keep only entries at or below
// its k-th entry. launch_topk_rows_fold then merges lists [l0, l1) of every
// row either into list 0 (out_d == nullptr) or into the final k per row at
// out_d / out_i (ids ~0 -> -1).
hipError_t launch_topk_rows_chunk(const float *dist, long m, long n, long stride, long seg0,
                                  uint32_t id0, int k, bool use_run, float *seg_d,
                                  uint32_t *seg_i, hipStream_t st);
hipError_t launch_topk_rows_fold(long m, long stride, long l0, long l1, int k, float *seg_d,
                                 uint32_t *seg_i, float *out_d, int64_t *out_i, hipStream_t st);

}  // namespace tq
