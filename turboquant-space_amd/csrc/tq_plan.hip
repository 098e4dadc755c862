// tq_plan.hip -- host side of the top-k search (include/tq_hip.h): search
// plans, their device workspace, the work queue, HIP graph capture and the
// autotuner table.
//
// A plan coalesces its jobs by collection (same space, codes and n): the
// queries of all jobs aimed at one collection are gathered into one batch,
// prepared and split once, and scored in one sweep over the collection's
// codes. Results land in plan-owned [rows, k] buffers and are scattered back
// to each job's outputs. Every step is a kernel launch on the caller's
// stream, so a plan replays as a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "tq_internal.h"
#include "tq_mfma.h"
#include "tq_search.h"

namespace tq {
namespace {

using capi::fail;
using capi::hip_fail;

// Contiguous float copies (query gather) and top-k result copies (scatter).
struct CopySeg {
  const float *src;
  float *dst;
  long count;  // floats
};
struct OutSeg {
  const float *src_d;
  const int64_t *src_i;
  float *dst_d;
  int64_t *dst_i;
  long count;  // entries
};

__global__ void k_gather(const CopySeg *__restrict__ segs) {
  const CopySeg sg = segs[blockIdx.y];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < sg.count;
       i += (long)gridDim.x * blockDim.x)
    sg.dst[i] = sg.src[i];
}

__global__ void k_scatter(const OutSeg *__restrict__ segs) {
  const OutSeg sg = segs[blockIdx.y];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < sg.count;
       i += (long)gridDim.x * blockDim.x) {
    sg.dst_d[i] = sg.src_d[i];
    sg.dst_i[i] = sg.src_i[i];
  }
}

// ---- autotuner table: (padded dim, bits, m bucket) -> knobs --------------

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, tq_search_tuning> g_tune;

int m_bucket(long m) {
  int b = 0;
  while ((1L << b) < m) ++b;
  return b;  // ceil(log2 m): 1 -> 0, 2 -> 1, 3..4 -> 2, ...
}

bool tune_lookup(int D, int bits, long m, tq_search_tuning *out) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find({D, bits, m_bucket(m)});
  if (it == g_tune.end()) return false;
  *out = it->second;
  return true;
}

void tune_store(int D, int bits, long m, const tq_search_tuning &t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[{D, bits, m_bucket(m)}] = t;
}

constexpr int kDefaultItemsPerWg = 4;
constexpr long kMinItemTiles = 8;  // fused form: code tiles per work item, at least
constexpr size_t kFp32BlockBytes = 512ull << 20;  // distance rows per block (FP32 plans)
// MFMA plans: by default a batch runs blocked: k_mfma writes the distances
// of all its queries against a chunk of codes into a block of at most
// kMfmaBlockBytes (or 4,096 columns), a row top-k keeps each chunk's
// candidates, and the candidates fold into one list per query. The fused
// persistent search (k_search) keeps its top-k lists in registers, which
// limits its query tiles to 32 rows at two wavefronts per SIMD (or 64 rows at
// one); the blocked form was faster at every shape measured (one job, dim
// 768 4-bit and dim 128 8-bit, n = 50,000 .. 2,000,000, m = 1 .. 1,024:
// 0.24-0.93x the fused time; 10M codes: 1.3-3x). The fused form stays for
// batches the blocked one cannot take (more than 65535 queries, no codes),
// and for callers and the autotuner to choose.
constexpr size_t kMfmaBlockBytes = 256ull << 20;

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace
}  // namespace tq

// One collection and the jobs aimed at it.
struct PlanBatch {
  tq_space space;
  const uint8_t *codes;
  long n;
  long m = 0;                 // rows over all its jobs
  long row0 = 0;              // first row in the plan-owned outputs
  std::vector<int> jobs;      // indices into plan->jobs, in row order
  std::vector<long> job_row;  // first row of each job within the batch
  int qr = 1;                 // fused query tile: 32 * qr
  int bqt = 0;                // blocked: k_mfma query tile (0: by m)
  size_t block_bytes = 0;     // blocked: distance block budget
  bool blocked = false;       // MFMA plans: k_mfma blocks + row top-k instead of k_search
  // The plain fp32 scan in row blocks + row top-k: every batch of an FP32
  // plan, and batches of padded dim below 32 (under the MFMA tile's
  // 32-coordinate chunk) in MFMA plans.
  bool plain = false;
  long nc = 0;                // blocked: codes per chunk
  long lists = 0;             // blocked: candidate lists per query (1 + segments of a chunk)
  // device pointers inside the workspace
  float *queries = nullptr, *prep = nullptr;
  _Float16 *planes = nullptr;
};

struct tq_search_plan_s {
  int device = 0;
  int k = 0;
  int mode = TQ_MODE_MFMA;
  int cus = 256;
  tq_search_tuning forced{};  // knobs the caller fixed (0 = automatic)
  tq_search_tuning used{};    // knobs of the current build
  std::vector<tq_search_job> jobs;
  std::vector<PlanBatch> batches;
  long rows = 0;

  // host tables (uploaded at build)
  std::vector<tq::SearchJobDev> h_jobs;
  std::vector<tq::SearchItem> h_items;
  std::vector<int> h_order;
  std::vector<tq::SearchGroup> h_groups;
  std::vector<tq::CopySeg> h_gather;
  std::vector<tq::OutSeg> h_scatter;
  int nitems_qr[2] = {0, 0};
  int grid_qr[2] = {0, 0};
  long gather_max = 0, scatter_max = 0;

  // workspace layout (byte offsets) of the current build
  size_t o_jobs = 0, o_items = 0, o_order = 0, o_groups = 0, o_gather = 0, o_scatter = 0,
         o_counter = 0, o_cand_d = 0, o_cand_i = 0, o_all_d = 0, o_all_i = 0, o_dist = 0,
         o_seg_d = 0, o_seg_i = 0, o_scan_ws = 0, o_thr = 0, scan_ws_bytes = 0, need = 0;
  std::vector<size_t> o_q, o_prep, o_planes;

  // device workspace: owned (tq_search_plan_create) or lent by the caller
  void *ws = nullptr;
  size_t ws_bytes = 0;
  bool owned = false;
  tq::SearchJobDev *d_jobs = nullptr;
  tq::SearchItem *d_items = nullptr;
  int *d_order = nullptr;
  tq::SearchGroup *d_groups = nullptr;
  tq::CopySeg *d_gather = nullptr;
  tq::OutSeg *d_scatter = nullptr;
  int *d_counter = nullptr;
  float *cand_d = nullptr;
  uint32_t *cand_i = nullptr;
  float *thr = nullptr;  // fused form: running k-th best bound per query row
  float *all_d = nullptr;
  int64_t *all_i = nullptr;
  // FP32 plans
  float *dist = nullptr;
  float *seg_d = nullptr;
  uint32_t *seg_i = nullptr;
  void *scan_ws = nullptr;

  // graph
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  void drop_ws() {
    if (ws && owned) (void)hipFree(ws);
    ws = nullptr;
    ws_bytes = 0;
    owned = false;
  }
  ~tq_search_plan_s() {
    drop_graph();
    drop_ws();
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {

using tq::capi::fail;
using tq::capi::hip_fail;

long fp32_block_rows(const PlanBatch &b) {
  long r = (long)(tq::kFp32BlockBytes / (sizeof(float) * (size_t)std::max(b.n, 1L)));
  r = std::max(1L, std::min(r, b.m));
  return std::min(r, 65535L);
}

// Knobs for one batch: forced > tuned table > built-in rule.
tq_search_tuning knobs_for(const tq_search_plan_s *p, const PlanBatch &b) {
  tq_search_tuning t{};
  const tq_space_info &info = b.space->info;
  tq_search_tuning tab{};
  const bool have = tq::tune_lookup((int)info.padded_dim, info.bits_per_coord, b.m, &tab);
  t.query_tile = p->forced.query_tile ? p->forced.query_tile
                 : have               ? tab.query_tile
                                      : (b.m <= 32 ? 32 : 64);
  t.items_per_wg = p->forced.items_per_wg ? p->forced.items_per_wg
                   : have                 ? tab.items_per_wg
                                          : tq::kDefaultItemsPerWg;
  t.form = p->forced.form ? p->forced.form : have && tab.form ? tab.form : TQ_FORM_BLOCKED;
  t.wg_per_cu = p->forced.wg_per_cu ? p->forced.wg_per_cu
                : have && tab.wg_per_cu ? tab.wg_per_cu
                                        : 2;
  t.blocked_query_tile = p->forced.blocked_query_tile ? p->forced.blocked_query_tile
                         : have                       ? tab.blocked_query_tile
                                                      : 0;
  t.block_mb = p->forced.block_mb ? p->forced.block_mb
               : have && tab.block_mb ? tab.block_mb
                                      : (int)(tq::kMfmaBlockBytes >> 20);
  return t;
}

// Work queue, groups and the workspace layout for the current knobs (host
// only; bind() places it in device memory).
int layout(tq_search_plan_s *p) {
  const char *what = "tq_search_plan";
  p->drop_graph();
  p->h_jobs.clear();
  p->h_items.clear();
  p->h_order.clear();
  p->h_groups.clear();
  p->h_gather.clear();
  p->h_scatter.clear();
  p->nitems_qr[0] = p->nitems_qr[1] = 0;
  p->grid_qr[0] = p->grid_qr[1] = 0;
  p->gather_max = p->scatter_max = 0;

  // Plan-level items per workgroup and persistent workgroups per CU: the
  // knobs of the costliest batch.
  int ipw = tq::kDefaultItemsPerWg, wpc = 2;
  double worst = -1.0;
  for (PlanBatch &b : p->batches) {
    const tq_search_tuning t = knobs_for(p, b);
    b.qr = t.query_tile == 64 ? 2 : 1;
    b.bqt = t.blocked_query_tile;
    b.block_bytes = (size_t)t.block_mb << 20;
    b.plain = p->mode == TQ_MODE_FP32 || b.space->info.padded_dim < 32;
    // k above the fused form's list length always runs blocked (plan_new
    // refuses plans where that is not possible).
    b.blocked = p->mode == TQ_MODE_MFMA &&
                (t.form == TQ_FORM_BLOCKED || p->k > tq::kSearchMaxK) &&
                b.space->info.padded_dim >= 32 && tq::plan_blocked(b.n, b.m);
    const double cost = (double)b.m * (double)b.n * (double)b.space->info.padded_dim;
    if (cost > worst) {
      worst = cost;
      ipw = std::max(1, (int)t.items_per_wg);
      wpc = t.wg_per_cu == 1 ? 1 : 2;
    }
  }
  p->used.items_per_wg = ipw;
  p->used.wg_per_cu = wpc;
  p->used.query_tile = 0;

  // Items per query-tile kind, sized so each kind's queue holds about
  // ipw items per persistent workgroup; within a kind the largest go first.
  const int ntile = tq::kSearchNTile;
  struct Pending {
    tq::SearchItem it;
    double cost;
    int kind;
  };
  std::vector<Pending> pend;
  if (p->mode == TQ_MODE_MFMA) {
    for (int kind = 1; kind <= 2; ++kind) {
      double total = 0.0;
      for (const PlanBatch &b : p->batches) {
        if (b.qr != kind || b.m == 0 || b.blocked || b.plain) continue;
        const long qtiles = (b.m + 32 * kind - 1) / (32 * kind);
        const long ntiles = (b.n + ntile - 1) / ntile;
        total += (double)qtiles * ntiles * b.space->info.padded_dim * kind;
      }
      const long wg = (long)p->cus * wpc;  // persistent workgroups for both query tiles
      const double target = total / (double)(wg * ipw);
      for (size_t bi = 0; bi < p->batches.size(); ++bi) {
        const PlanBatch &b = p->batches[bi];
        if (b.qr != kind || b.m == 0 || b.blocked || b.plain) continue;
        const long qtiles = (b.m + 32 * kind - 1) / (32 * kind);
        const long ntiles = (b.n + ntile - 1) / ntile;
        const double tile_cost = (double)b.space->info.padded_dim * kind;
        long tpi = (long)std::llround(target / tile_cost);
        // At least kMinItemTiles code tiles per item: every item writes 256
        // candidate lists and the merge reads them per query, so one-tile
        // items (small batches over many workgroups) cost more than they
        // balance.
        tpi = std::max(tpi, tq::kMinItemTiles);
        tpi = std::max(1L, std::min(tpi, std::max(ntiles, 1L)));
        for (long qt = 0; qt < qtiles; ++qt) {
          tq::SearchGroup g{(int)bi, (int)qt, 0, 0, kind};
          g.ibeg = (int)pend.size();
          for (long t0 = 0; t0 < ntiles; t0 += tpi) {
            const long t1 = std::min(ntiles, t0 + tpi);
            pend.push_back({{(int)bi, (int)qt, (int)t0, (int)t1}, (double)(t1 - t0) * tile_cost,
                            kind});
          }
          g.iend = (int)pend.size();
          p->h_groups.push_back(g);
        }
      }
    }
    if (pend.size() > (size_t)INT32_MAX / 4)
      return fail(TQ_EINVAL, "%s: too many work items (%zu)", what, pend.size());
    p->h_items.reserve(pend.size());
    for (const Pending &pi : pend) p->h_items.push_back(pi.it);
    std::vector<int> ord(pend.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
    // Queue order: per kind, the costliest items first; among items of equal
    // cost, by code range and then query tile, so the workgroups that run
    // side by side score the same codes for different query tiles and share
    // the code bytes in L2 (items of one query tile stay contiguous in the
    // item list, which the groups index; only the queue visits them apart).
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      const Pending &x = pend[a], &y = pend[b];
      if (x.kind != y.kind) return x.kind < y.kind;
      if (x.cost != y.cost) return x.cost > y.cost;
      if (x.it.job != y.it.job) return x.it.job < y.it.job;
      if (x.it.t0 != y.it.t0) return x.it.t0 < y.it.t0;
      return x.it.qtile < y.it.qtile;
    });
    p->h_order = ord;
    for (const Pending &pi : pend) p->nitems_qr[pi.kind - 1]++;
    for (int kind = 1; kind <= 2; ++kind) {
      const long wg = (long)p->cus * wpc;
      p->grid_qr[kind - 1] = (int)std::min<long>(wg, p->nitems_qr[kind - 1]);
    }
  }

  // Workspace layout.
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += tq::align_up(bytes ? bytes : 4);
    return at;
  };
  const size_t nb = p->batches.size();
  p->o_jobs = take(nb * sizeof(tq::SearchJobDev));
  p->o_items = take(p->h_items.size() * sizeof(tq::SearchItem));
  p->o_order = take(p->h_order.size() * sizeof(int));
  p->o_groups = take(p->h_groups.size() * sizeof(tq::SearchGroup));
  p->o_gather = take(p->jobs.size() * sizeof(tq::CopySeg));
  p->o_scatter = take(p->jobs.size() * sizeof(tq::OutSeg));
  p->o_counter = take(2 * sizeof(int));
  const size_t cand = p->h_items.size() * 4 * 64 * (size_t)tq::kSearchMaxK;
  p->o_cand_d = take(cand * sizeof(float));
  p->o_cand_i = take(cand * sizeof(uint32_t));
  p->o_thr = take((size_t)p->rows * sizeof(float));
  p->o_all_d = take((size_t)p->rows * p->k * sizeof(float));
  p->o_all_i = take((size_t)p->rows * p->k * sizeof(int64_t));
  p->o_q.assign(nb, 0);
  p->o_prep.assign(nb, 0);
  p->o_planes.assign(nb, 0);
  size_t dist_elems = 0, seg_elems = 0, scan_ws = 0;
  for (size_t bi = 0; bi < nb; ++bi) {
    PlanBatch &b = p->batches[bi];
    const tq_space_info &info = b.space->info;
    p->o_q[bi] = take((size_t)b.m * info.dim * sizeof(float));
    p->o_prep[bi] = take((size_t)b.m * info.prepared_stride * sizeof(float));
    if (!b.plain) {
      p->o_planes[bi] = take(tq::split_planes_bytes(b.space->dev, (long)b.m));
      if (b.blocked) {
        long nc = (long)(b.block_bytes / (sizeof(float) * (size_t)b.m));
        nc = std::max(4096L, nc - nc % tq::kSearchNTile);
        b.nc = std::min(nc, b.n);
        b.lists = 1 + tq::topk_rows_segments(b.m, b.nc);
        dist_elems = std::max(dist_elems, (size_t)b.m * b.nc);
        seg_elems = std::max(seg_elems, (size_t)b.m * b.lists * p->k);
      }
    } else if (b.m > 0) {
      const long rows = fp32_block_rows(b);
      dist_elems = std::max(dist_elems, (size_t)rows * b.n);
      // A block of r <= rows rows uses r * topk_rows_segments(r, n) lists.
      seg_elems = std::max(seg_elems,
                           (size_t)rows * tq::topk_rows_segments(rows, b.n) * p->k);
      scan_ws = std::max(scan_ws, tq::scan_fp32_ws_bytes(b.space->dev, rows));
    }
  }
  p->o_dist = take(dist_elems * sizeof(float));
  p->o_seg_d = take(seg_elems * sizeof(float));
  p->o_seg_i = take(seg_elems * sizeof(uint32_t));
  p->o_scan_ws = take(scan_ws);
  p->scan_ws_bytes = scan_ws;
  p->need = off;
  return TQ_OK;
}

// Places the layout at `ws` (at least p->need bytes), fills the device-side
// tables and uploads them: with hipMemcpy, or enqueued on `st` when `async`
// (the plan's host tables must then live until the copies have run).
int bind(tq_search_plan_s *p, void *ws, hipStream_t st, bool async) {
  const char *what = "tq_search_plan";
  p->drop_graph();
  char *w = static_cast<char *>(ws);
  p->d_jobs = reinterpret_cast<tq::SearchJobDev *>(w + p->o_jobs);
  p->d_items = reinterpret_cast<tq::SearchItem *>(w + p->o_items);
  p->d_order = reinterpret_cast<int *>(w + p->o_order);
  p->d_groups = reinterpret_cast<tq::SearchGroup *>(w + p->o_groups);
  p->d_gather = reinterpret_cast<tq::CopySeg *>(w + p->o_gather);
  p->d_scatter = reinterpret_cast<tq::OutSeg *>(w + p->o_scatter);
  p->d_counter = reinterpret_cast<int *>(w + p->o_counter);
  p->cand_d = reinterpret_cast<float *>(w + p->o_cand_d);
  p->cand_i = reinterpret_cast<uint32_t *>(w + p->o_cand_i);
  p->thr = reinterpret_cast<float *>(w + p->o_thr);
  p->all_d = reinterpret_cast<float *>(w + p->o_all_d);
  p->all_i = reinterpret_cast<int64_t *>(w + p->o_all_i);
  p->dist = reinterpret_cast<float *>(w + p->o_dist);
  p->seg_d = reinterpret_cast<float *>(w + p->o_seg_d);
  p->seg_i = reinterpret_cast<uint32_t *>(w + p->o_seg_i);
  p->scan_ws = p->scan_ws_bytes ? static_cast<void *>(w + p->o_scan_ws) : nullptr;

  p->h_jobs.clear();
  p->h_gather.clear();
  p->h_scatter.clear();
  p->gather_max = p->scatter_max = 0;
  const size_t nb = p->batches.size();
  for (size_t bi = 0; bi < nb; ++bi) {
    PlanBatch &b = p->batches[bi];
    b.queries = reinterpret_cast<float *>(w + p->o_q[bi]);
    b.prep = reinterpret_cast<float *>(w + p->o_prep[bi]);
    b.planes =
        !b.plain ? reinterpret_cast<_Float16 *>(w + p->o_planes[bi]) : nullptr;
    tq::SearchJobDev jd{};
    jd.s = b.space->dev;
    jd.planes = b.planes;
    jd.prep = b.prep;
    jd.codes = b.codes;
    jd.m = b.m;
    jd.n = b.n;
    jd.out_d = p->all_d + b.row0 * p->k;
    jd.out_i = p->all_i + b.row0 * p->k;
    jd.thr = p->thr + b.row0;
    p->h_jobs.push_back(jd);
    const long dim = (long)b.space->info.dim;
    for (size_t t = 0; t < b.jobs.size(); ++t) {
      const tq_search_job &j = p->jobs[b.jobs[t]];
      if (j.m == 0) continue;
      const long r0 = b.row0 + b.job_row[t];
      p->h_gather.push_back({j.queries, b.queries + b.job_row[t] * dim, (long)j.m * dim});
      p->h_scatter.push_back({p->all_d + r0 * p->k, p->all_i + r0 * p->k, j.out_dist, j.out_ids,
                              (long)j.m * p->k});
      p->gather_max = std::max(p->gather_max, (long)j.m * dim);
      p->scatter_max = std::max(p->scatter_max, (long)j.m * p->k);
    }
  }
  hipError_t e = hipSuccess;
  auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
    if (!bytes) return hipSuccess;
    return async ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st)
                 : hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  if ((e = up(p->d_jobs, p->h_jobs.data(), nb * sizeof(tq::SearchJobDev))) != hipSuccess ||
      (e = up(p->d_items, p->h_items.data(), p->h_items.size() * sizeof(tq::SearchItem))) !=
          hipSuccess ||
      (e = up(p->d_order, p->h_order.data(), p->h_order.size() * sizeof(int))) != hipSuccess ||
      (e = up(p->d_groups, p->h_groups.data(), p->h_groups.size() * sizeof(tq::SearchGroup))) !=
          hipSuccess ||
      (e = up(p->d_gather, p->h_gather.data(), p->h_gather.size() * sizeof(tq::CopySeg))) !=
          hipSuccess ||
      (e = up(p->d_scatter, p->h_scatter.data(), p->h_scatter.size() * sizeof(tq::OutSeg))) !=
          hipSuccess)
    return hip_fail(e, what);
  return TQ_OK;
}

// layout() + bind() into memory the plan owns (grown when the layout needs
// more than it has).
int build(tq_search_plan_s *p) {
  const char *what = "tq_search_plan";
  if (int r = layout(p)) return r;
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return hip_fail(e, what);
  if (!p->owned || p->need > p->ws_bytes) {
    p->drop_graph();
    p->drop_ws();
    e = hipMalloc(&p->ws, p->need);
    if (e != hipSuccess) {
      p->ws = nullptr;
      return hip_fail(e, what);
    }
    p->ws_bytes = p->need;
    p->owned = true;
  }
  return bind(p, p->ws, nullptr, false);
}

unsigned copy_blocks(long max_count) {
  long b = (max_count + 255) / 256;
  return (unsigned)std::max(1L, std::min(b, 1024L));
}

// Everything the plan does, enqueued on `st`.
int enqueue(tq_search_plan_s *p, hipStream_t st) {
  const char *what = "tq_search_plan_run";
  hipError_t e;
  if (!p->h_gather.empty()) {
    hipLaunchKernelGGL(tq::k_gather, dim3(copy_blocks(p->gather_max), (unsigned)p->h_gather.size()),
                       dim3(256), 0, st, p->d_gather);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, what);
  }
  for (const PlanBatch &b : p->batches) {
    if (b.m == 0) continue;
    if ((e = tq::launch_prepare(b.space->dev, b.queries, b.m, b.prep, st)) != hipSuccess)
      return hip_fail(e, what);
    if (!b.plain &&
        (e = tq::launch_split_queries(b.space->dev, b.prep, b.m, b.planes, st)) != hipSuccess)
      return hip_fail(e, what);
  }
  if (p->mode == TQ_MODE_MFMA) {
    // Fused batches (none when every batch runs blocked, as for k > 16).
    if (!p->h_items.empty() || !p->h_groups.empty()) {
      if ((e = tq::launch_fill_f32(p->thr, (long)p->rows, __builtin_inff(), st)) != hipSuccess)
        return hip_fail(e, what);
      e = tq::launch_search(p->d_jobs, p->d_items, p->d_order, p->nitems_qr, p->d_groups,
                            (int)p->h_groups.size(), p->k, p->d_counter, p->cand_d, p->cand_i,
                            p->grid_qr, st);
      if (e != hipSuccess) return hip_fail(e, what);
    }
    // Blocked batches: per chunk of codes, the distance block, its row top-k
    // as lists 1.. of each query, folded into list 0; the last fold writes
    // the query's k best. After the first chunk, list 0's k-th entries bound
    // each query's final k-th best: the coarse form's blocks correct only the
    // near pairs that can still beat them.
    for (const PlanBatch &b : p->batches) {
      if (!b.blocked) continue;
      const long cs = (long)b.space->info.code_size;
      for (long c0 = 0; c0 < b.n; c0 += b.nc) {
        const long nc = std::min(b.nc, b.n - c0);
        const long l1 = 1 + tq::topk_rows_segments(b.m, nc);
        const bool last = c0 + nc >= b.n;
        const float *thr = c0 > 0 ? p->seg_d + (p->k - 1) : nullptr;
        if ((e = tq::launch_mfma_planes(b.space->dev, b.planes, b.prep, b.m, b.codes + c0 * cs, nc,
                                        p->dist, st, true, b.bqt, thr,
                                        (long)b.lists * p->k)) != hipSuccess ||
            (e = tq::launch_topk_rows_chunk(p->dist, b.m, nc, b.lists, 1, (uint32_t)c0, p->k,
                                            c0 > 0, p->seg_d, p->seg_i, st)) != hipSuccess ||
            (e = tq::launch_topk_rows_fold(b.m, b.lists, c0 == 0 ? 1 : 0, l1, p->k, p->seg_d,
                                           p->seg_i, last ? p->all_d + b.row0 * p->k : nullptr,
                                           last ? p->all_i + b.row0 * p->k : nullptr, st)) !=
                hipSuccess)
          return hip_fail(e, what);
      }
    }
  }
  {
    for (const PlanBatch &b : p->batches) {
      if (b.m == 0 || !b.plain) continue;
      float *od = p->all_d + b.row0 * p->k;
      int64_t *oi = p->all_i + b.row0 * p->k;
      // With no codes the scan is skipped and the top-k writes empty entries.
      const long rows = fp32_block_rows(b);
      const long stride = (long)b.space->info.prepared_stride;
      for (long r0 = 0; r0 < b.m; r0 += rows) {
        const long rc = std::min(rows, b.m - r0);
        e = tq::launch_scan_fp32(b.space->dev, b.prep + r0 * stride, rc, b.codes, b.n, p->dist,
                                 false, st, p->scan_ws);
        if (e != hipSuccess) return hip_fail(e, what);
        e = tq::launch_topk_rows(p->dist, rc, b.n, p->k, od + r0 * p->k, oi + r0 * p->k,
                                 p->seg_d, p->seg_i, st);
        if (e != hipSuccess) return hip_fail(e, what);
      }
    }
  }
  if (!p->h_scatter.empty()) {
    hipLaunchKernelGGL(tq::k_scatter,
                       dim3(copy_blocks(p->scatter_max), (unsigned)p->h_scatter.size()), dim3(256),
                       0, st, p->d_scatter);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, what);
  }
  return TQ_OK;
}

}  // namespace

namespace tq {

// Checks the jobs and builds the plan's host side and layout; no device
// memory is touched. Shared by tq_search_plan_create and the server.
int plan_new(const tq_search_job *jobs, uint64_t njobs, int k, int mode,
             const tq_search_tuning *tuning, tq_search_plan *out) {
  const char *what = "tq_search_plan_create";
  if (!out) return fail(TQ_EINVAL, "%s: out is NULL", what);
  *out = nullptr;
  if (!jobs || njobs == 0) return fail(TQ_EINVAL, "%s: no jobs", what);
  if (njobs > 65535) return fail(TQ_EINVAL, "%s: at most 65535 jobs per plan", what);
  if (k < 1 || k > TQ_SEARCH_MAX_K)
    return fail(TQ_EINVAL, "%s: k must be in [1, %d], got %d", what, TQ_SEARCH_MAX_K, k);
  if (mode != TQ_MODE_MFMA && mode != TQ_MODE_FP32)
    return fail(TQ_EINVAL, "%s: unknown mode %d", what, mode);
  if (tuning) {
    if (tuning->query_tile != 0 && tuning->query_tile != 32 && tuning->query_tile != 64)
      return fail(TQ_EINVAL, "%s: query_tile must be 0, 32 or 64", what);
    if (tuning->items_per_wg < 0 || tuning->items_per_wg > 1024)
      return fail(TQ_EINVAL, "%s: items_per_wg must be in [0, 1024]", what);
    if (tuning->form != TQ_FORM_AUTO && tuning->form != TQ_FORM_FUSED &&
        tuning->form != TQ_FORM_BLOCKED)
      return fail(TQ_EINVAL, "%s: form must be 0, TQ_FORM_FUSED or TQ_FORM_BLOCKED", what);
    if (tuning->wg_per_cu < 0 || tuning->wg_per_cu > 2)
      return fail(TQ_EINVAL, "%s: wg_per_cu must be 0, 1 or 2", what);
    if (tuning->blocked_query_tile != 0 && tuning->blocked_query_tile != 32 &&
        tuning->blocked_query_tile != 64)
      return fail(TQ_EINVAL, "%s: blocked_query_tile must be 0, 32 or 64", what);
    if (tuning->block_mb != 0 && (tuning->block_mb < 16 || tuning->block_mb > 256))
      return fail(TQ_EINVAL, "%s: block_mb must be 0 or in [16, 256]", what);
  }
  int device = -1;
  for (uint64_t i = 0; i < njobs; ++i) {
    const tq_search_job &j = jobs[i];
    if (!j.space) return fail(TQ_EINVAL, "%s: job %llu has no space", what, (unsigned long long)i);
    if (device < 0) device = j.space->device;
    if (j.space->device != device)
      return fail(TQ_EINVAL, "%s: job %llu is on device %d, job 0 on %d", what,
                  (unsigned long long)i, j.space->device, device);
    if (j.n >= 0xFFFFFFFFull)
      return fail(TQ_EINVAL, "%s: job %llu: n must be below 2^32 - 1", what, (unsigned long long)i);
    if (j.m > (1ull << 31))
      return fail(TQ_EINVAL, "%s: job %llu: m must be at most 2^31", what, (unsigned long long)i);
    if (j.m > 0 && (!j.queries || !j.out_dist || !j.out_ids || (j.n > 0 && !j.codes)))
      return fail(TQ_EINVAL, "%s: job %llu has a NULL buffer", what, (unsigned long long)i);
  }
  // k above the fused form's per-lane lists runs blocked only.
  if (mode == TQ_MODE_MFMA && k > tq::kSearchMaxK) {
    if (tuning && tuning->form == TQ_FORM_FUSED)
      return fail(TQ_EINVAL, "%s: the fused form takes k <= %d, got %d", what, tq::kSearchMaxK, k);
    uint64_t rows = 0;
    for (uint64_t i = 0; i < njobs; ++i) rows += jobs[i].m;
    if (rows > 65535)
      return fail(TQ_EINVAL, "%s: k > %d runs blocked, which takes at most 65535 queries per plan",
                  what, tq::kSearchMaxK);
  }
  tq_search_plan_s *p = new (std::nothrow) tq_search_plan_s();
  if (!p) return fail(TQ_ENOMEM, "%s: host allocation failed", what);
  p->device = device;
  p->k = k;
  p->mode = mode;
  p->cus = tq::capi::cu_count(device);
  if (tuning) p->forced = *tuning;
  p->jobs.assign(jobs, jobs + njobs);
  // Coalesce by collection, in first-seen order.
  for (uint64_t i = 0; i < njobs; ++i) {
    const tq_search_job &j = jobs[i];
    PlanBatch *b = nullptr;
    for (PlanBatch &c : p->batches)
      if (c.space == j.space && c.codes == j.codes && c.n == (long)j.n) b = &c;
    if (!b) {
      p->batches.push_back(PlanBatch{});
      b = &p->batches.back();
      b->space = j.space;
      b->codes = j.codes;
      b->n = (long)j.n;
    }
    b->jobs.push_back((int)i);
    b->job_row.push_back(b->m);
    b->m += (long)j.m;
  }
  for (PlanBatch &b : p->batches) {
    b.row0 = p->rows;
    p->rows += b.m;
    if (mode == TQ_MODE_MFMA && b.m > (1L << 31) - 64)
      return delete p, fail(TQ_EINVAL, "%s: too many queries for one collection", what);
  }
  if (int r = layout(p)) return delete p, r;
  *out = p;
  return TQ_OK;
}

size_t plan_ws_bytes(tq_search_plan p) { return p->need; }

// Binds the plan to caller memory (>= plan_ws_bytes) and enqueues the table
// uploads on `st`; the plan must outlive that work.
int plan_bind(tq_search_plan p, void *ws, size_t bytes, hipStream_t st) {
  if (bytes < p->need)
    return fail(TQ_EINVAL, "tq_search_plan: workspace of %zu bytes, %zu needed", bytes, p->need);
  p->drop_ws();
  p->ws = ws;
  p->ws_bytes = bytes;
  p->owned = false;
  return bind(p, ws, st, true);
}

int plan_enqueue(tq_search_plan p, hipStream_t st) { return enqueue(p, st); }

bool plan_blocked(long n, long rows) {
  return rows > 0 && rows <= 65535 && n > 0 && n < 0xFFFFFFFFL;
}

size_t plan_blocked_ws_bound(long rows, long n_max, int k) {
  // The distance block: rows x nc floats, nc <= min(n, max(4096,
  // kMfmaBlockBytes / (4 * rows))); and (1 + segments) lists per query over a
  // chunk, segments = topk_rows_segments(rows, nc) = ceil(nc / 8192):
  // rows * lists <= 2 * rows + rows * nc / 8192, which the conservative
  // 4 * rows + kMfmaBlockBytes / 8192 covers (rows * nc * 4 bytes is at most
  // max(kMfmaBlockBytes, 4096 * 4 * rows)).
  const size_t r = (size_t)std::max(rows, 1L);
  const size_t block = std::min(std::max(kMfmaBlockBytes, r * 4096 * sizeof(float)),
                                r * (size_t)std::max(n_max, 1L) * sizeof(float));
  const size_t lists = 4 * r + kMfmaBlockBytes / 8192;
  return align_up(block) + 2 * align_up(lists * (size_t)std::max(k, 1) * sizeof(float)) + 4096;
}

}  // namespace tq

extern "C" {

int tq_search_plan_create(const tq_search_job *jobs, uint64_t njobs, int k, int mode,
                          const tq_search_tuning *tuning, tq_search_plan *out) {
  const char *what = "tq_search_plan_create";
  if (!out) return fail(TQ_EINVAL, "%s: out is NULL", what);
  *out = nullptr;
  tq_search_plan p = nullptr;
  if (int r = tq::plan_new(jobs, njobs, k, mode, tuning, &p)) return r;
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return delete p, hip_fail(e, what);
  if (int r = build(p)) return delete p, r;
  *out = p;
  return TQ_OK;
}

int tq_search_plan_run(tq_search_plan p, void *stream) {
  if (!p) return fail(TQ_EINVAL, "tq_search_plan_run: NULL plan");
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return hip_fail(e, "tq_search_plan_run");
  return enqueue(p, static_cast<hipStream_t>(stream));
}

int tq_search_plan_run_graph(tq_search_plan p, void *stream) {
  const char *what = "tq_search_plan_run_graph";
  if (!p) return fail(TQ_EINVAL, "%s: NULL plan", what);
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return hip_fail(e, what);
  if (!p->exec) {
    if (!p->cap && (e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking)) != hipSuccess)
      return hip_fail(e, what);
    if ((e = hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hip_fail(e, what);
    const int r = enqueue(p, p->cap);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(p->cap, &g);
    if (r) {
      if (g) (void)hipGraphDestroy(g);
      return r;
    }
    if (e != hipSuccess) return hip_fail(e, what);
    p->graph = g;
    if ((e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0)) != hipSuccess) {
      p->drop_graph();
      return hip_fail(e, what);
    }
  }
  e = hipGraphLaunch(p->exec, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, what);
  return TQ_OK;
}

int tq_search_tuning_lookup(int32_t padded_dim, int32_t bits, uint64_t m, tq_search_tuning *out) {
  if (!out) return fail(TQ_EINVAL, "tq_search_tuning_lookup: out is NULL");
  if (!tq::tune_lookup(padded_dim, bits, (long)m, out))
    return fail(TQ_EINVAL, "tq_search_tuning_lookup: (%d, %d, m = %llu) has not been tuned",
                padded_dim, bits, (unsigned long long)m);
  return TQ_OK;
}

int tq_search_plan_tune(tq_search_plan p, void *stream, tq_search_tuning *chosen) {
  const char *what = "tq_search_plan_tune";
  if (!p) return fail(TQ_EINVAL, "%s: NULL plan", what);
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return hip_fail(e, what);
  if (p->mode == TQ_MODE_MFMA) {
    hipEvent_t t0, t1;
    if ((e = hipEventCreate(&t0)) != hipSuccess) return hip_fail(e, what);
    if ((e = hipEventCreate(&t1)) != hipSuccess) {
      (void)hipEventDestroy(t0);
      return hip_fail(e, what);
    }
    const tq_search_tuning saved = p->forced;
    int rc = TQ_OK;
    // Each collection batch alone: a sub-plan over that batch's jobs.
    for (size_t bi = 0; bi < p->batches.size() && rc == TQ_OK; ++bi) {
      const PlanBatch &b = p->batches[bi];
      if (b.m == 0 || b.n == 0 || b.plain) continue;  // plain batches have no knobs
      std::vector<tq_search_job> sub;
      for (int ji : b.jobs) sub.push_back(p->jobs[ji]);
      // Candidates: the fused form's knobs (query tile, items per
      // workgroup, persistent workgroups per CU), then the blocked form's
      // (MFMA query tile, distance block budget) when the batch allows it.
      // The table keeps the fastest knobs of each form and the fastest form.
      tq_search_tuning best{}, best_blk{};
      float best_ms = 0.0f, best_blk_ms = 0.0f, best_form_ms = 0.0f;
      int best_form = TQ_FORM_FUSED;
      std::vector<tq_search_tuning> cands;
      if (p->k <= tq::kSearchMaxK)
        for (int qt : {32, 64})
          for (int ipw : {1, 2, 4, 8})
            for (int wpc : {1, 2})
              cands.push_back(tq_search_tuning{qt, ipw, TQ_FORM_FUSED, wpc, 0, 0});
      if (b.space->info.padded_dim >= 32 && tq::plan_blocked(b.n, b.m))
        for (int bqt : {32, 64}) {
          if (bqt == 64 && b.m <= 32) continue;  // one 32-row tile covers the batch
          for (int mb : {64, 128, 256})
            cands.push_back(tq_search_tuning{0, 0, TQ_FORM_BLOCKED, 0, bqt, mb});
        }
      for (size_t ci = 0; ci < cands.size(); ++ci) {
        const tq_search_tuning &cand = cands[ci];
        tq_search_plan sp = nullptr;
        rc = tq_search_plan_create(sub.data(), sub.size(), p->k, p->mode, &cand, &sp);
        if (rc) break;
        float ms = 0.0f, run_best = 0.0f;
        for (int rep = 0; rep < 4 && rc == TQ_OK; ++rep) {
          if ((e = hipEventRecord(t0, st)) != hipSuccess) rc = hip_fail(e, what);
          if (!rc) rc = enqueue(sp, st);
          if (!rc && (e = hipEventRecord(t1, st)) != hipSuccess) rc = hip_fail(e, what);
          if (!rc && (e = hipEventSynchronize(t1)) != hipSuccess) rc = hip_fail(e, what);
          if (!rc && (e = hipEventElapsedTime(&ms, t0, t1)) != hipSuccess) rc = hip_fail(e, what);
          if (!rc && rep > 0 && (rep == 1 || ms < run_best)) run_best = ms;  // rep 0 warms up
        }
        tq_search_plan_destroy(sp);
        if (rc) break;
        if (cand.form == TQ_FORM_FUSED && (best.query_tile == 0 || run_best < best_ms)) {
          best = cand;
          best_ms = run_best;
        }
        if (cand.form == TQ_FORM_BLOCKED &&
            (best_blk.blocked_query_tile == 0 || run_best < best_blk_ms)) {
          best_blk = cand;
          best_blk_ms = run_best;
        }
        if (ci == 0 || run_best < best_form_ms) {
          best_form = cand.form;
          best_form_ms = run_best;
        }
      }
      if (rc == TQ_OK) {
        best.form = best_form;
        best.blocked_query_tile = best_blk.blocked_query_tile;
        best.block_mb = best_blk.block_mb;
        tq::tune_store((int)b.space->info.padded_dim, b.space->info.bits_per_coord, b.m, best);
      }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return rc;
    p->forced = saved;
    if (int r = build(p)) return r;
  }
  if (chosen) {
    *chosen = p->used;
    // The knobs of the costliest batch.
    double worst = -1.0;
    for (const PlanBatch &b : p->batches) {
      const double cost = (double)b.m * (double)b.n * (double)b.space->info.padded_dim;
      if (cost > worst) {
        worst = cost;
        chosen->query_tile = 32 * b.qr;
        chosen->form = b.blocked ? TQ_FORM_BLOCKED : TQ_FORM_FUSED;
        chosen->blocked_query_tile = b.bqt ? b.bqt : (b.m <= 32 ? 32 : 64);
        chosen->block_mb = (int)(b.block_bytes >> 20);
      }
    }
  }
  return TQ_OK;
}

int tq_search_plan_stats(tq_search_plan p, uint64_t *items, uint64_t *workgroups) {
  if (!p) return fail(TQ_EINVAL, "tq_search_plan_stats: NULL plan");
  if (items) *items = p->h_items.size();
  if (workgroups) *workgroups = (uint64_t)p->grid_qr[0] + (uint64_t)p->grid_qr[1];
  return TQ_OK;
}

int tq_search_plan_destroy(tq_search_plan p) {
  delete p;
  return TQ_OK;
}

int tq_search(const tq_search_job *jobs, uint64_t njobs, int k, int mode, void *stream) {
  tq_search_plan p = nullptr;
  if (int r = tq_search_plan_create(jobs, njobs, k, mode, nullptr, &p)) return r;
  int r = tq_search_plan_run(p, stream);
  if (r == TQ_OK) {
    hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) r = hip_fail(e, "tq_search");
  }
  tq_search_plan_destroy(p);
  return r;
}

int tq_topk_rows(const float *dist, uint64_t m, uint64_t n, int k, float *out_dist,
                 int64_t *out_ids, void *stream) {
  const char *what = "tq_topk_rows";
  if (k < 1 || k > TQ_SEARCH_MAX_K)
    return fail(TQ_EINVAL, "%s: k must be in [1, %d], got %d", what, TQ_SEARCH_MAX_K, k);
  if (m > 65535) return fail(TQ_EINVAL, "%s: m must be at most 65535", what);
  if (n >= 0xFFFFFFFFull) return fail(TQ_EINVAL, "%s: n must be below 2^32 - 1", what);
  if (m == 0) return TQ_OK;
  if (!out_dist || !out_ids || (n > 0 && !dist)) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t seg = (size_t)m * tq::topk_rows_segments((long)m, (long)n) * (size_t)k;
  void *buf = nullptr;
  hipError_t e = hipMallocAsync(&buf, tq::align_up(seg * sizeof(float)) + seg * sizeof(uint32_t), st);
  if (e != hipSuccess) return hip_fail(e, what);
  float *sd = static_cast<float *>(buf);
  uint32_t *si = reinterpret_cast<uint32_t *>(static_cast<char *>(buf) +
                                              tq::align_up(seg * sizeof(float)));
  e = tq::launch_topk_rows(dist, (long)m, (long)n, k, out_dist, out_ids, sd, si, st);
  hipError_t f = hipFreeAsync(buf, st);
  if (e != hipSuccess) return hip_fail(e, what);
  if (f != hipSuccess) return hip_fail(f, what);
  return TQ_OK;
}

}  // extern "C"
