// tq_serve.hip -- the request server of include/tq_hip.h: resident
// collections, admission against a device-memory budget, continuous batching
// over a fixed number of batch slots, and completion tracking.
//
// Threads: callers submit and wait; one scheduler thread forms batches and
// enqueues them; one completer thread waits for batches in order and hands
// results back. A batch slot owns a stream, a device workspace, pinned host
// staging and two events (append fence, done). Batch i + 1 may be enqueued
// on its own slot while batch i still runs, so their kernels overlap on the
// device.
//
// Ordering: an append reserves its code rows at submission; a search records
// the collection size at submission and reads only those rows, so a search
// never reads rows an append is still writing. Each batch first waits on the
// previous batch's append fence, so appends land in submission order and
// before any search that counts them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "tq_internal.h"
#include "tq_search.h"

namespace {

using tq::capi::fail;
using tq::capi::hip_fail;
using Clock = std::chrono::steady_clock;

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

enum Kind { kSearch = 0, kAppend = 1 };

struct Request {
  uint64_t id = 0;
  Kind kind = kSearch;
  uint32_t coll = 0;
  // The collection's space and code buffer, copied at submission under the
  // server lock: the scheduler reads them without the lock while other
  // threads may add collections.
  tq_space space = nullptr;
  uint8_t *codes = nullptr;
  const float *X = nullptr;  // queries or vectors (host)
  uint64_t rows = 0;
  uint64_t n_at = 0;  // search: codes visible to it; append: first row it writes
  float *out_d = nullptr;
  int64_t *out_i = nullptr;
  Clock::time_point t_submit, t_done;
  int status = TQ_OK;
  std::string err;
  bool done = false;
};

struct Collection {
  tq_space space = nullptr;
  uint8_t *codes = nullptr;
  uint64_t capacity = 0;
  uint64_t n = 0;  // rows reserved by submitted appends
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t fence = nullptr;  // after this batch's appends
  hipEvent_t done = nullptr;
  void *ws = nullptr;
  size_t ws_bytes = 0;
  void *pin = nullptr;  // pinned staging: inputs in, results out
  size_t pin_bytes = 0;
  std::vector<Request *> reqs;
  tq_search_plan plan = nullptr;
  // results in the pinned area, per search request in `reqs` order
  std::vector<size_t> res_off;
  size_t res_d = 0, res_i = 0;
  bool busy = false;
  int status = TQ_OK;
  std::string err;
};

}  // namespace

struct tq_server_s {
  tq_server_config cfg{};
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done, cv_complete;
  std::deque<Request *> pending;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  // A deque: adding a collection leaves the others where they are.
  std::deque<Collection> colls;
  std::vector<Slot> slots;
  std::deque<Slot *> inflight;
  Slot *last = nullptr;  // slot of the most recently enqueued batch
  uint64_t next_id = 1;
  uint64_t resident = 0, reserved = 0;
  int cus = 256;
  bool stop = false;
  bool sched_done = false;
  tq_server_stats stats{};
  std::thread sched, completer;
};

namespace {

void finish(tq_server_s *s, Request *r, int status, const std::string &err) {
  r->status = status;
  r->err = err;
  r->t_done = Clock::now();
  r->done = true;
  if (status != TQ_OK) s->stats.failed++;
}

// Grows the slot's pinned staging to `bytes`.
int ensure_pinned(Slot &sl, size_t bytes) {
  if (bytes <= sl.pin_bytes) return TQ_OK;
  if (sl.pin) (void)hipHostFree(sl.pin);
  sl.pin = nullptr;
  sl.pin_bytes = 0;
  hipError_t e = hipHostMalloc(&sl.pin, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return hip_fail(e, "tq_server: pinned staging");
  sl.pin_bytes = bytes;
  return TQ_OK;
}

// Enqueues one batch on its slot. Device layout: [queries and vectors |
// results_d | results_i | plan]; pinned layout: [queries and vectors |
// results_d | results_i].
int enqueue_batch(tq_server_s *s, Slot &sl, Slot *prev) {
  const char *what = "tq_server";
  hipError_t e;
  const int k = s->cfg.k;
  size_t in_bytes = 0, nres = 0;
  std::vector<size_t> in_off;
  for (Request *r : sl.reqs) {
    in_off.push_back(in_bytes);
    in_bytes += align_up(r->rows * r->space->info.dim * sizeof(float));
    if (r->kind == kSearch) nres += r->rows * k;
  }
  const size_t o_res_d = in_bytes;
  const size_t o_res_i = o_res_d + align_up(nres * sizeof(float));
  const size_t o_plan = o_res_i + align_up(nres * sizeof(int64_t));
  char *dws = static_cast<char *>(sl.ws);

  // The plan over this batch's searches, outputs in the result staging.
  std::vector<tq_search_job> jobs;
  sl.res_off.clear();
  size_t res = 0;
  for (size_t t = 0; t < sl.reqs.size(); ++t) {
    Request *r = sl.reqs[t];
    if (r->kind != kSearch) continue;
    jobs.push_back({r->space, reinterpret_cast<const float *>(dws + in_off[t]), r->rows, r->codes,
                    r->n_at, reinterpret_cast<float *>(dws + o_res_d) + res,
                    reinterpret_cast<int64_t *>(dws + o_res_i) + res});
    sl.res_off.push_back(res);
    res += r->rows * k;
  }
  size_t plan_bytes = 0;
  if (!jobs.empty()) {
    if (int r = tq::plan_new(jobs.data(), jobs.size(), k, s->cfg.mode, nullptr, &sl.plan))
      return r;
    plan_bytes = tq::plan_ws_bytes(sl.plan);
  }
  if (o_plan + plan_bytes > sl.ws_bytes)
    return fail(TQ_ENOMEM, "%s: batch needs %zu workspace bytes, the slot has %zu", what,
                o_plan + plan_bytes, sl.ws_bytes);

  // Inputs go through pinned memory so the uploads are asynchronous.
  const size_t pin_need = o_plan;
  if (int r = ensure_pinned(sl, pin_need ? pin_need : 256)) return r;
  char *pin = static_cast<char *>(sl.pin);
  for (size_t t = 0; t < sl.reqs.size(); ++t) {
    Request *r = sl.reqs[t];
    std::memcpy(pin + in_off[t], r->X, r->rows * r->space->info.dim * sizeof(float));
  }
  sl.res_d = o_res_d;
  sl.res_i = o_res_i;

  if (prev && (e = hipStreamWaitEvent(sl.st, prev->fence, 0)) != hipSuccess)
    return hip_fail(e, what);
  if (in_bytes &&
      (e = hipMemcpyAsync(dws, pin, in_bytes, hipMemcpyHostToDevice, sl.st)) != hipSuccess)
    return hip_fail(e, what);
  // Appends in submission order.
  for (size_t t = 0; t < sl.reqs.size(); ++t) {
    Request *r = sl.reqs[t];
    if (r->kind != kAppend || r->rows == 0) continue;
    e = tq::launch_encode(r->space->dev, reinterpret_cast<const float *>(dws + in_off[t]),
                          (long)r->rows, r->codes + r->n_at * r->space->info.code_size, sl.st);
    if (e != hipSuccess) return hip_fail(e, what);
  }
  if ((e = hipEventRecord(sl.fence, sl.st)) != hipSuccess) return hip_fail(e, what);

  if (sl.plan) {
    if (int r = tq::plan_bind(sl.plan, dws + o_plan, sl.ws_bytes - o_plan, sl.st)) return r;
    if (int r = tq::plan_enqueue(sl.plan, sl.st)) return r;
  }
  if (nres) {
    if ((e = hipMemcpyAsync(pin + o_res_d, dws + o_res_d, nres * sizeof(float),
                            hipMemcpyDeviceToHost, sl.st)) != hipSuccess ||
        (e = hipMemcpyAsync(pin + o_res_i, dws + o_res_i, nres * sizeof(int64_t),
                            hipMemcpyDeviceToHost, sl.st)) != hipSuccess)
      return hip_fail(e, what);
  }
  if ((e = hipEventRecord(sl.done, sl.st)) != hipSuccess) return hip_fail(e, what);
  return TQ_OK;
}

// Conservative device bytes of a batch: staging, plus per query its
// gathered, prepared and split copies and results, plus the candidate lists
// of the most work items a plan makes (3 workgroups per CU x 8 items each,
// and one more per query tile), plus an FP32 plan's distance block.
size_t batch_bound(tq_server_s *s, const std::vector<Request *> &b) {
  size_t bytes = 1 << 20;
  size_t qtiles = 0;
  bool any = false;
  for (Request *r : b) {
    const tq_space_info &in = r->space->info;
    bytes += align_up(r->rows * in.dim * sizeof(float)) + 2048;
    if (r->kind != kSearch) continue;
    any = true;
    // split query planes over rows rounded up to 32: at most 5 D halves per
    // row (4 planes, or the coarse form's 3 planes and its 2 D-halves
    // correction row) plus the correction row's two norms; the fused form's
    // running threshold per row
    const size_t planes_row = in.padded_dim * 10 + 8;
    bytes += r->rows * (in.dim * 4 + in.prepared_stride * 4 + planes_row + 4 +
                        (size_t)s->cfg.k * 24) + 31 * planes_row + 8 * 256;
    qtiles += (r->rows + 31) / 32;
  }
  // Searches the plan runs plain (fp32 scan in row blocks + row top-k): all of
  // an FP32 server's, and those over collections of padded dim below 32.
  bool plain = false;
  for (Request *r : b)
    plain = plain || (r->kind == kSearch && (s->cfg.mode == TQ_MODE_FP32 ||
                                             r->space->info.padded_dim < 32));
  if (plain)
    // a 512 MiB distance block and its row top-k lists: at most
    // (2^27 / 8192 + 65535) lists of k entries of 8 bytes
    bytes += (512ull << 20) + (size_t)s->cfg.k * (1ull << 20);
  if (any) {
    if (s->cfg.mode == TQ_MODE_MFMA) {
      if (s->cfg.k <= tq::kSearchMaxK)  // fused-form candidate lists
        bytes += (3 * (size_t)s->cus * 8 + qtiles + 16) * 4 * 64 * tq::kSearchMaxK * 8;
      long rows = 0, n_max = 0;
      bool blocked = false;
      for (Request *r : b)
        if (r->kind == kSearch) {
          rows += (long)r->rows;
          n_max = std::max(n_max, (long)r->n_at);
          blocked = blocked || tq::plan_blocked((long)r->n_at, 1);
        }
      if (blocked) bytes += tq::plan_blocked_ws_bound(rows, n_max, s->cfg.k);
    }
  }
  return bytes;
}

// Takes requests from the front of the queue while the batch stays within
// max_batch_queries and (by the bound above) the slot workspace; the first
// request is always taken. Called with the lock held.
std::vector<Request *> take_batch(tq_server_s *s, const Slot &sl) {
  std::vector<Request *> b;
  uint64_t queries = 0;
  while (!s->pending.empty()) {
    Request *r = s->pending.front();
    const uint64_t q = r->kind == kSearch ? r->rows : 0;
    if (!b.empty() && queries + q > s->cfg.max_batch_queries) break;
    b.push_back(r);
    if (b.size() > 1 && batch_bound(s, b) > sl.ws_bytes) {
      b.pop_back();
      break;
    }
    s->pending.pop_front();
    queries += q;
  }
  return b;
}

void scheduler(tq_server_s *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    Slot *free_slot = nullptr;
    s->cv_sched.wait(lk, [&] {
      if (s->stop && s->pending.empty()) return true;
      if (s->pending.empty()) return false;
      for (Slot &sl : s->slots)
        if (!sl.busy) return true;
      return false;
    });
    if (s->stop && s->pending.empty()) break;
    for (Slot &sl : s->slots)
      if (!sl.busy) {
        free_slot = &sl;
        break;
      }
    Slot &sl = *free_slot;
    sl.reqs = take_batch(s, sl);
    sl.busy = true;
    Slot *prev = s->last;
    s->last = &sl;
    lk.unlock();
    sl.status = TQ_OK;
    sl.err.clear();
    int rc = enqueue_batch(s, sl, prev);
    if (rc != TQ_OK) {
      sl.status = rc;
      sl.err = tq::capi::last_error();
      // The fence and done events still mark this point of the stream, so
      // later batches and the completer do not wait on stale records.
      (void)hipEventRecord(sl.fence, sl.st);
      (void)hipEventRecord(sl.done, sl.st);
    }
    lk.lock();
    s->stats.batches++;
    if (sl.reqs.size() > s->stats.max_batch_requests) s->stats.max_batch_requests = sl.reqs.size();
    s->inflight.push_back(&sl);
    s->cv_complete.notify_one();
  }
  s->sched_done = true;
  s->cv_complete.notify_all();
}

void completer(tq_server_s *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    // Exits only after the scheduler has, so no batch is left behind.
    s->cv_complete.wait(lk, [&] { return !s->inflight.empty() || s->sched_done; });
    if (s->inflight.empty()) break;
    Slot &sl = *s->inflight.front();
    lk.unlock();
    hipError_t e = hipEventSynchronize(sl.done);
    int status = sl.status;
    std::string err = sl.err;
    if (status == TQ_OK && e != hipSuccess) {
      status = hip_fail(e, "tq_server: batch");
      err = tq::capi::last_error();
    }
    const int k = s->cfg.k;
    if (status == TQ_OK) {
      const char *pin = static_cast<const char *>(sl.pin);
      size_t si = 0;
      for (Request *r : sl.reqs) {
        if (r->kind != kSearch) continue;
        const size_t off = sl.res_off[si++];
        std::memcpy(r->out_d, pin + sl.res_d + off * sizeof(float), r->rows * k * sizeof(float));
        std::memcpy(r->out_i, pin + sl.res_i + off * sizeof(int64_t),
                    r->rows * k * sizeof(int64_t));
      }
    }
    if (sl.plan) {
      tq_search_plan_destroy(sl.plan);
      sl.plan = nullptr;
    }
    lk.lock();
    for (Request *r : sl.reqs) {
      finish(s, r, status, err);
      if (status == TQ_OK) {
        if (r->kind == kSearch) {
          s->stats.searches_done++;
          s->stats.queries_done += r->rows;
        } else {
          s->stats.appends_done++;
        }
      }
    }
    sl.reqs.clear();
    sl.busy = false;
    s->inflight.pop_front();
    s->cv_done.notify_all();
    s->cv_sched.notify_one();
  }
}

void release(tq_server_s *s) {
  for (Slot &sl : s->slots) {
    if (sl.plan) tq_search_plan_destroy(sl.plan);
    if (sl.ws) (void)hipFree(sl.ws);
    if (sl.pin) (void)hipHostFree(sl.pin);
    if (sl.fence) (void)hipEventDestroy(sl.fence);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.st) (void)hipStreamDestroy(sl.st);
  }
  for (Collection &c : s->colls)
    if (c.codes) (void)hipFree(c.codes);
}

Request *new_request(tq_server_s *s, Kind kind, uint32_t coll, const float *X, uint64_t rows) {
  auto r = std::make_unique<Request>();
  r->id = s->next_id++;
  r->kind = kind;
  r->coll = coll;
  r->X = X;
  r->rows = rows;
  r->t_submit = Clock::now();
  Request *raw = r.get();
  s->reqs.emplace(raw->id, std::move(r));
  return raw;
}

}  // namespace

extern "C" {

int tq_server_create(const tq_server_config *cfg, tq_server *out) {
  const char *what = "tq_server_create";
  if (!cfg || !out) return fail(TQ_EINVAL, "%s: NULL argument", what);
  *out = nullptr;
  if (cfg->mode != TQ_MODE_MFMA && cfg->mode != TQ_MODE_FP32)
    return fail(TQ_EINVAL, "%s: unknown mode %d", what, cfg->mode);
  if (cfg->k < 1 || cfg->k > TQ_SEARCH_MAX_K)
    return fail(TQ_EINVAL, "%s: k must be in [1, %d], got %d", what, TQ_SEARCH_MAX_K, cfg->k);
  tq_server_config c = *cfg;
  if (!c.slot_bytes) c.slot_bytes = 1ull << 30;
  if (!c.max_inflight) c.max_inflight = 2;
  if (!c.max_batch_queries) c.max_batch_queries = 4096;
  if (c.max_inflight > 8) return fail(TQ_EINVAL, "%s: max_inflight must be at most 8", what);
  if (c.mode == TQ_MODE_MFMA && c.k > tq::kSearchMaxK && c.max_batch_queries > 65535)
    return fail(TQ_EINVAL, "%s: k > %d runs blocked, which takes at most 65535 queries per batch",
                what, tq::kSearchMaxK);
  const uint64_t reserved = (uint64_t)c.max_inflight * c.slot_bytes;
  if (reserved > c.hbm_budget)
    return fail(TQ_ENOMEM, "%s: %u slots of %llu bytes exceed the budget of %llu bytes", what,
                c.max_inflight, (unsigned long long)c.slot_bytes,
                (unsigned long long)c.hbm_budget);
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return hip_fail(e, what);
  tq_server_s *s = new (std::nothrow) tq_server_s();
  if (!s) return fail(TQ_ENOMEM, "%s: host allocation failed", what);
  s->cfg = c;
  s->device = device;
  s->cus = tq::capi::cu_count(device);
  s->slots.resize(c.max_inflight);
  for (Slot &sl : s->slots) {
    if ((e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&sl.fence, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipMalloc(&sl.ws, c.slot_bytes)) != hipSuccess) {
      if (e == hipErrorOutOfMemory) sl.ws = nullptr;
      release(s);
      delete s;
      return hip_fail(e, what);
    }
    sl.ws_bytes = c.slot_bytes;
  }
  s->reserved = reserved;
  s->stats.reserved_bytes = reserved;
  s->sched = std::thread(scheduler, s);
  s->completer = std::thread(completer, s);
  *out = s;
  return TQ_OK;
}

int tq_server_destroy(tq_server s) {
  if (!s) return TQ_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_sched.notify_all();
  s->cv_complete.notify_all();
  s->sched.join();
  s->completer.join();
  release(s);
  delete s;
  return TQ_OK;
}

int tq_server_add_collection(tq_server s, tq_space space, uint64_t capacity, uint32_t *coll) {
  const char *what = "tq_server_add_collection";
  if (!s || !space || !coll) return fail(TQ_EINVAL, "%s: NULL argument", what);
  if (capacity == 0 || capacity >= 0xFFFFFFFFull)
    return fail(TQ_EINVAL, "%s: capacity must be in [1, 2^32 - 2]", what);
  if (space->device != s->device)
    return fail(TQ_EINVAL, "%s: space is on device %d, the server on %d", what, space->device,
                s->device);
  const uint64_t bytes = capacity * space->info.code_size;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->reserved + s->resident + bytes > s->cfg.hbm_budget)
    return fail(TQ_ENOMEM, "%s: %llu bytes would exceed the budget (%llu held of %llu)", what,
                (unsigned long long)bytes, (unsigned long long)(s->reserved + s->resident),
                (unsigned long long)s->cfg.hbm_budget);
  Collection c;
  c.space = space;
  c.capacity = capacity;
  hipError_t e = hipSetDevice(s->device);
  if (e == hipSuccess) e = hipMalloc(&c.codes, bytes);
  if (e != hipSuccess) return hip_fail(e, what);
  s->colls.push_back(c);
  s->resident += bytes;
  s->stats.resident_bytes = s->resident;
  *coll = (uint32_t)(s->colls.size() - 1);
  return TQ_OK;
}

int tq_server_submit_append(tq_server s, uint32_t coll, const float *X, uint64_t n,
                            uint64_t *ticket) {
  const char *what = "tq_server_submit_append";
  if (!s || !ticket) return fail(TQ_EINVAL, "%s: NULL argument", what);
  if (n > 0 && !X) return fail(TQ_EINVAL, "%s: NULL vectors", what);
  std::lock_guard<std::mutex> lk(s->mu);
  if (coll >= s->colls.size()) return fail(TQ_EINVAL, "%s: no collection %u", what, coll);
  Collection &c = s->colls[coll];
  if (c.n + n > c.capacity)
    return fail(TQ_ENOMEM, "%s: collection %u holds %llu of %llu codes; %llu more do not fit",
                what, coll, (unsigned long long)c.n, (unsigned long long)c.capacity,
                (unsigned long long)n);
  if (align_up(n * c.space->info.dim * sizeof(float)) > s->cfg.slot_bytes / 2)
    return fail(TQ_ENOMEM, "%s: %llu vectors exceed the batch workspace", what,
                (unsigned long long)n);
  Request *r = new_request(s, kAppend, coll, X, n);
  r->space = c.space;
  r->codes = c.codes;
  r->n_at = c.n;
  c.n += n;
  s->pending.push_back(r);
  *ticket = r->id;
  s->cv_sched.notify_one();
  return TQ_OK;
}

int tq_server_submit_search(tq_server s, uint32_t coll, const float *Q, uint64_t m,
                            float *out_dist, int64_t *out_ids, uint64_t *ticket) {
  const char *what = "tq_server_submit_search";
  if (!s || !ticket) return fail(TQ_EINVAL, "%s: NULL argument", what);
  if (m > 0 && (!Q || !out_dist || !out_ids)) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  std::lock_guard<std::mutex> lk(s->mu);
  if (coll >= s->colls.size()) return fail(TQ_EINVAL, "%s: no collection %u", what, coll);
  if (m > s->cfg.max_batch_queries)
    return fail(TQ_EINVAL, "%s: %llu queries exceed max_batch_queries %u", what,
                (unsigned long long)m, s->cfg.max_batch_queries);
  Request *r = new_request(s, kSearch, coll, Q, m);
  r->space = s->colls[coll].space;
  r->codes = s->colls[coll].codes;
  r->n_at = s->colls[coll].n;
  r->out_d = out_dist;
  r->out_i = out_ids;
  s->pending.push_back(r);
  *ticket = r->id;
  s->cv_sched.notify_one();
  return TQ_OK;
}

int tq_server_wait(tq_server s, uint64_t ticket, double *latency_s) {
  const char *what = "tq_server_wait";
  if (!s) return fail(TQ_EINVAL, "%s: NULL server", what);
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end())
    return fail(TQ_EINVAL, "%s: unknown ticket %llu", what, (unsigned long long)ticket);
  Request *r = it->second.get();
  s->cv_done.wait(lk, [&] { return r->done; });
  if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
  const int status = r->status;
  const std::string err = r->err;
  s->reqs.erase(it);
  lk.unlock();
  if (status != TQ_OK) return fail(status, "%s", err.c_str());
  return TQ_OK;
}

int tq_server_collection(tq_server s, uint32_t coll, uint64_t *n, const uint8_t **codes) {
  if (!s) return fail(TQ_EINVAL, "tq_server_collection: NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (coll >= s->colls.size()) return fail(TQ_EINVAL, "tq_server_collection: no collection %u", coll);
  if (n) *n = s->colls[coll].n;
  if (codes) *codes = s->colls[coll].codes;
  return TQ_OK;
}

int tq_server_load_codes(tq_server s, uint32_t coll, const uint8_t *codes, uint64_t n,
                         int mem) {
  const char *what = "tq_server_load_codes";
  if (!s || (n && !codes)) return fail(TQ_EINVAL, "%s: NULL argument", what);
  if (mem != TQ_MEM_HOST && mem != TQ_MEM_DEVICE)
    return fail(TQ_EINVAL, "%s: mem must be TQ_MEM_HOST or TQ_MEM_DEVICE", what);
  uint8_t *dst = nullptr;
  size_t cs = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (coll >= s->colls.size()) return fail(TQ_EINVAL, "%s: no collection %u", what, coll);
    Collection &c = s->colls[coll];
    if (c.n + n > c.capacity)
      return fail(TQ_ENOMEM, "%s: collection %u holds %llu of %llu codes; %llu more do not fit",
                  what, coll, (unsigned long long)c.n, (unsigned long long)c.capacity,
                  (unsigned long long)n);
    cs = c.space->info.code_size;
    dst = c.codes + c.n * cs;
    c.n += n;
  }
  if (!n) return TQ_OK;
  hipError_t e = hipSetDevice(s->device);
  if (e == hipSuccess)
    e = hipMemcpy(dst, codes, n * cs,
                  mem == TQ_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, what);
  return TQ_OK;
}

int tq_server_read_codes(tq_server s, uint32_t coll, uint64_t start, uint64_t count,
                         uint8_t *host) {
  const char *what = "tq_server_read_codes";
  if (!s || (count && !host)) return fail(TQ_EINVAL, "%s: NULL argument", what);
  const uint8_t *codes = nullptr;
  size_t cs = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (coll >= s->colls.size()) return fail(TQ_EINVAL, "%s: no collection %u", what, coll);
    const Collection &c = s->colls[coll];
    if (start > c.n || count > c.n - start)
      return fail(TQ_EINVAL, "%s: rows [%llu, %llu) beyond the %llu codes held", what,
                  (unsigned long long)start, (unsigned long long)(start + count),
                  (unsigned long long)c.n);
    codes = c.codes;
    cs = c.space->info.code_size;
  }
  if (!count) return TQ_OK;
  hipError_t e = hipSetDevice(s->device);
  if (e == hipSuccess) e = hipMemcpy(host, codes + start * cs, count * cs, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, what);
  return TQ_OK;
}

int tq_server_stats_get(tq_server s, tq_server_stats *out) {
  if (!s || !out) return fail(TQ_EINVAL, "tq_server_stats_get: NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  return TQ_OK;
}

}  // extern "C"
