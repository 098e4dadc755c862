// tq_internal.h -- host-side state shared by the C-ABI translation units of
// libtq_hip.so (tq_capi.hip, tq_plan.hip). Not installed.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/tq_hip.h"
#include "tq_common.h"

// The space handle: host tables (what tq_space_tables returns) and their
// device copy, as the kernels see them.
struct tq_space_s {
  tq_space_info info{};
  int device = 0;
  std::vector<float> cent, bnd, rot, qjl;
  float *d_tables = nullptr;  // [cent L | bnd L-1 (+pad) | rot D | qjl D]
  tq::SpaceDev dev{};
};

namespace tq {
namespace capi {
// Records the thread-local error text and returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// TQ_ENOMEM for hipErrorOutOfMemory, else TQ_EHIP with the HIP text.
int hip_fail(hipError_t e, const char *what);
// Compute units of `device` (256 if the query fails).
int cu_count(int device);
const char *last_error();
}  // namespace capi

// Search plans without their own memory (tq_plan.hip): plan_new checks the
// jobs and lays out the workspace on the host; plan_bind places it in caller
// memory of at least plan_ws_bytes and enqueues the table uploads on `st`;
// plan_enqueue is tq_search_plan_run. Destroy with tq_search_plan_destroy.
int plan_new(const tq_search_job *jobs, uint64_t njobs, int k, int mode,
             const tq_search_tuning *tuning, tq_search_plan *out);
size_t plan_ws_bytes(tq_search_plan p);
int plan_bind(tq_search_plan p, void *ws, size_t bytes, hipStream_t st);
int plan_enqueue(tq_search_plan p, hipStream_t st);
// MFMA plans can run a batch of `rows` queries over `n` codes blocked
// (distance blocks + row top-k) when plan_blocked(n, rows), and do unless
// their knobs ask for the fused form; plan_blocked_ws_bound(rows, n_max, k)
// bounds the workspace that adds for batches of up to `rows` queries over at
// most n_max codes.
bool plan_blocked(long n, long rows);
size_t plan_blocked_ws_bound(long rows, long n_max, int k);
}  // namespace tq
