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
}  // namespace tq
