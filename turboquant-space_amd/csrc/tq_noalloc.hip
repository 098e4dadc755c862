// tq_noalloc.hip -- distance_m_to_n for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc). The same kernels as tq_distance_m_to_n; their scratch
// (prepared queries, then the MFMA query planes or the fp32 scan's QJL
// digit words) lives in one fixed per-device array of this library instead
// of stream-ordered allocations.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tq_internal.h"
#include "tq_mfma.h"

namespace {

using tq::capi::fail;
using tq::capi::hip_fail;

// Every GPU the library runs on holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[TQ_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device. In host code the array's name is only a stand-in for the symbol,
// never a device address.
int scratch_ptr(int device, void **out, const char *what) {
  if (device < 0 || device >= kMaxDevices)
    return fail(TQ_EUNSUPPORTED, "%s: device %d beyond the %d this library tracks", what, device,
                kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, what);
    if (!p) return fail(TQ_EHIP, "%s: no device address for the scratch area", what);
    g_ptr[device] = p;
  }
  *out = g_ptr[device];
  return TQ_OK;
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int tq_distance_m_to_n_noalloc(tq_space s, const float *Q, uint64_t m, const uint8_t *codes,
                               uint64_t n, float *out, int mode, void *stream) {
  const char *what = "tq_distance_m_to_n_noalloc";
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (mode != TQ_MODE_FP32 && mode != TQ_MODE_MFMA)
    return fail(TQ_EINVAL, "%s: unknown mode %d", what, mode);
  if (m == 0 || n == 0) return TQ_OK;
  if (!Q || !codes || !out) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (m > 65535L * 64)
    return fail(TQ_EINVAL, "%s: at most %ld queries per call", what, 65535L * 64);
  const size_t prep_bytes = align_up(m * s->info.prepared_stride * sizeof(float));
  // Padded dims below 32 (no MFMA tile) run the plain kernels in MFMA mode too.
  const bool mfma = mode == TQ_MODE_MFMA && s->dev.D >= 32;
  const size_t extra = mfma ? tq::split_planes_bytes(s->dev, (long)m)
                            : tq::scan_fp32_ws_bytes(s->dev, (long)m);
  const size_t need = prep_bytes + align_up(extra);
  if (need > TQ_NOALLOC_SCRATCH_BYTES)
    return fail(TQ_ENOMEM,
                "%s: %llu queries at padded dim %d need %zu scratch bytes, more than the %u the "
                "library holds",
                what, (unsigned long long)m, s->dev.D, need, (unsigned)TQ_NOALLOC_SCRATCH_BYTES);
  hipError_t e = hipSetDevice(s->device);
  if (e != hipSuccess) return hip_fail(e, what);
  void *base = nullptr;
  if (int r = scratch_ptr(s->device, &base, what)) return r;
  float *prep = static_cast<float *>(base);
  void *ws = static_cast<char *>(base) + prep_bytes;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // Every scratch byte a kernel reads is written earlier in the same call
  // (prepared rows, planes including their zero rows past m, digit words),
  // so nothing needs clearing.
  e = tq::launch_prepare(s->dev, Q, (long)m, prep, st);
  if (e == hipSuccess) {
    if (mfma) {
      auto *planes = static_cast<_Float16 *>(ws);
      e = tq::launch_split_queries(s->dev, prep, (long)m, planes, st);
      if (e == hipSuccess)
        e = tq::launch_mfma_planes(s->dev, planes, prep, (long)m, codes, (long)n, out, st);
    } else {
      e = tq::launch_scan_fp32(s->dev, prep, (long)m, codes, (long)n, out, false, st, ws);
    }
  }
  if (e != hipSuccess) return hip_fail(e, what);
  return TQ_OK;
}

}  // extern "C"
