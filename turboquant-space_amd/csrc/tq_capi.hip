// tq_capi.hip -- the extern "C" boundary of libtq_hip.so (declared in
// include/tq_hip.h): space construction on the host, device tables, staging
// of host buffers, and dispatch to the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "tq_internal.h"
#include "tq_mfma.h"
#include "tq_symfull.h"

namespace tq {
namespace capi {

namespace {
thread_local std::string g_err;
}

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory)
    return fail(TQ_ENOMEM, "%s: out of device memory", what);
  return fail(TQ_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int cu_count(int device) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return cus;
}

const char *last_error() { return g_err.c_str(); }

}  // namespace capi
}  // namespace tq

namespace {

using tq::capi::fail;
using tq::capi::hip_fail;

// ---- host tables ---------------------------------------------------------

// splitmix64 (include/turboquant/turbo_quant.h:27-39) and the sign rule of
// generateSigns (:74-86): bit 0 of each output set -> +1, else -1.
std::vector<float> make_signs(size_t d, uint64_t seed) {
  std::vector<float> out(d);
  uint64_t state = seed;
  for (size_t i = 0; i < d; ++i) {
    uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    out[i] = (z & 1ULL) ? 1.0f : -1.0f;
  }
  return out;
}

// Lloyd-Max quantizer for N(0,1) with 2^mse_bits levels, by the fixed-point
// iteration of computeLloydMax (space_turbo_quant.h:76-131): centroids start
// at (i+0.5)*3.5/half, boundaries are midpoints, each centroid is the
// conditional mean over its cell, at most 1000 sweeps or until the largest
// move is below 1e-12; the positive half is mirrored and rounded to fp32.
void make_codebook(int mse_bits, std::vector<float> &cent, std::vector<float> &bnd) {
  const int levels = 1 << mse_bits, half = levels / 2;
  const double kPi = 3.14159265358979323846;
  auto pdf = [&](double x) { return std::exp(-0.5 * x * x) / std::sqrt(2.0 * kPi); };
  auto cdf = [](double x) { return 0.5 * std::erfc(-x * 0.70710678118654752440); };
  std::vector<double> c(half), nc(half), b(half > 1 ? half - 1 : 1);
  for (int i = 0; i < half; ++i) c[i] = (i + 0.5) * 3.5 / half;
  for (int it = 0; it < 1000; ++it) {
    for (int i = 0; i + 1 < half; ++i) b[i] = 0.5 * (c[i] + c[i + 1]);
    double worst = 0.0;
    for (int i = 0; i < half; ++i) {
      double lo = i == 0 ? 0.0 : b[i - 1];
      double hi = i == half - 1 ? 1e10 : b[i];
      double mass = cdf(hi) - cdf(lo);
      nc[i] = mass < 1e-15 ? 0.5 * (lo + hi) : (pdf(lo) - pdf(hi)) / mass;
      worst = std::max(worst, std::fabs(nc[i] - c[i]));
    }
    c = nc;
    if (worst < 1e-12) break;
  }
  cent.assign(levels, 0.0f);
  bnd.assign(levels - 1, 0.0f);
  for (int i = 0; i < half; ++i) {
    cent[half + i] = static_cast<float>(c[i]);
    cent[half - 1 - i] = -static_cast<float>(c[i]);
  }
  bnd[half - 1] = 0.0f;
  for (int i = 0; i + 1 < half; ++i) {
    double m = 0.5 * (c[i] + c[i + 1]);
    bnd[half + i] = static_cast<float>(m);
    bnd[half - 2 - i] = -static_cast<float>(m);
  }
}

// ---- scoped device buffer ------------------------------------------------

struct DevBuf {
  void *p = nullptr;
  hipError_t alloc(size_t bytes) {
    if (bytes == 0) bytes = 4;
    return hipMalloc(&p, bytes);
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};


// Centroid scale of the fp16 MFMA operand. Byte codes and 2-3-bit nibble
// codes keep 2^8 and carry each centroid as c_hi + c_lo (two fp16 terms).
// The coarse form of 4-bit codes (tq::mfma_coarse) carries c_hi alone, so its
// scale alpha in [256, 512) (multiples of 2^-12, exact in fp32) is chosen so
// that fp16(fl32(c * alpha)) / alpha is within 2^-15 of each of the four
// centroid magnitudes (1.85e-5 for the 3-bit Lloyd-Max codebook). The search
// mirrors the device arithmetic: fp32 product, fp16 round to nearest even,
// fp32 reciprocal scale 1 / (2^10 alpha).
double fp16_rne(double v) {  // v > 0 in fp16's normal range
  int e = 0;
  std::frexp(v, &e);  // v = f * 2^e, f in [0.5, 1)
  const double ulp = std::ldexp(1.0, e - 11);
  return std::nearbyint(v / ulp) * ulp;
}

float coarse_centroid_scale(const std::vector<float> &cent) {
  static std::mutex mu;
  static float cached = 0.0f;
  std::lock_guard<std::mutex> lk(mu);
  if (cached != 0.0f) return cached;  // one 3-bit codebook for every 4-bit space
  const size_t L = cent.size();
  double best = 1.0, best_a = 256.0;
  for (int k = 0; k < (1 << 20); ++k) {
    const double a = 256.0 + std::ldexp((double)k, -12);
    const double u = (double)(float)(1.0 / (1024.0 * a));
    double worst = 0.0;
    for (size_t l = L / 2; l < L && worst < best; ++l) {
      const double c = std::fabs((double)cent[l]);
      const double v = (double)(float)(c * a);
      worst = std::max(worst, std::fabs(fp16_rne(v) * u * 1024.0 - c) / c);
    }
    if (worst < best) {
      best = worst;
      best_a = a;
    }
  }
  cached = best <= std::ldexp(1.0, -15) ? (float)best_a : 256.0f;
  return cached;
}

}  // namespace

extern "C" {

const char *tq_last_error(void) { return tq::capi::last_error(); }

int tq_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int tq_space_create(uint64_t dim, int bits, uint64_t rot_seed, uint64_t qjl_seed,
                    tq_space *out) {
  if (!out) return fail(TQ_EINVAL, "tq_space_create: out is NULL");
  *out = nullptr;
  if (dim < 1) return fail(TQ_EINVAL, "TurboQuantSpace: dim must be >= 1");
  if (bits < 2 || bits > 9)
    return fail(TQ_EINVAL, "TurboQuantSpace: bits_per_coord must be in [2, 9], got %d", bits);
  uint64_t D = 1;
  const uint64_t want = dim < 4 ? 4 : dim;
  while (D < want) D <<= 1;
  if (!tq::encode_supported((int)D) || D > (1ULL << 20))
    return fail(TQ_EUNSUPPORTED, "TurboQuantSpace: padded dim %llu is above the GPU limit 8192",
                (unsigned long long)D);
  tq_space_s *s = new (std::nothrow) tq_space_s();
  if (!s) return fail(TQ_ENOMEM, "tq_space_create: host allocation failed");
  s->info.dim = dim;
  s->info.padded_dim = D;
  s->info.bits_per_coord = bits;
  s->info.packed_bytes = bits <= 4 ? (D + 1) / 2 : D;
  s->info.code_size = s->info.packed_bytes + 12;
  s->info.num_levels = 1 << (bits - 1);
  s->info.scale = std::sqrt(static_cast<float>(3.14159265358979323846) / 2.0f) /
                  std::sqrt(static_cast<float>(D));
  s->info.padded = D != dim;
  s->info.prepared_stride = (uint64_t)tq::prep_stride((int)D);
  s->info.sym_prepared_stride = (uint64_t)tq::sym_prep_stride((int)D);
  make_codebook(bits - 1, s->cent, s->bnd);
  s->rot = make_signs(D, rot_seed);
  s->qjl = make_signs(D, qjl_seed);

  hipError_t e = hipGetDevice(&s->device);
  if (e != hipSuccess) {
    delete s;
    return hip_fail(e, "tq_space_create: no HIP device");
  }
  const size_t L = s->cent.size();
  const size_t nfloats = L + L + 2 * D;
  e = hipMalloc(&s->d_tables, nfloats * sizeof(float));
  if (e != hipSuccess) {
    delete s;
    return hip_fail(e, "tq_space_create: table allocation");
  }
  std::vector<float> host(nfloats, 0.0f);
  std::memcpy(host.data(), s->cent.data(), L * sizeof(float));
  std::memcpy(host.data() + L, s->bnd.data(), (L - 1) * sizeof(float));
  std::memcpy(host.data() + 2 * L, s->rot.data(), D * sizeof(float));
  std::memcpy(host.data() + 2 * L + D, s->qjl.data(), D * sizeof(float));
  e = hipMemcpy(s->d_tables, host.data(), nfloats * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(s->d_tables);
    delete s;
    return hip_fail(e, "tq_space_create: table upload");
  }
  tq::SpaceDev &d = s->dev;
  d.dim = (int)dim;
  d.D = (int)D;
  d.log2D = 0;
  while ((1ULL << d.log2D) < D) ++d.log2D;
  d.bits = bits;
  d.mse_bits = bits - 1;
  d.nib = bits <= 4;
  d.packed = (int)s->info.packed_bytes;
  d.cs = (int)s->info.code_size;
  d.L = (int)L;
  d.scale = s->info.scale;
  d.inv_sqrt_D = 1.0f / std::sqrt(static_cast<float>(D));
  // Centroids times 2^8 stay within fp16's normal range together with their
  // fp16 remainders (|c| < 4 at every bit width); the coarse form of 4-bit
  // spaces takes its single-term scale (coarse_centroid_scale: the 8
  // centroids of 3 MSE bits fit one; byte codes' 128 do not and keep 2^8).
  d.cscale = tq::mfma_coarse(d) && d.nib ? coarse_centroid_scale(s->cent) : 256.0f;
  d.mse_unscale = (float)(1.0 / (1024.0 * (double)d.cscale));
  // The same fp32 product and fp16 rounding as the device's cent_lo; code
  // units index at most 128 centroids (a byte holds 7 index bits).
  d.clo_max = 0.0f;
  if (tq::mfma_coarse(d))
    for (int k = 0; k < (int)std::min<size_t>(L, 128); ++k) {
      const float c = s->cent[k] * d.cscale;
      d.clo_max = std::max(d.clo_max, std::fabs(c - (float)(_Float16)c));
    }
  d.cent = s->d_tables;
  d.bnd = s->d_tables + L;
  d.rot = s->d_tables + 2 * L;
  d.qjl = s->d_tables + 2 * L + D;
  *out = s;
  return TQ_OK;
}

int tq_space_destroy(tq_space s) {
  if (!s) return TQ_OK;
  if (s->d_tables) (void)hipFree(s->d_tables);
  delete s;
  return TQ_OK;
}

int tq_space_info_get(tq_space s, tq_space_info *out) {
  if (!s || !out) return fail(TQ_EINVAL, "tq_space_info_get: NULL argument");
  *out = s->info;
  return TQ_OK;
}

int tq_space_tables(tq_space s, float *centroids, float *boundaries, float *rot_signs,
                    float *qjl_signs) {
  if (!s) return fail(TQ_EINVAL, "tq_space_tables: NULL space");
  if (centroids) std::memcpy(centroids, s->cent.data(), s->cent.size() * sizeof(float));
  if (boundaries) std::memcpy(boundaries, s->bnd.data(), s->bnd.size() * sizeof(float));
  if (rot_signs) std::memcpy(rot_signs, s->rot.data(), s->rot.size() * sizeof(float));
  if (qjl_signs) std::memcpy(qjl_signs, s->qjl.data(), s->qjl.size() * sizeof(float));
  return TQ_OK;
}

}  // extern "C"

namespace {

// Runs `body(device pointers...)` on device memory. For TQ_MEM_HOST the inputs
// are copied in, the outputs copied back and the stream synchronised.
struct Staged {
  const void *src;
  size_t bytes;
  bool is_output;
  void *dev = nullptr;
  DevBuf buf;
};

int stage_in(Staged &b, int mem, hipStream_t st, const char *what) {
  if (mem == TQ_MEM_DEVICE) {
    b.dev = const_cast<void *>(b.src);
    return TQ_OK;
  }
  hipError_t e = b.buf.alloc(b.bytes);
  if (e != hipSuccess) return hip_fail(e, what);
  b.dev = b.buf.p;
  if (!b.is_output && b.bytes) {
    e = hipMemcpyAsync(b.dev, b.src, b.bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, what);
  }
  return TQ_OK;
}

int stage_out(Staged &b, int mem, hipStream_t st, const char *what) {
  if (mem == TQ_MEM_DEVICE || !b.bytes) return TQ_OK;
  hipError_t e = hipMemcpyAsync(const_cast<void *>(b.src), b.dev, b.bytes,
                                hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return hip_fail(e, what);
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, what);
  return TQ_OK;
}

int check_mem(int mem, const char *what) {
  if (mem != TQ_MEM_HOST && mem != TQ_MEM_DEVICE)
    return fail(TQ_EINVAL, "%s: mem must be TQ_MEM_HOST or TQ_MEM_DEVICE", what);
  return TQ_OK;
}

int select_device(tq_space s, const char *what) {
  hipError_t e = hipSetDevice(s->device);
  if (e != hipSuccess) return hip_fail(e, what);
  return TQ_OK;
}

// Scan of prepared queries against codes, all pointers on the device.
int scan_prepared(tq_space s, const float *prep, uint64_t m, const uint8_t *codes, uint64_t n,
                  float *out, int mode, hipStream_t st, const char *what) {
  hipError_t e;
  if (mode == TQ_MODE_MFMA) {
    e = tq::launch_scan_mfma(s->dev, prep, (long)m, codes, (long)n, out, st);
    if (e == hipErrorNotSupported)
      return fail(TQ_EUNSUPPORTED, "%s: MFMA mode does not support padded dim %d", what,
                  s->dev.D);
  } else if (mode == TQ_MODE_FP32) {
    e = tq::launch_scan_fp32(s->dev, prep, (long)m, codes, (long)n, out, false, st);
  } else {
    return fail(TQ_EINVAL, "%s: unknown mode %d", what, mode);
  }
  if (e != hipSuccess) return hip_fail(e, what);
  return TQ_OK;
}

}  // namespace

extern "C" {

int tq_encode_batch(tq_space s, const float *X, uint64_t n, uint8_t *codes, int mem,
                    void *stream) {
  const char *what = "tq_encode_batch";
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (int r = check_mem(mem, what)) return r;
  if (n == 0) return TQ_OK;
  if (!X || !codes) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (int r = select_device(s, what)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Staged in{X, n * s->info.dim * sizeof(float), false};
  Staged outb{codes, n * s->info.code_size, true};
  if (int r = stage_in(in, mem, st, what)) return r;
  if (int r = stage_in(outb, mem, st, what)) return r;
  hipError_t e = tq::launch_encode(s->dev, static_cast<const float *>(in.dev), (long)n,
                                   static_cast<uint8_t *>(outb.dev), st);
  if (e != hipSuccess) return hip_fail(e, what);
  return stage_out(outb, mem, st, what);
}

struct tq_encode_graph_s {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~tq_encode_graph_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {
// True when `p` is device (or managed) memory of `device`, so a kernel
// replayed from a graph can read it; host memory is not.
bool on_device(const void *p, int device) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) &&
         (a.type == hipMemoryTypeManaged || a.device == device);
}
}  // namespace

int tq_encode_graph_create(tq_space s, const float *X, uint64_t n, uint8_t *codes,
                           tq_encode_graph *out) {
  const char *what = "tq_encode_graph_create";
  if (!out) return fail(TQ_EINVAL, "%s: out is NULL", what);
  *out = nullptr;
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (n == 0 || n > (uint64_t)INT64_MAX / 4096) return fail(TQ_EINVAL, "%s: bad n", what);
  if (!X || !codes) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (int r = select_device(s, what)) return r;
  if (!on_device(X, s->device) || !on_device(codes, s->device))
    return fail(TQ_EINVAL, "%s: X and codes must be device memory of the space's device", what);
  hipStream_t cap = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, what);
  hipGraph_t g = nullptr;
  e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    const hipError_t le = tq::launch_encode(s->dev, X, (long)n, codes, cap);
    e = hipStreamEndCapture(cap, &g);
    if (le != hipSuccess) e = le;
  }
  (void)hipStreamDestroy(cap);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(e, what);
  }
  auto *eg = new (std::nothrow) tq_encode_graph_s;
  if (!eg) {
    (void)hipGraphDestroy(g);
    return fail(TQ_ENOMEM, "%s: out of host memory", what);
  }
  eg->device = s->device;
  eg->graph = g;
  if ((e = hipGraphInstantiate(&eg->exec, g, nullptr, nullptr, 0)) != hipSuccess) {
    eg->exec = nullptr;
    delete eg;
    return hip_fail(e, what);
  }
  *out = eg;
  return TQ_OK;
}

int tq_encode_graph_run(tq_encode_graph g, void *stream) {
  const char *what = "tq_encode_graph_run";
  if (!g) return fail(TQ_EINVAL, "%s: NULL graph", what);
  hipError_t e = hipSetDevice(g->device);
  if (e == hipSuccess) e = hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? TQ_OK : hip_fail(e, what);
}

int tq_encode_graph_destroy(tq_encode_graph g) {
  if (!g) return TQ_OK;
  (void)hipSetDevice(g->device);
  delete g;
  return TQ_OK;
}

int tq_prepare_queries(tq_space s, const float *Q, uint64_t m, float *prepared, int mem,
                       void *stream) {
  const char *what = "tq_prepare_queries";
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (int r = check_mem(mem, what)) return r;
  if (m == 0) return TQ_OK;
  if (!Q || !prepared) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (int r = select_device(s, what)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Staged in{Q, m * s->info.dim * sizeof(float), false};
  Staged outb{prepared, m * s->info.prepared_stride * sizeof(float), true};
  if (int r = stage_in(in, mem, st, what)) return r;
  if (int r = stage_in(outb, mem, st, what)) return r;
  hipError_t e = tq::launch_prepare(s->dev, static_cast<const float *>(in.dev), (long)m,
                                    static_cast<float *>(outb.dev), st);
  if (e != hipSuccess) return hip_fail(e, what);
  return stage_out(outb, mem, st, what);
}

int tq_distance_prepared(tq_space s, const float *prepared, uint64_t m, const uint8_t *codes,
                         uint64_t n, float *out, int mode, void *stream) {
  const char *what = "tq_distance_prepared";
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (m == 0 || n == 0) return TQ_OK;
  if (!prepared || !codes || !out) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (int r = select_device(s, what)) return r;
  return scan_prepared(s, prepared, m, codes, n, out, mode, static_cast<hipStream_t>(stream),
                       what);
}

int tq_distance_m_to_n(tq_space s, const float *Q, uint64_t m, const uint8_t *codes, uint64_t n,
                       float *out, int mode, int mem, void *stream) {
  const char *what = "tq_distance_m_to_n";
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (int r = check_mem(mem, what)) return r;
  if (mode != TQ_MODE_FP32 && mode != TQ_MODE_MFMA)
    return fail(TQ_EINVAL, "%s: unknown mode %d", what, mode);
  if (m == 0 || n == 0) return TQ_OK;
  if (!Q || !codes || !out) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (int r = select_device(s, what)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Staged qb{Q, m * s->info.dim * sizeof(float), false};
  Staged cb{codes, n * s->info.code_size, false};
  Staged ob{out, m * n * sizeof(float), true};
  if (int r = stage_in(qb, mem, st, what)) return r;
  if (int r = stage_in(cb, mem, st, what)) return r;
  if (int r = stage_in(ob, mem, st, what)) return r;
  DevBuf prep;
  hipError_t e = prep.alloc(m * s->info.prepared_stride * sizeof(float));
  if (e != hipSuccess) return hip_fail(e, what);
  e = tq::launch_prepare(s->dev, static_cast<const float *>(qb.dev), (long)m,
                         static_cast<float *>(prep.p), st);
  if (e != hipSuccess) return hip_fail(e, what);
  if (int r = scan_prepared(s, static_cast<const float *>(prep.p), m,
                            static_cast<const uint8_t *>(cb.dev), n, static_cast<float *>(ob.dev),
                            mode, st, what))
    return r;
  if (mem == TQ_MEM_DEVICE) {
    // `prep` is freed on return; make sure the scan that reads it is done.
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, what);
    return TQ_OK;
  }
  return stage_out(ob, mem, st, what);
}

int tq_distance_1_to_n(tq_space s, const float *query, const uint8_t *codes, uint64_t n,
                       float *out, int mem, void *stream) {
  return tq_distance_m_to_n(s, query, 1, codes, n, out, TQ_MODE_FP32, mem, stream);
}

int tq_distance_m_to_n_symmetric_ex(tq_space s, const uint8_t *codes_a, uint64_t m,
                                    const uint8_t *codes_b, uint64_t n, float *out, int variant,
                                    int mode, int mem, void *stream) {
  const char *what = "tq_distance_m_to_n_symmetric";
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (int r = check_mem(mem, what)) return r;
  if (variant != TQ_SYM_MSE && variant != TQ_SYM_FULL && variant != TQ_SYM_LIGHT)
    return fail(TQ_EINVAL, "%s: unknown variant %d", what, variant);
  if (mode != TQ_MODE_FP32 && mode != TQ_MODE_MFMA)
    return fail(TQ_EINVAL, "%s: unknown mode %d", what, mode);
  if (mode == TQ_MODE_MFMA && variant != TQ_SYM_MSE)
    return fail(TQ_EUNSUPPORTED, "%s: the QJL-corrected variants run in fp32 only", what);
  if (m == 0 || n == 0) return TQ_OK;
  if (!codes_a || !codes_b || !out) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (int r = select_device(s, what)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Staged ab{codes_a, m * s->info.code_size, false};
  Staged bb{codes_b, n * s->info.code_size, false};
  Staged ob{out, m * n * sizeof(float), true};
  if (int r = stage_in(ab, mem, st, what)) return r;
  if (int r = stage_in(bb, mem, st, what)) return r;
  if (int r = stage_in(ob, mem, st, what)) return r;
  const auto *da = static_cast<const uint8_t *>(ab.dev);
  const auto *db = static_cast<const uint8_t *>(bb.dev);
  float *dout = static_cast<float *>(ob.dev);
  DevBuf prep;
  hipError_t e;
  if (variant == TQ_SYM_MSE) {
    e = prep.alloc(m * s->info.prepared_stride * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, what);
    auto *dp = static_cast<float *>(prep.p);
    e = tq::launch_prepare_from_codes(s->dev, da, (long)m, dp, st);
    if (e == hipSuccess) {
      if (mode == TQ_MODE_MFMA) {
        e = tq::launch_scan_mfma(s->dev, dp, (long)m, db, (long)n, dout, st, /*qjl=*/false);
        if (e == hipErrorNotSupported)
          return fail(TQ_EUNSUPPORTED, "%s: MFMA mode does not support padded dim %d", what,
                      s->dev.D);
      } else {
        e = tq::launch_scan_fp32(s->dev, dp, (long)m, db, (long)n, dout, true, st);
      }
    }
  } else {
    e = prep.alloc(m * s->info.sym_prepared_stride * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, what);
    auto *dp = static_cast<float *>(prep.p);
    e = tq::launch_prepare_sym(s->dev, da, (long)m, variant == TQ_SYM_LIGHT, dp, st);
    if (e == hipSuccess) e = tq::launch_sym_qjl(s->dev, dp, (long)m, db, (long)n, dout, st);
  }
  if (e != hipSuccess) return hip_fail(e, what);
  if (mem == TQ_MEM_DEVICE) {
    // `prep` is freed on return; make sure the scan that reads it is done.
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, what);
    return TQ_OK;
  }
  return stage_out(ob, mem, st, what);
}

int tq_distance_m_to_n_symmetric(tq_space s, const uint8_t *codes_a, uint64_t m,
                                 const uint8_t *codes_b, uint64_t n, float *out, int mem,
                                 void *stream) {
  return tq_distance_m_to_n_symmetric_ex(s, codes_a, m, codes_b, n, out, TQ_SYM_MSE,
                                         TQ_MODE_FP32, mem, stream);
}

int tq_prepare_symmetric(tq_space s, const uint8_t *codes, uint64_t m, float *prepared, int mem,
                         void *stream) {
  const char *what = "tq_prepare_symmetric";
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (int r = check_mem(mem, what)) return r;
  if (m == 0) return TQ_OK;
  if (!codes || !prepared) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (int r = select_device(s, what)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Staged cb{codes, m * s->info.code_size, false};
  Staged ob{prepared, m * s->info.sym_prepared_stride * sizeof(float), true};
  if (int r = stage_in(cb, mem, st, what)) return r;
  if (int r = stage_in(ob, mem, st, what)) return r;
  hipError_t e = tq::launch_prepare_sym(s->dev, static_cast<const uint8_t *>(cb.dev), (long)m,
                                        false, static_cast<float *>(ob.dev), st);
  if (e != hipSuccess) return hip_fail(e, what);
  return stage_out(ob, mem, st, what);
}

int tq_distance_symmetric_prepared(tq_space s, const float *prepared, uint64_t m,
                                   const uint8_t *codes, uint64_t n, float *out, int mem,
                                   void *stream) {
  const char *what = "tq_distance_symmetric_prepared";
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (int r = check_mem(mem, what)) return r;
  if (m == 0 || n == 0) return TQ_OK;
  if (!prepared || !codes || !out) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (int r = select_device(s, what)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Staged pb{prepared, m * s->info.sym_prepared_stride * sizeof(float), false};
  Staged cb{codes, n * s->info.code_size, false};
  Staged ob{out, m * n * sizeof(float), true};
  if (int r = stage_in(pb, mem, st, what)) return r;
  if (int r = stage_in(cb, mem, st, what)) return r;
  if (int r = stage_in(ob, mem, st, what)) return r;
  hipError_t e = tq::launch_sym_qjl(s->dev, static_cast<const float *>(pb.dev), (long)m,
                                    static_cast<const uint8_t *>(cb.dev), (long)n,
                                    static_cast<float *>(ob.dev), st);
  if (e != hipSuccess) return hip_fail(e, what);
  return stage_out(ob, mem, st, what);
}

int tq_rotated_coords_batch(tq_space s, const float *X, uint64_t n, float *out, int mem,
                            void *stream) {
  const char *what = "tq_rotated_coords_batch";
  if (!s) return fail(TQ_EINVAL, "%s: NULL space", what);
  if (int r = check_mem(mem, what)) return r;
  if (n == 0) return TQ_OK;
  if (!X || !out) return fail(TQ_EINVAL, "%s: NULL buffer", what);
  if (int r = select_device(s, what)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Staged in{X, n * s->info.dim * sizeof(float), false};
  Staged ob{out, n * s->info.padded_dim * sizeof(float), true};
  if (int r = stage_in(in, mem, st, what)) return r;
  if (int r = stage_in(ob, mem, st, what)) return r;
  hipError_t e = tq::launch_rotated(s->dev, static_cast<const float *>(in.dev), (long)n,
                                    static_cast<float *>(ob.dev), st);
  if (e != hipSuccess) return hip_fail(e, what);
  return stage_out(ob, mem, st, what);
}

}  // extern "C"
