// tq_mfma.hip -- distance_m_to_n as a matrix-core contraction on gfx950.
//
// The asymmetric estimator (space_turbo_quant.h:659-686, :951-983) is, over a
// batch, two dense contractions between the queries and the decoded codes:
//   MSE[q, j] = sum_i q_rot[q, i] * c[idx_{j,i}]
//   QJL[q, j] = sum_i s_q[q, i]  * (qjl_{j,i} ? +1 : -1)
// followed by a per-pair fp32 epilogue with the code's norm/gamma/sigma and
// the query's norm (identical to the plain path's).
//
// Precision: fp16 holds neither the queries nor the Lloyd-Max centroids to
// fp32 accuracy, so both are split into hi + lo fp16 parts (v = hi + lo with
// hi = fp16(v), lo = fp16(v - hi), ~22 significant bits) and the MSE term is
// accumulated as q_hi*c_hi + q_lo*c_hi + q_hi*c_lo (the dropped q_lo*c_lo is
// ~2^-22 relative); the +-1 QJL factors are exact in fp16, so that term is
// s_hi*t + s_lo*t. Everything accumulates in fp32 (v_mfma_f32_32x32x16_f16).
// Queries are scaled by 2^10 and centroids by 2^8 before the split so the lo
// parts stay in fp16's normal range; the epilogue removes the exact powers of
// two. Measured error against the fp64 restatement of the estimator is held
// below 1e-4 of (||q||^2 + ||x||^2).
//
// Tiling: a workgroup (4 wavefronts) owns 64 queries x 256 codes; each
// wavefront owns 64 queries x 64 codes as 2 x 2 blocks of 32 x 32 MFMA
// outputs, walking the padded dim in steps of 16. Per step a lane loads one
// 32-bit word of code bytes per 32-code block (8 coordinates of one code:
// exactly the B-operand slice it contributes) and decodes it in registers
// against a (c_hi | c_lo << 16) table in LDS; the A operands (query hi/lo
// planes, written once per batch by k_split_queries) are 16-byte loads that
// stay in L2.
#include "tq_mfma.h"

namespace tq {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr float kQScale = 1024.0f;              // 2^10 on q_rot and s_q
constexpr float kCScale = 256.0f;               // 2^8 on the centroids
constexpr float kMseUnscale = 1.0f / 262144.0f; // 2^-18
constexpr float kQjlUnscale = 1.0f / 1024.0f;   // 2^-10

__device__ __forceinline__ uint32_t h_bits(_Float16 h) {
  return (uint32_t)__builtin_bit_cast(uint16_t, h);
}

// planes: [4][m][D] fp16 = q_hi, q_lo, s_hi, s_lo (scaled by 2^10).
__global__ __launch_bounds__(256) void k_split_queries(SpaceDev s, const float *__restrict__ prep,
                                                       long m, _Float16 *__restrict__ planes) {
  const long stride = prep_stride(s.D);
  const long mD = m * (long)s.D;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < mD;
       e += (long)gridDim.x * blockDim.x) {
    const long q = e / s.D;
    const int i = (int)(e - q * s.D);
    const float v = prep[q * stride + i] * kQScale;
    const float w = prep[q * stride + s.D + i] * kQScale;
    const _Float16 vh = (_Float16)v, wh = (_Float16)w;
    planes[e] = vh;
    planes[mD + e] = (_Float16)(v - (float)vh);
    planes[2 * mD + e] = wh;
    planes[3 * mD + e] = (_Float16)(w - (float)wh);
  }
}

// Units of 8 consecutive coordinates -> B-operand fragments.
__device__ __forceinline__ void decode8(const uint32_t (&u)[8], const uint32_t *l_tab, h8 &bh,
                                        h8 &bl, h8 &bt) {
  uint32_t ph[4], pl[4], pt[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const uint32_t t0 = l_tab[u[2 * p] >> 1];
    const uint32_t t1 = l_tab[u[2 * p + 1] >> 1];
    ph[p] = (t0 & 0xFFFFu) | (t1 << 16);
    pl[p] = (t0 >> 16) | (t1 & 0xFFFF0000u);
    pt[p] = 0xBC00BC00u ^ ((u[2 * p] & 1u) << 15) ^ ((u[2 * p + 1] & 1u) << 31);
  }
  bh = __builtin_bit_cast(h8, ph);
  bl = __builtin_bit_cast(h8, pl);
  bt = __builtin_bit_cast(h8, pt);
}

template <bool NIB>
__global__ __launch_bounds__(256) void k_mfma(SpaceDev s, const _Float16 *__restrict__ planes,
                                              const float *__restrict__ prep, long m,
                                              const uint8_t *__restrict__ codes, long n,
                                              float *__restrict__ out) {
  __shared__ uint32_t l_tab[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    const float c = i < s.L ? s.cent[i] * kCScale : 0.0f;
    const _Float16 ch = (_Float16)c;
    const _Float16 cl = (_Float16)(c - (float)ch);
    l_tab[i] = h_bits(ch) | (h_bits(cl) << 16);
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lr = lane & 31, kg = lane >> 5;
  const long q0 = (long)blockIdx.y * 64;
  const long j0 = ((long)blockIdx.x * 4 + wave) * 64;
  const long mD = m * (long)s.D;

  const _Float16 *arow[2];
  bool qok[2];
  const uint8_t *crow[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const long q = q0 + 32 * r + lr;
    qok[r] = q < m;
    arow[r] = planes + (qok[r] ? q : 0L) * s.D + 8 * kg;
    const long j = j0 + 32 * r + lr;
    crow[r] = codes + (j < n ? j : 0L) * (long)s.cs;
  }

  f16v acc_m[2][2], acc_q[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        acc_m[r][c][v] = 0.0f;
        acc_q[r][c][v] = 0.0f;
      }

  const h8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  const int nkb = s.D / 16;
  for (int kb = 0; kb < nkb; ++kb) {
    h8 a[2][4];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int p = 0; p < 4; ++p)
        a[r][p] = qok[r] ? *reinterpret_cast<const h8 *>(arow[r] + p * mD + kb * 16) : zero8;
    h8 bh[2], bl[2], bt[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      uint32_t u[8];
      if constexpr (NIB) {
        const uint32_t w = *reinterpret_cast<const uint32_t *>(crow[c] + kb * 8 + 4 * kg);
#pragma unroll
        for (int v = 0; v < 8; ++v) u[v] = (w >> (4 * v)) & 0xFu;
      } else {
        const uint32_t *wp = reinterpret_cast<const uint32_t *>(crow[c] + kb * 16 + 8 * kg);
        const uint32_t w0 = wp[0], w1 = wp[1];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          u[v] = (w0 >> (8 * v)) & 0xFFu;
          u[4 + v] = (w1 >> (8 * v)) & 0xFFu;
        }
      }
      decode8(u, l_tab, bh[c], bl[c], bt[c]);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        acc_m[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[r][0], bh[c], acc_m[r][c], 0, 0, 0);
        acc_m[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[r][1], bh[c], acc_m[r][c], 0, 0, 0);
        acc_m[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[r][0], bl[c], acc_m[r][c], 0, 0, 0);
        acc_q[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[r][2], bt[c], acc_q[r][c], 0, 0, 0);
        acc_q[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[r][3], bt[c], acc_q[r][c], 0, 0, 0);
      }
  }

  // Epilogue: output element v of a 32x32 block sits at row
  // 8*(v/4) + 4*kg + (v%4), column lr.
  const long pstride = prep_stride(s.D);
  const bool al = (s.cs & 3) == 0;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const long j = j0 + 32 * c + lr;
    if (j >= n) continue;
    const uint8_t *cp = codes + j * (long)s.cs;
    const float xn = load_f32(cp + s.packed, al);
    const float gamma = load_f32(cp + s.packed + 4, al);
    const float sigma = load_f32(cp + s.packed + 8, al);
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const long q = q0 + 32 * r + 8 * (v >> 2) + 4 * kg + (v & 3);
        if (q >= m) continue;
        const float qns = prep[q * pstride + 2 * s.D];
        const float qn = prep[q * pstride + 2 * s.D + 1];
        const float ip_mse = (acc_m[r][c][v] * kMseUnscale) * sigma;
        const float corr = s.scale * gamma * (acc_q[r][c][v] * kQjlUnscale);
        const float ip = (ip_mse + corr) * xn * qn;
        const float d = qns + xn * xn - 2.0f * ip;
        out[q * n + j] = d > 0.0f ? d : 0.0f;
      }
  }
}

}  // namespace

hipError_t launch_scan_mfma(const SpaceDev &s, const float *prep, long m, const uint8_t *codes,
                            long n, float *out, hipStream_t st) {
  if (s.D < 16) return hipErrorNotSupported;
  if (m <= 0 || n <= 0) return hipSuccess;
  // grid.y covers 64 queries per workgroup row; larger batches go in chunks.
  const long chunk_max = 65535L * 64;
  const long chunk = m < chunk_max ? m : chunk_max;
  _Float16 *planes = nullptr;
  const size_t bytes = 4 * (size_t)chunk * s.D * sizeof(_Float16);
  hipError_t e = hipMallocAsync(reinterpret_cast<void **>(&planes), bytes, st);
  if (e != hipSuccess) return e;
  const long pstride = prep_stride(s.D);
  for (long q0 = 0; q0 < m && e == hipSuccess; q0 += chunk) {
    const long mc = (m - q0) < chunk ? (m - q0) : chunk;
    const float *pc = prep + q0 * pstride;
    float *oc = out + q0 * n;
    long blocks = (mc * (long)s.D + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_split_queries, dim3((unsigned)blocks), dim3(256), 0, st, s, pc, mc,
                       planes);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)((mc + 63) / 64));
    if (s.nib)
      hipLaunchKernelGGL(k_mfma<true>, grid, dim3(256), 0, st, s, planes, pc, mc, codes, n, oc);
    else
      hipLaunchKernelGGL(k_mfma<false>, grid, dim3(256), 0, st, s, planes, pc, mc, codes, n, oc);
    e = hipGetLastError();
  }
  hipError_t f = hipFreeAsync(planes, st);
  return e != hipSuccess ? e : f;
}

}  // namespace tq
