// tq_symfull.hip -- symmetric code x code distances with QJL correction:
// the "full" and "light" estimators (distBuildFullScalar, distBuildLightScalar,
// distBuildFullPreparedScalar; include/turboquant/space_turbo_quant.h:731-945).
//
// For codes a, b with reconstructions r_a = c[idx_a]*sigma_a, QJL signs s_a,
// and H = RHT_qjl (sign flip by the QJL signs, WHT, times fl(1/sqrt(D))):
//   ip_mse = <r_a, r_b>
//   full:  corr = scale*g_b*<H r_a, s_b> + scale*g_a*<s_a, H r_b>
//                 + pi/(2D)*g_a*g_b*<s_a, s_b>
//   light: corr = pi/(2D)*g_a*g_b*<s_a, s_b>
//   corr clamped to [-|ip_mse|, |ip_mse|]; ip = (ip_mse + corr)*n_a*n_b;
//   d = max(0, n_a^2 + n_b^2 - 2*ip).
// The a side is prepared once per code (k_prepare_sym, the GPU form of
// prepareSymmetricQuery :498-532): since <s_a, H r_b> = <H^T s_a, r_b> and
// r_b = sigma_b * c[idx_b], every term is a contraction of a per-a vector
// with the b code's centroids or signs:
//   ip_mse = sigma_b * <r_a, c_b>
//   corr   = sigma_b * <G_a, c_b> + scale*g_b * <h'_a, s_b>
// with G_a = scale*g_a*H^T s_a and h'_a = H r_a + k_a*s_a (full) or
// G_a = 0, h'_a = k_a*s_a (light), k_a = pi*g_a / (2*D*scale). H^T s_a is a
// WHT of a +-1 vector (exact integers) scaled and sign-flipped. The
// reassociation moves results by rounding only (the correction terms carry
// factors ~1e-2), well inside the 1e-5 bar.
#include "tq_common.h"
#include "tq_symfull.h"

namespace tq {
namespace {

constexpr float kPi = 3.14159265358979323846f;

// One workgroup per code of `codes`; writes row a of `sp` (stride
// sym_prep_stride(D)): [r (D) | G (D) | h' (D) | n^2 | n | 0 | 0].
__global__ __launch_bounds__(256) void k_prepare_sym(SpaceDev s, const uint8_t *__restrict__ codes,
                                                     long m, int light, float *__restrict__ sp) {
  extern __shared__ float lds[];
  float *w_r = lds;         // r_a, then H r_a
  float *w_s = lds + s.D;   // s_a, then H^T s_a
  const int D = s.D;
  const bool al = (s.cs & 3) == 0;
  for (long a = blockIdx.x; a < m; a += gridDim.x) {
    const uint8_t *cp = codes + a * (long)s.cs;
    const float nrm = load_f32(cp + s.packed, al);
    const float gamma = load_f32(cp + s.packed + 4, al);
    const float sigma = load_f32(cp + s.packed + 8, al);
    float *row = sp + a * sym_prep_stride(D);
    __syncthreads();  // the previous code is done with the workspace
    for (int i = threadIdx.x; i < D; i += blockDim.x) {
      uint32_t u;
      if (s.nib) {
        const uint32_t byte = cp[i >> 1];
        u = (i & 1) ? (byte >> 4) : (byte & 0xFu);
      } else {
        u = cp[i];
      }
      const float r = s.cent[u >> 1] * sigma;
      row[i] = r;
      w_r[i] = r * s.qjl[i];  // randomizedHadamard: sign flip first
      w_s[i] = (u & 1u) ? 1.0f : -1.0f;
    }
    __syncthreads();
    // Butterflies in ascending step order (turbo_quant.h:46-60) on both.
    for (int step = 1; step < D; step <<= 1) {
      for (int p = threadIdx.x; p < D / 2; p += blockDim.x) {
        const int i = (p / step) * 2 * step + (p % step);
        const float a0 = w_r[i], b0 = w_r[i + step];
        w_r[i] = a0 + b0;
        w_r[i + step] = a0 - b0;
        const float a1 = w_s[i], b1 = w_s[i + step];
        w_s[i] = a1 + b1;
        w_s[i + step] = a1 - b1;
      }
      __syncthreads();
    }
    // pi/(2d) as the reference forms it (:819-820), folded with 1/scale.
    const float pi_over_2d = kPi / (2.0f * (float)D);
    const float k_a = pi_over_2d * gamma / s.scale;
    const float sg = s.scale * gamma;
    for (int i = threadIdx.x; i < D; i += blockDim.x) {
      uint32_t u;
      if (s.nib) {
        const uint32_t byte = cp[i >> 1];
        u = (i & 1) ? (byte >> 4) : (byte & 0xFu);
      } else {
        u = cp[i];
      }
      const float sa = (u & 1u) ? 1.0f : -1.0f;
      const float h = w_r[i] * s.inv_sqrt_D;                 // (H r_a)_i
      const float hts = (w_s[i] * s.inv_sqrt_D) * s.qjl[i];  // (H^T s_a)_i
      row[D + i] = light ? 0.0f : sg * hts;
      row[2 * D + i] = light ? k_a * sa : h + k_a * sa;
    }
    if (threadIdx.x == 0) {
      row[3 * D] = nrm * nrm;
      row[3 * D + 1] = nrm;
      row[3 * D + 2] = 0.0f;
      row[3 * D + 3] = 0.0f;
    }
  }
}

// One thread per b code; the prepared a row is staged in LDS and read as a
// broadcast. Each thread walks its code in words of 8 coordinates.
template <bool NIB>
__global__ __launch_bounds__(256) void k_sym_qjl(SpaceDev s, const float *__restrict__ sp, long m,
                                                 const uint8_t *__restrict__ codes, long n,
                                                 float *__restrict__ out) {
  extern __shared__ float lds[];
  float *l_r = lds, *l_g = lds + s.D, *l_h = lds + 2 * s.D;
  float *l_cent = lds + 3 * s.D;
  const int D = s.D;
  for (int i = threadIdx.x; i < s.L; i += blockDim.x) l_cent[i] = s.cent[i];
  const bool al = (s.cs & 3) == 0;
  for (long qi = blockIdx.y; qi < m; qi += gridDim.y) {
    const float *row = sp + qi * sym_prep_stride(D);
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * D; i += blockDim.x) lds[i] = row[i];
    __syncthreads();
    const float qns = row[3 * D];
    const float qn = row[3 * D + 1];
    float *orow = out + qi * n;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (long)gridDim.x * blockDim.x) {
      const uint8_t *cp = codes + j * (long)s.cs;
      float am = 0.0f, ag = 0.0f, aq = 0.0f;
      if (D < 8) {  // D = 4: coordinate by coordinate
        for (int i = 0; i < D; ++i) {
          const uint32_t uu = NIB ? ((cp[i >> 1] >> (4 * (i & 1))) & 0xFu) : cp[i];
          const float c = l_cent[uu >> 1];
          am = fmaf(l_r[i], c, am);
          ag = fmaf(l_g[i], c, ag);
          aq += (uu & 1u) ? l_h[i] : -l_h[i];
        }
      }
      // Eight partial sums per term (one per coordinate residue mod 8),
      // folded pairwise at the end: same-sign sums over large D (self-
      // distances) keep ~1/8 of a single chain's rounding.
      float pm[8], pg[8], pq[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) pm[k] = pg[k] = pq[k] = 0.0f;
      for (int w = 0; w < D / 8; ++w) {
        uint32_t u[8];
        if constexpr (NIB) {
          const uint32_t word = (uint32_t)cp[4 * w] | ((uint32_t)cp[4 * w + 1] << 8) |
                                ((uint32_t)cp[4 * w + 2] << 16) | ((uint32_t)cp[4 * w + 3] << 24);
#pragma unroll
          for (int k = 0; k < 8; ++k) u[k] = (word >> (4 * k)) & 0xFu;
        } else {
#pragma unroll
          for (int k = 0; k < 8; ++k) u[k] = cp[8 * w + k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int i = 8 * w + k;
          const float c = l_cent[u[k] >> 1];
          pm[k] = fmaf(l_r[i], c, pm[k]);
          pg[k] = fmaf(l_g[i], c, pg[k]);
          pq[k] += (u[k] & 1u) ? l_h[i] : -l_h[i];
        }
      }
#pragma unroll
      for (int h = 4; h >= 1; h >>= 1)
#pragma unroll
        for (int k = 0; k < h; ++k) {
          pm[k] += pm[k + h];
          pg[k] += pg[k + h];
          pq[k] += pq[k + h];
        }
      am += pm[0];
      ag += pg[0];
      aq += pq[0];
      const float xn = load_f32(cp + s.packed, al);
      const float gamma = load_f32(cp + s.packed + 4, al);
      const float sigma = load_f32(cp + s.packed + 8, al);
      const float ip_mse = am * sigma;
      float corr = ag * sigma + s.scale * gamma * aq;
      const float mc = fabsf(ip_mse);
      corr = fmaxf(-mc, fminf(mc, corr));
      const float ip = (ip_mse + corr) * qn * xn;
      const float d = qns + xn * xn - 2.0f * ip;
      orow[j] = d > 0.0f ? d : 0.0f;
    }
  }
}

}  // namespace

bool sym_qjl_supported(const SpaceDev &s) { return s.D >= 4 && s.D <= 8192; }

hipError_t launch_prepare_sym(const SpaceDev &s, const uint8_t *codes, long m, bool light,
                              float *sp, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  if (!sym_qjl_supported(s)) return hipErrorNotSupported;
  const long blocks = m < 65535 ? m : 65535;
  hipLaunchKernelGGL(k_prepare_sym, dim3((unsigned)blocks), dim3(256),
                     2 * sizeof(float) * (size_t)s.D, st, s, codes, m, light ? 1 : 0, sp);
  return hipGetLastError();
}

hipError_t launch_sym_qjl(const SpaceDev &s, const float *sp, long m, const uint8_t *codes,
                          long n, float *out, hipStream_t st) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (!sym_qjl_supported(s)) return hipErrorNotSupported;
  long bx = (n + 255) / 256;
  if (bx > 1024) bx = 1024;
  const long by = m < 65535 ? m : 65535;
  const size_t lds = sizeof(float) * (3 * (size_t)s.D + 256);
  auto kern = s.nib ? k_sym_qjl<true> : k_sym_qjl<false>;
  if (lds > 65536) {  // padded dim 8192: more than the default dynamic LDS
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)bx, (unsigned)by), dim3(256), lds, st, s, sp, m, codes,
                     n, out);
  return hipGetLastError();
}

}  // namespace tq
