// tq_scan.hip -- plain fp32 asymmetric / symmetric distance scans for gfx950.
//
// Per (prepared query, code) pair this evaluates the reference estimator
// (distSearchScalar / distSearchScalarB4, include/turboquant/space_turbo_quant.h
// :659-686 and :951-983):
//   ip_mse = sigma * sum_i q_rot[i] * c[idx_i]
//   dot    = sum_i s_q[i] * (qjl_i ? +1 : -1)
//   ip     = (ip_mse + scale*gamma*dot) * ||x|| * ||q||
//   d      = max(0, q_norm_sq + ||x||^2 - 2*ip)
// in fp32. The symmetric MSE-only estimator (distBuildScalar(B4), :688-714,
// :985-1019) is the same expression with the query side replaced by a
// decoded code (q_rot[i] = c[a_i]*sigma_a, ||q|| = ||a||, no QJL term); the
// prepared form of a code is written by k_prepare_from_codes.
//
// Layout: a code is owned by LPC = D/CPL lanes, each holding CPL coordinates
// of the prepared query in registers for the whole scan; the lanes stream
// their CPL/2 (4-bit) or CPL (8-bit) code bytes with 4-byte loads (one
// wavefront-wide load covers whole codes contiguously), decode against a
// centroid table in LDS, and fold their partial sums across the LPC lanes.
#include "tq_common.h"

namespace tq {
namespace {

template <int CPL, bool NIB, bool SYM>
__global__ __launch_bounds__(256) void k_scan(SpaceDev s, const float *__restrict__ prep,
                                              long m, const uint8_t *__restrict__ codes,
                                              long n, float *__restrict__ out) {
  __shared__ float l_cent[256];
  for (int i = threadIdx.x; i < s.L; i += blockDim.x) l_cent[i] = s.cent[i];
  __syncthreads();

  const int LPC = s.D / CPL;  // lanes per code (power of two, <= 64)
  const int CPW = 64 / LPC;   // codes per wavefront pass
  const int lane = threadIdx.x & 63;
  const int sl = lane & (LPC - 1);
  const int cslot = lane / LPC;
  const long pstride = prep_stride(s.D);

  for (long qi = blockIdx.y; qi < m; qi += gridDim.y) {
    const float *pq = prep + qi * pstride;
    float qv[CPL], sv[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      qv[c] = pq[sl * CPL + c];
      sv[c] = SYM ? 0.0f : pq[s.D + sl * CPL + c];
    }
    const float qns = pq[2 * s.D];
    const float qn = pq[2 * s.D + 1];
    float *orow = out + qi * n;

    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    for (long base = wave * CPW; base < n; base += nwaves * CPW) {
      const long code = base + cslot;
      const bool act = code < n;
      const uint8_t *cp = codes + (act ? code : 0L) * (long)s.cs;
      float am = 0.0f, aq = 0.0f;
      if constexpr (NIB) {
        if constexpr (CPL >= 8) {
#pragma unroll
          for (int w = 0; w < CPL / 8; ++w) {
            uint32_t word = *reinterpret_cast<const uint32_t *>(cp + sl * (CPL / 2) + 4 * w);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              uint32_t u = (word >> (4 * k)) & 0xFu;
              am = fmaf(qv[8 * w + k], l_cent[u >> 1], am);
              if constexpr (!SYM) aq += (u & 1u) ? sv[8 * w + k] : -sv[8 * w + k];
            }
          }
        } else {
          uint32_t half = *reinterpret_cast<const uint16_t *>(cp + sl * (CPL / 2));
#pragma unroll
          for (int k = 0; k < CPL; ++k) {
            uint32_t u = (half >> (4 * k)) & 0xFu;
            am = fmaf(qv[k], l_cent[u >> 1], am);
            if constexpr (!SYM) aq += (u & 1u) ? sv[k] : -sv[k];
          }
        }
      } else {
#pragma unroll
        for (int w = 0; w < CPL / 4; ++w) {
          uint32_t word = *reinterpret_cast<const uint32_t *>(cp + sl * CPL + 4 * w);
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            uint32_t u = (word >> (8 * k)) & 0xFFu;
            am = fmaf(qv[4 * w + k], l_cent[u >> 1], am);
            if constexpr (!SYM) aq += (u & 1u) ? sv[4 * w + k] : -sv[4 * w + k];
          }
        }
      }
      for (int o = 1; o < LPC; o <<= 1) {
        am += __shfl_xor(am, o);
        if constexpr (!SYM) aq += __shfl_xor(aq, o);
      }
      if (act && sl == 0) {
        const bool al = (s.cs & 3) == 0;
        const float xn = load_f32(cp + s.packed, al);
        const float gamma = load_f32(cp + s.packed + 4, al);
        const float sigma = load_f32(cp + s.packed + 8, al);
        const float ip_mse = am * sigma;
        const float corr = SYM ? 0.0f : s.scale * gamma * aq;
        const float ip = (ip_mse + corr) * xn * qn;
        const float d = qns + xn * xn - 2.0f * ip;
        orow[code] = d > 0.0f ? d : 0.0f;
      }
    }
  }
}

// Prepared form of codes for the symmetric scan: q_rot[i] = c[idx_i]*sigma,
// q_norm_sq = norm*norm, q_norm = norm (the query side of distBuildScalar,
// space_turbo_quant.h:699-712).
__global__ __launch_bounds__(256) void k_prepare_from_codes(SpaceDev s,
                                                            const uint8_t *__restrict__ codes,
                                                            long m, float *__restrict__ prep) {
  const long stride = prep_stride(s.D);
  const long total = m * (long)s.D;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long a = e / s.D;
    const int i = (int)(e - a * s.D);
    const uint8_t *cp = codes + a * (long)s.cs;
    const bool al = (s.cs & 3) == 0;
    const float sigma = load_f32(cp + s.packed + 8, al);
    uint32_t u;
    if (s.nib) {
      uint32_t b = cp[i >> 1];
      u = (i & 1) ? (b >> 4) : (b & 0xFu);
    } else {
      u = cp[i];
    }
    float *pq = prep + a * stride;
    pq[i] = s.cent[u >> 1] * sigma;
    pq[s.D + i] = 0.0f;
    if (i == 0) {
      const float nrm = load_f32(cp + s.packed, al);
      pq[2 * s.D] = nrm * nrm;
      pq[2 * s.D + 1] = nrm;
      pq[2 * s.D + 2] = 0.0f;
      pq[2 * s.D + 3] = 0.0f;
    }
  }
}

template <int CPL>
hipError_t launch_cpl(const SpaceDev &s, const float *prep, long m, const uint8_t *codes,
                      long n, float *out, bool sym, hipStream_t st) {
  const int LPC = s.D / CPL;
  const long cpb = 4L * (64 / LPC);  // codes per block pass (4 wavefronts)
  long bx = (n + cpb - 1) / cpb;
  const long cap = m >= 64 ? 512 : 2048;
  if (bx > cap) bx = cap;
  if (bx < 1) bx = 1;
  long by = m < 65535 ? m : 65535;
  dim3 grid((unsigned)bx, (unsigned)by);
  if (s.nib) {
    if (sym)
      hipLaunchKernelGGL((k_scan<CPL, true, true>), grid, dim3(256), 0, st, s, prep, m, codes, n, out);
    else
      hipLaunchKernelGGL((k_scan<CPL, true, false>), grid, dim3(256), 0, st, s, prep, m, codes, n, out);
  } else {
    if (sym)
      hipLaunchKernelGGL((k_scan<CPL, false, true>), grid, dim3(256), 0, st, s, prep, m, codes, n, out);
    else
      hipLaunchKernelGGL((k_scan<CPL, false, false>), grid, dim3(256), 0, st, s, prep, m, codes, n, out);
  }
  return hipGetLastError();
}

}  // namespace

bool scan_supported(int D) { return D >= 4 && D <= 8192; }

hipError_t launch_scan_fp32(const SpaceDev &s, const float *prep, long m, const uint8_t *codes,
                            long n, float *out, bool symmetric, hipStream_t st, void *ws) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (lut_supported(s)) return launch_scan_lut(s, prep, m, codes, n, out, symmetric, st, ws);
  switch (s.D) {
    case 4: return launch_cpl<4>(s, prep, m, codes, n, out, symmetric, st);
    case 8: return launch_cpl<8>(s, prep, m, codes, n, out, symmetric, st);
    case 16:
    case 32:
    case 64:
    case 128:
    case 256:
    case 512:
    case 1024: return launch_cpl<16>(s, prep, m, codes, n, out, symmetric, st);
    case 2048: return launch_cpl<32>(s, prep, m, codes, n, out, symmetric, st);
    case 4096: return launch_cpl<64>(s, prep, m, codes, n, out, symmetric, st);
    case 8192: return launch_cpl<128>(s, prep, m, codes, n, out, symmetric, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_prepare_from_codes(const SpaceDev &s, const uint8_t *codes, long m, float *prep,
                                     hipStream_t st) {
  if (m <= 0) return hipSuccess;
  long total = m * (long)s.D;
  long blocks = (total + 255) / 256;
  if (blocks > 65535) blocks = 65535;
  hipLaunchKernelGGL(k_prepare_from_codes, dim3((unsigned)blocks), dim3(256), 0, st, s, codes, m,
                     prep);
  return hipGetLastError();
}

}  // namespace tq
