// tq_lut.hip -- HBM-streaming fp32 scan for 4-bit (nibble-packed) codes.
//
// Same estimator as tq_scan.hip (distSearchScalarB4,
// include/turboquant/space_turbo_quant.h:951-983; distBuildScalarB4 :985-1019
// for the symmetric form), arranged so that a lane spends about two VALU
// operations per coordinate and the scan runs at HBM speed:
//
//  * Lanes own whole codes (64 codes per wavefront pass), so every lane of a
//    wavefront works on the same coordinates at the same time and the query
//    side is wavefront-uniform.
//  * MSE term through a pair table in LDS: for byte position p (coordinates
//    2p, 2p+1) the 64 entries T_p[lo + 8*hi] = q[2p]*c[lo] + q[2p+1]*c[hi]
//    cover every pair of 3-bit centroid indices. One lookup and one add per
//    code byte. The byte offset 4*(lo + 8*hi) of an entry is obtained for all
//    four bytes of a code word at once as (w & 0xEE..) + (w & 0x0E..) (no
//    carries between bytes), and one v_perm per byte splices it under the
//    table's base. All lanes read the same 64-entry table, which covers each
//    LDS bank once: no bank conflicts.
//  * QJL term as an exact integer dot product: s_q is rounded to a 15-bit
//    fixed-point integer (scale 2^e per query) split into four signed 4-bit
//    digits, and sum_{qjl bit set} S_i is four v_dot8 (i4 x i4 -> i32) per
//    code word against the word masked to its sign bits; then
//    dot = (2*T - sum_i S_i) * 2^-e. The rounding of s_q to 15 bits moves
//    the distance by ~1e-7 of its scale (the QJL term carries a factor
//    scale*gamma ~ 1e-2).
// The plain fp32 path and its 1e-5 bar are unchanged: the per-code epilogue
// is the reference's fp32 expression.
#include "tq_common.h"

namespace tq {
namespace {

// Per query: [e, S_int, 0, 0 | D/2 words]; word 4*wi + j holds digit j of the
// fixed-point s_q of coordinates 8*wi .. 8*wi+7 (nibble v = coordinate 8wi+v).
__host__ __device__ inline long qlimb_stride(int D) { return 4 + D / 2; }

__global__ __launch_bounds__(256) void k_qjl_limbs(SpaceDev s, const float *__restrict__ prep,
                                                   long m, int32_t *__restrict__ ql) {
  __shared__ float red_f[256];
  __shared__ int red_i[256];
  const long pstride = prep_stride(s.D);
  const long lstride = qlimb_stride(s.D);
  for (long q = blockIdx.x; q < m; q += gridDim.x) {
    const float *sq = prep + q * pstride + s.D;
    float mx = 0.0f;
    for (int i = threadIdx.x; i < s.D; i += blockDim.x) mx = fmaxf(mx, fabsf(sq[i]));
    red_f[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red_f[threadIdx.x] = fmaxf(red_f[threadIdx.x], red_f[threadIdx.x + o]);
      __syncthreads();
    }
    mx = red_f[0];
    int ex = 0;
    if (mx > 0.0f) frexpf(mx, &ex);  // mx < 2^ex
    const int e = mx > 0.0f ? 14 - ex : 0;  // |s| * 2^e < 2^14
    int32_t *out = ql + q * lstride;
    int ssum = 0;
    for (int wi = threadIdx.x; wi < s.D / 8; wi += blockDim.x) {
      uint32_t L[4] = {0u, 0u, 0u, 0u};
      for (int v = 0; v < 8; ++v) {
        int S = (int)rintf(ldexpf(sq[8 * wi + v], e));
        ssum += S;
        for (int j = 0; j < 4; ++j) {
          const int d = ((S + 8) & 15) - 8;
          L[j] |= (uint32_t)(d & 15) << (4 * v);
          S = (S - d) >> 4;
        }
      }
      for (int j = 0; j < 4; ++j) out[4 + 4 * wi + j] = (int32_t)L[j];
    }
    red_i[threadIdx.x] = ssum;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red_i[threadIdx.x] += red_i[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      out[0] = e;
      out[1] = red_i[0];
      out[2] = 0;
      out[3] = 0;
    }
    __syncthreads();
  }
}

// Measured on MI355X at dim 768 / 4-bit over 10M codes: 512 threads, one
// code per lane per block and one block of prefetch beat 1024 threads
// (register spills / L1 thrash), two codes per lane and deeper prefetch.
constexpr int kLutThreads = 512;
constexpr int kLutWaves = kLutThreads / 64;
constexpr int kLutCodesPerLane = 1;
constexpr int kLutPrefetch = 1;  // blocks in flight ahead of the one computed

// Work split of one block: SL coordinate slices x G code groups = 8
// wavefronts. Wavefront w takes slice w % SL (WPW consecutive 32-bit words of
// every code) of the CPL*64 codes of group w / SL; lane l owns codes
// (group*CPL + c)*64 + l. One thread per code reduces the slices. The 8 wavefronts together read each block's
// codes as one contiguous run of bytes, so every cache line is used whole
// while it is still in L2.
template <int D>
struct LutShape {
  static constexpr int NW = D / 8;                      // words per code
  static constexpr int SL = NW / 4 < 8 ? NW / 4 : 8;    // slices (>= 16 bytes each)
  static constexpr int G = kLutWaves / SL;              // code groups
  static constexpr int WPW = NW / SL;                   // words per slice
  static constexpr int V4 = WPW / 4;                    // 16-byte loads per slice
  static constexpr int CPL = G * kLutCodesPerLane * 64 <= kLutThreads ? kLutCodesPerLane : 1;  // codes per lane
  static constexpr int CODES = G * CPL * 64;            // codes per block (<= 512)
};

template <bool SYM, int D>
__global__ __launch_bounds__(kLutThreads) void k_lut_nib(SpaceDev s, const float *__restrict__ prep,
                                                         const int32_t *__restrict__ ql, long m,
                                                         const uint8_t *__restrict__ codes, long n,
                                                         float *__restrict__ out) {
  using Sh = LutShape<D>;
  constexpr int SL = Sh::SL, WPW = Sh::WPW, V4 = Sh::V4, CODES = Sh::CODES, kLutCPL = Sh::CPL;
  static_assert(CODES <= kLutThreads, "one reducing thread per code");
  static_assert(kLutWaves % SL == 0 && Sh::G * SL == kLutWaves, "every wavefront owns a (slice, group)");
  // (D/2) tables of 64 floats; fixed-size arrays keep LDS addresses
  // compile-time constants. part_*: per-slice partial sums, double-buffered.
  __shared__ float tab[(D / 2) * 64];
  __shared__ float part_m[2][SL][CODES];
  __shared__ int part_q[2][SL][CODES];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int sl = wave % SL, grp = wave / SL;
  const long pstride = prep_stride(s.D);
  const long lstride = qlimb_stride(s.D);
  const long nblocks = (n + CODES - 1) / CODES;
  float cent[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) cent[k] = s.cent[k < s.L ? k : 0];

  auto load_block = [&](long b, uint4 (&r)[kLutCPL][V4]) {
#pragma unroll
    for (int c = 0; c < kLutCPL; ++c) {
      const long code = b * CODES + (grp * kLutCPL + c) * 64 + lane;
      const uint32_t *src = reinterpret_cast<const uint32_t *>(
          codes + (code < n ? code : 0L) * (long)s.cs + sl * WPW * 4);
#pragma unroll
      for (int v = 0; v < V4; ++v)
        r[c][v] = make_uint4(src[4 * v], src[4 * v + 1], src[4 * v + 2], src[4 * v + 3]);
    }
  };

  for (long qi = 0; qi < m; ++qi) {
    const float *pq = prep + qi * pstride;
    __syncthreads();  // the previous query is done with the tables
    for (int e = tid; e < (D / 2) * 64; e += kLutThreads) {
      const int p = e >> 6, lo = e & 7, hi = (e >> 3) & 7;
      tab[e] = fmaf(pq[2 * p + 1], cent[hi], pq[2 * p] * cent[lo]);
    }
    __syncthreads();
    const float qns = pq[2 * s.D];
    const float qn = pq[2 * s.D + 1];
    const int32_t *L = ql + qi * lstride;
    const int qe = SYM ? 0 : L[0];
    const int qsum = SYM ? 0 : L[1];
    float *orow = out + qi * n;

    // Blocks b, b + grid, ... stream through a ring of kLutPrefetch register
    // buffers: the loads of a block are issued kLutPrefetch blocks before use.
    uint4 ring[kLutPrefetch][kLutCPL][V4];
    long b = blockIdx.x;
#pragma unroll
    for (int d = 0; d < kLutPrefetch; ++d)
      if (b + (long)d * gridDim.x < nblocks) load_block(b + (long)d * gridDim.x, ring[d]);
    int buf = 0;
    for (; b < nblocks; b += gridDim.x, buf ^= 1) {
      uint4 cur[kLutCPL][V4];
#pragma unroll
      for (int c = 0; c < kLutCPL; ++c)
#pragma unroll
        for (int v = 0; v < V4; ++v) cur[c][v] = ring[0][c][v];
#pragma unroll
      for (int d = 0; d + 1 < kLutPrefetch; ++d)
#pragma unroll
        for (int c = 0; c < kLutCPL; ++c)
#pragma unroll
          for (int v = 0; v < V4; ++v) ring[d][c][v] = ring[d + 1][c][v];
      if (b + (long)kLutPrefetch * gridDim.x < nblocks)
        load_block(b + (long)kLutPrefetch * gridDim.x, ring[kLutPrefetch - 1]);
      // The reducing thread's trailer, fetched early.
      const long rcode = b * CODES + tid;
      const bool ract = tid < CODES && rcode < n;
      float xn = 0.0f, gamma = 0.0f, sigma = 0.0f;
      if (ract) {
        const float *meta = reinterpret_cast<const float *>(codes + rcode * (long)s.cs + s.packed);
        xn = meta[0];
        gamma = meta[1];
        sigma = meta[2];
      }
#pragma unroll
      for (int c = 0; c < kLutCPL; ++c) {
        float am0 = 0.0f, am1 = 0.0f;
        int aq0 = 0, aq1 = 0, aq2 = 0, aq3 = 0;
#pragma unroll
        for (int k2 = 0; k2 < WPW; ++k2) {
          const uint4 q4 = cur[c][k2 >> 2];
          const uint32_t w = (k2 & 3) == 0 ? q4.x : (k2 & 3) == 1 ? q4.y : (k2 & 3) == 2 ? q4.z : q4.w;
          const int wi = sl * WPW + k2;
          const uint32_t m2 = (w & 0xEEEEEEEEu) + (w & 0x0E0E0E0Eu);
          const uint32_t tb = (uint32_t)(wi * 1024);  // byte offset of table 4*wi
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const uint32_t addr =
                __builtin_amdgcn_perm(tb + 256u * k, m2, 0x07060500u | (uint32_t)k);
            const float v = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(tab) + addr);
            if (k & 1)
              am1 += v;
            else
              am0 += v;
          }
          if constexpr (!SYM) {
            const int t = (int)(w & 0x11111111u);
            aq0 = __builtin_amdgcn_sdot8(t, L[4 + 4 * wi + 0], aq0, false);
            aq1 = __builtin_amdgcn_sdot8(t, L[4 + 4 * wi + 1], aq1, false);
            aq2 = __builtin_amdgcn_sdot8(t, L[4 + 4 * wi + 2], aq2, false);
            aq3 = __builtin_amdgcn_sdot8(t, L[4 + 4 * wi + 3], aq3, false);
          }
        }
        const int slot = (grp * kLutCPL + c) * 64 + lane;
        part_m[buf][sl][slot] = am0 + am1;
        part_q[buf][sl][slot] = aq0 + 16 * aq1 + 256 * aq2 + 4096 * aq3;
      }
      __syncthreads();
      if (ract) {
        float am = 0.0f;
        int T = 0;
#pragma unroll
        for (int k = 0; k < SL; ++k) {
          am += part_m[buf][k][tid];
          T += part_q[buf][k][tid];
        }
        const float ip_mse = am * sigma;
        float corr = 0.0f;
        if constexpr (!SYM) {
          const float dot = ldexpf((float)(2 * T - qsum), -qe);
          corr = s.scale * gamma * dot;
        }
        const float ip = (ip_mse + corr) * xn * qn;
        const float d = qns + xn * xn - 2.0f * ip;
        orow[rcode] = d > 0.0f ? d : 0.0f;
      }
    }
  }
}

int cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  return cus;
}

}  // namespace

bool lut_supported(const SpaceDev &s) { return s.nib && s.D >= 32 && s.D <= 1024; }

namespace {

template <bool SYM, int D>
hipError_t launch_lut_d(const SpaceDev &s, const float *prep, const int32_t *ql, long m,
                        const uint8_t *codes, long n, float *out, hipStream_t st) {
  constexpr size_t lds =
      sizeof(float) * 64 * (D / 2) + 2 * 2 * 4 * LutShape<D>::SL * LutShape<D>::CODES;
  const long per_cu = (long)((160 * 1024) / lds) < 4 ? (long)((160 * 1024) / lds) : 4;
  long blocks = (long)cu_count() * (per_cu > 0 ? per_cu : 1);
  const long need = (n + LutShape<D>::CODES - 1) / LutShape<D>::CODES;
  if (blocks > need) blocks = need;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((k_lut_nib<SYM, D>), dim3((unsigned)blocks), dim3(kLutThreads), 0, st, s,
                     prep, ql, m, codes, n, out);
  return hipGetLastError();
}

template <bool SYM>
hipError_t launch_lut(const SpaceDev &s, const float *prep, const int32_t *ql, long m,
                      const uint8_t *codes, long n, float *out, hipStream_t st) {
  switch (s.D) {
    case 32: return launch_lut_d<SYM, 32>(s, prep, ql, m, codes, n, out, st);
    case 64: return launch_lut_d<SYM, 64>(s, prep, ql, m, codes, n, out, st);
    case 128: return launch_lut_d<SYM, 128>(s, prep, ql, m, codes, n, out, st);
    case 256: return launch_lut_d<SYM, 256>(s, prep, ql, m, codes, n, out, st);
    case 512: return launch_lut_d<SYM, 512>(s, prep, ql, m, codes, n, out, st);
    case 1024: return launch_lut_d<SYM, 1024>(s, prep, ql, m, codes, n, out, st);
    default: return hipErrorNotSupported;
  }
}

}  // namespace

size_t scan_fp32_ws_bytes(const SpaceDev &s, long m) {
  return lut_supported(s) ? sizeof(int32_t) * (size_t)m * qlimb_stride(s.D) : 0;
}

hipError_t launch_scan_lut(const SpaceDev &s, const float *prep, long m, const uint8_t *codes,
                           long n, float *out, bool sym, hipStream_t st, void *ws) {
  if (!lut_supported(s)) return hipErrorNotSupported;
  if (m <= 0 || n <= 0) return hipSuccess;
  int32_t *ql = nullptr;
  bool owned = false;
  hipError_t e = hipSuccess;
  if (!sym) {
    if (ws) {
      ql = static_cast<int32_t *>(ws);
    } else {
      e = hipMallocAsync(reinterpret_cast<void **>(&ql),
                         sizeof(int32_t) * (size_t)m * qlimb_stride(s.D), st);
      if (e != hipSuccess) return e;
      owned = true;
    }
    const long g = m < 65535 ? m : 65535;
    hipLaunchKernelGGL(k_qjl_limbs, dim3((unsigned)g), dim3(256), 0, st, s, prep, m, ql);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = sym ? launch_lut<true>(s, prep, ql, m, codes, n, out, st)
            : launch_lut<false>(s, prep, ql, m, codes, n, out, st);
  if (owned) {
    hipError_t f = hipFreeAsync(ql, st);
    if (e == hipSuccess) e = f;
  }
  return e;
}

}  // namespace tq
