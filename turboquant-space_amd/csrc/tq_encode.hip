// tq_encode.hip -- TurboQuant encoder and query preparation for gfx950.
//
// Follows encodeVector (include/turboquant/space_turbo_quant.h:339-396) and
// prepareQuery (:439-460). Codes must be byte-identical to the pinned
// reference build, so every floating-point step reproduces that build:
//   * this file is compiled with -ffp-contract=off; the one multiply-add the
//     reference build fuses (`rotated - centroid*sigma`, :366) is an explicit
//     fmaf, everything else is a single correctly rounded IEEE op;
//   * the three `s += x*x` reductions (:342-344, :353-355, :369-371) follow
//     the order that build vectorises them in: eight FMA chains over the
//     residues mod 8, folded ((a0+a4)+(a2+a6))+((a1+a5)+(a3+a7)), then a
//     4-wide block of plain products folded (m0+m2)+(m1+m3), then scalar FMA;
//   * the Walsh-Hadamard butterflies run in ascending step order
//     (include/turboquant/turbo_quant.h:46-60) followed by one multiply by
//     fl(1/sqrt(D)) (:68-71);
//   * division and sqrt are HIP's correctly rounded fp32 forms.
//
// Layout: one vector is owned by G lanes (G = 8, 16, 32 or 64) of a
// wavefront; lane (h, j) (j = lane & 7, h = (lane >> 3) & (G/8 - 1)) holds
// coordinates i = 8*(h*TP + t) + j for t < TP in registers, D = G*TP. Each
// lane therefore owns whole FMA chains of one residue class mod 8, which is
// what makes the reference's reduction order cheap to reproduce: the chain
// runs in registers, the 8-way fold is three cross-lane adds. Butterfly
// steps 1, 2, 4 cross lanes j, steps 8..8*TP/2 stay in registers, larger
// steps cross lanes h (DPP within rows of 16 lanes, shuffles beyond).
#include "tq_common.h"

namespace tq {
namespace {

// What k_encode writes: codes, prepared queries, or rotated coordinates.
constexpr int kModeEncode = 0, kModePrepare = 1, kModeRotated = 2;

// Value of `v` in another lane through DPP (no LDS): CTRL 0xB1 / 0x4E =
// quad_perm lane^1 / lane^2; 0x104 = row_shl:4 (lane+4 within a row of 16);
// 0x114 = row_shr:4 (lane-4); 0x128 = row_ror:8 (lane^8 within a row);
// 0x141 = row_half_mirror (7-i within each 8 lanes). Lanes whose source is
// outside their row read 0.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

// One butterfly step across lanes lane and lane^M: the lane with the lower
// index keeps a + b, the other a - b (a = lower lane's value). Steps 1, 2, 4
// and 8 stay inside a row of 16 lanes and go through DPP; wider ones use a
// cross-lane shuffle.
template <int TP, int M>
__device__ __forceinline__ void lane_butterfly(float (&r)[TP], int lane) {
  const bool hi = (lane & M) != 0;
#pragma unroll
  for (int t = 0; t < TP; ++t) {
    float plo, phi;  // partner's value as seen by the lower / upper lane
    if constexpr (M == 1) {
      plo = phi = dpp_f<0xB1>(r[t]);
    } else if constexpr (M == 2) {
      plo = phi = dpp_f<0x4E>(r[t]);
    } else if constexpr (M == 4) {
      plo = dpp_f<0x104>(r[t]);
      phi = dpp_f<0x114>(r[t]);
    } else if constexpr (M == 8) {
      plo = phi = dpp_f<0x128>(r[t]);
    } else {
      plo = phi = __shfl_xor(r[t], M);
    }
    r[t] = hi ? (phi - r[t]) : (r[t] + plo);
  }
}

// Unnormalised WHT over the D = G*TP coordinates held by a lane group, in
// ascending step order.
template <int G, int TP>
__device__ __forceinline__ void wht_group(float (&r)[TP], int lane) {
  lane_butterfly<TP, 1>(r, lane);
  lane_butterfly<TP, 2>(r, lane);
  lane_butterfly<TP, 4>(r, lane);
#pragma unroll
  for (int st = 1; st < TP; st <<= 1) {
#pragma unroll
    for (int t = 0; t < TP; ++t) {
      if ((t & st) == 0) {
        float a = r[t], b = r[t + st];
        r[t] = a + b;
        r[t + st] = a - b;
      }
    }
  }
  if constexpr (G > 8) lane_butterfly<TP, 8>(r, lane);
  if constexpr (G > 16) lane_butterfly<TP, 16>(r, lane);
  if constexpr (G > 32) lane_butterfly<TP, 32>(r, lane);
}

// Sum of squares over the first `nblocks` blocks of 8 coordinates, in the
// reference build's order (see file header). Chains of residue j run through
// lanes h = 0, 1, ... in turn; every lane of the group returns the sum.
template <int G, int TP>
__device__ __forceinline__ float chain_sumsq(const float (&r)[TP], int nblocks,
                                             int h, int j, int lane, int grp0) {
  constexpr int HG = G / 8;
  float acc = 0.0f;
#pragma unroll
  for (int hh = 0; hh < HG; ++hh) {
    if (h == hh) {
#pragma unroll
      for (int t = 0; t < TP; ++t)
        if (hh * TP + t < nblocks) acc = fmaf(r[t], r[t], acc);
    }
    if (hh + 1 < HG) {
      float a = __shfl(acc, (lane - 8) & 63);
      if (h == hh + 1) acc = a;
    }
  }
  float a = __shfl(acc, grp0 + (HG - 1) * 8 + j);
  float hs = a + __shfl_xor(a, 4);
  float g = hs + __shfl_xor(hs, 2);
  return g + __shfl_xor(g, 1);
}

// The part of the norm loop the reference build does not vectorise 8-wide:
// a 4-wide block of plain products, then scalar FMA (elements i0 .. n-1).
__device__ __forceinline__ float tail_sumsq(float s, const float *x, int i0, int n) {
  int i = i0;
  if (n - i >= 4) {
    float m0 = x[i] * x[i], m1 = x[i + 1] * x[i + 1];
    float m2 = x[i + 2] * x[i + 2], m3 = x[i + 3] * x[i + 3];
    s = s + ((m0 + m2) + (m1 + m3));
    i += 4;
  }
  for (; i < n; ++i) s = fmaf(x[i], x[i], s);
  return s;
}

// idx = #{k : v > bnd[k]} over 2^mbits - 1 ascending boundaries (quantize,
// space_turbo_quant.h:332-337); NaN compares false everywhere and maps to 0.
__device__ __forceinline__ int quantize_idx(float v, const float *bnd, int mbits) {
  int idx = 0;
  for (int k = mbits - 1; k >= 0; --k) {
    int step = 1 << k;
    if (v > bnd[idx + step - 1]) idx += step;
  }
  return idx;
}

template <int G, int TP, bool NIB, int MODE>
__global__ __launch_bounds__(256) void k_encode(SpaceDev s, const float *__restrict__ X,
                                                long n, uint8_t *__restrict__ codes,
                                                float *__restrict__ prep) {
  constexpr int VPW = 64 / G;
  extern __shared__ float lds[];
  float *l_rot = lds;
  float *l_qjl = lds + s.D;
  float *l_cent = l_qjl + s.D;
  float *l_bnd = l_cent + s.L;
  for (int i = threadIdx.x; i < s.D; i += blockDim.x) {
    l_rot[i] = s.rot[i];
    l_qjl[i] = s.qjl[i];
  }
  for (int i = threadIdx.x; i < s.L; i += blockDim.x) l_cent[i] = s.cent[i];
  for (int i = threadIdx.x; i < s.L - 1; i += blockDim.x) l_bnd[i] = s.bnd[i];
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int j = lane & 7;
  const int h = (lane >> 3) & (G / 8 - 1);
  const int grp0 = lane & ~(G - 1);
  const long vec = ((long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * VPW + lane / G;
  const bool active = vec < n;
  const float *x = X + (active ? vec : 0L) * (long)s.dim;

  float r[TP];
#pragma unroll
  for (int t = 0; t < TP; ++t) {
    int i = 8 * (h * TP + t) + j;
    r[t] = (active && i < s.dim) ? x[i] : 0.0f;
  }

  // ||x||^2 over the input dim. encodeVector (space_turbo_quant.h:342-345)
  // is vectorised by the reference build; prepareQuery (:443-446) adds into
  // the returned struct, which the compiler must assume aliases the query, so
  // it stays one scalar FMA chain (run by the group's first lane).
  constexpr bool PREP = MODE == kModePrepare;
  float nsq;
  if constexpr (PREP) {
    float acc = 0.0f;
    if (active && lane == grp0)
      for (int i = 0; i < s.dim; ++i) acc = fmaf(x[i], x[i], acc);
    nsq = __shfl(acc, grp0);
  } else {
    const int q8 = s.dim >> 3;
    nsq = chain_sumsq<G, TP>(r, q8, h, j, lane, grp0);
    if (active) nsq = tail_sumsq(nsq, x, q8 * 8, s.dim);
  }
  const float norm = sqrtf(nsq);
  // rotatedCoordsBatch (:418-419) tests ||x||^2 and divides by its root;
  // encodeVector and prepareQuery (:345-348, :446-447) test ||x||.
  const float inv_norm = MODE == kModeRotated ? ((nsq > 1e-20f) ? (1.0f / sqrtf(nsq)) : 0.0f)
                                              : ((norm > 1e-10f) ? (1.0f / norm) : 0.0f);

  // rotated = RHT_rot(x / ||x||), zero-padded (:346-350, turbo_quant.h:89-99).
#pragma unroll
  for (int t = 0; t < TP; ++t) {
    int i = 8 * (h * TP + t) + j;
    r[t] = r[t] * inv_norm;
    r[t] = r[t] * l_rot[i];
  }
  wht_group<G, TP>(r, lane);
#pragma unroll
  for (int t = 0; t < TP; ++t) r[t] = r[t] * s.inv_sqrt_D;

  if constexpr (MODE == kModeRotated) {
    // rotatedCoordsBatch (:411-435): the rotated coordinates times
    // sqrt(D / sum r^2) (1 when the sum is below 1e-20), D floats per vector.
    const float var = chain_sumsq<G, TP>(r, G / 8 * TP, h, j, lane, grp0);
    const float inv_sigma = (var > 1e-20f) ? sqrtf((float)s.D / var) : 1.0f;
    float *dst = prep + (active ? vec : 0L) * (long)s.D;
    if (active) {
#pragma unroll
      for (int t = 0; t < TP; ++t) dst[8 * (h * TP + t) + j] = r[t] * inv_sigma;
    }
    return;
  } else if constexpr (PREP) {
    // TurboQuantPreparedQuery: q_rot, then s_q = RHT_qjl(q_rot) (:449-458).
    float *pq = prep + (active ? vec : 0L) * prep_stride(s.D);
    if (active) {
#pragma unroll
      for (int t = 0; t < TP; ++t) pq[8 * (h * TP + t) + j] = r[t];
    }
#pragma unroll
    for (int t = 0; t < TP; ++t) r[t] = r[t] * l_qjl[8 * (h * TP + t) + j];
    wht_group<G, TP>(r, lane);
    if (active) {
#pragma unroll
      for (int t = 0; t < TP; ++t) pq[s.D + 8 * (h * TP + t) + j] = r[t] * s.inv_sqrt_D;
      if (lane == grp0) {
        pq[2 * s.D] = nsq;
        pq[2 * s.D + 1] = norm;
        pq[2 * s.D + 2] = 0.0f;
        pq[2 * s.D + 3] = 0.0f;
      }
    }
    return;
  } else {
    // sigma = RMS of the rotated coordinates, clamped (:352-359).
    const float var = chain_sumsq<G, TP>(r, G / 8 * TP, h, j, lane, grp0);
    float sigma = sqrtf(var / (float)s.D);
    if (sigma < 1e-10f) sigma = 1e-10f;
    const float inv_sigma = 1.0f / sigma;

    uint8_t *out = codes + (active ? vec : 0L) * (long)s.cs;
    if constexpr (NIB && TP >= 8) {
      // Lloyd-Max index and residual (:361-367). At most 8 centroids: the
      // count #{k : v > b_k} is a three-level search over the boundaries
      // padded to 7 with +inf (the padding never compares true), each
      // level's threshold picked by selects from registers.
      float b[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) b[k] = k < s.L - 1 ? l_bnd[k] : __builtin_inff();
      // nw[k] gathers the units of t = 8k + m as nibble m.
      uint32_t nw[TP / 8];
#pragma unroll
      for (int k = 0; k < TP / 8; ++k) nw[k] = 0u;
#pragma unroll
      for (int t = 0; t < TP; ++t) {
        const float v = r[t] * inv_sigma;
        const bool c1 = v > b[3];
        const bool c2 = v > (c1 ? b[5] : b[1]);
        const bool c3 = v > (c1 ? (c2 ? b[6] : b[4]) : (c2 ? b[2] : b[0]));
        const uint32_t q = (c1 ? 4u : 0u) | (c2 ? 2u : 0u) | (c3 ? 1u : 0u);
        nw[t >> 3] |= q << (4 * (t & 7) + 1);
        r[t] = fmaf(-l_cent[q], sigma, r[t]);
      }
      const float gamma = sqrtf(chain_sumsq<G, TP>(r, G / 8 * TP, h, j, lane, grp0));
#pragma unroll
      for (int t = 0; t < TP; ++t) r[t] = r[t] * l_qjl[8 * (h * TP + t) + j];
      wht_group<G, TP>(r, lane);
#pragma unroll
      for (int t = 0; t < TP; ++t)
        nw[t >> 3] |= (r[t] * s.inv_sqrt_D >= 0.0f ? 1u : 0u) << (4 * (t & 7));
      // Lane j holds nibble j of words h*TP + 8k + m (m = 0..7); an 8x8
      // nibble transpose over the 8 lanes j (swapping lane bit 2 with nibble
      // bit 2, then bit 1, then bit 0) leaves lane j with word h*TP + 8k + j.
      const bool j4 = (j & 4) != 0, j2 = (j & 2) != 0, j1 = (j & 1) != 0;
      const uint32_t selA = j4 ? 0x07060302u : 0x01000504u;  // halves
      const uint32_t selB = j2 ? 0x07030501u : 0x02060004u;  // bytes
      const uint32_t rotC = j1 ? 4u : 28u;                   // nibbles
      const uint32_t keepC = j1 ? 0xF0F0F0F0u : 0x0F0F0F0Fu;
#pragma unroll
      for (int k = 0; k < TP / 8; ++k) {
        uint32_t w = nw[k];
        // Both reads run in every lane (a DPP read from a lane switched off
        // by a branch returns 0); each lane then keeps its partner's.
        const uint32_t pl = dpp_u<0x104>(w), ph = dpp_u<0x114>(w);
        uint32_t p = j4 ? ph : pl;
        w = __builtin_amdgcn_perm(w, p, selA);
        p = dpp_u<0x4E>(w);
        w = __builtin_amdgcn_perm(w, p, selB);
        p = dpp_u<0xB1>(w);
        p = __builtin_amdgcn_alignbit(p, p, rotC);
        w = (w & keepC) | (p & ~keepC);
        if (active) *reinterpret_cast<uint32_t *>(out + 4 * (h * TP + 8 * k + j)) = w;
      }
      if (active && lane == grp0) {
        float *meta = reinterpret_cast<float *>(out + s.packed);
        meta[0] = norm;
        meta[1] = gamma;
        meta[2] = sigma;
      }
      return;
    }

    if constexpr (!NIB && TP >= 4) {
      // Lloyd-Max index and residual (:361-367) for byte codes (up to 255
      // boundaries): binary search in LDS, QG coordinates side by side so
      // each level issues QG independent lookups. bw[k] gathers the units of
      // t = 4k + m as byte m; (idx << 1) keeps 8 bits as the reference's u8
      // unit does (:388-389).
      constexpr int QG = TP < 8 ? TP : 8;
      uint32_t bw[TP / 4];
#pragma unroll
      for (int k = 0; k < TP / 4; ++k) bw[k] = 0u;
#pragma unroll
      for (int t0 = 0; t0 < TP; t0 += QG) {
        float v[QG];
        int idx[QG];
#pragma unroll
        for (int e = 0; e < QG; ++e) {
          v[e] = r[t0 + e] * inv_sigma;
          idx[e] = 0;
        }
        for (int k = s.mse_bits - 1; k >= 0; --k) {
          const int step = 1 << k;
#pragma unroll
          for (int e = 0; e < QG; ++e) idx[e] += v[e] > l_bnd[idx[e] + step - 1] ? step : 0;
        }
#pragma unroll
        for (int e = 0; e < QG; ++e) {
          const int t = t0 + e;
          bw[t >> 2] |= ((uint32_t)(idx[e] << 1) & 0xFFu) << (8 * (t & 3));
          r[t] = fmaf(-l_cent[idx[e]], sigma, r[t]);
        }
      }
      const float gamma = sqrtf(chain_sumsq<G, TP>(r, G / 8 * TP, h, j, lane, grp0));
#pragma unroll
      for (int t = 0; t < TP; ++t) r[t] = r[t] * l_qjl[8 * (h * TP + t) + j];
      wht_group<G, TP>(r, lane);
#pragma unroll
      for (int t = 0; t < TP; ++t)
        bw[t >> 2] |= (r[t] * s.inv_sqrt_D >= 0.0f ? 1u : 0u) << (8 * (t & 3));
      // Lane j holds byte j & 3 of words 2*(h*TP + 4k + m) + (j >> 2)
      // (m = 0..3); a 4x4 byte transpose over the lanes of a quad (lane bit 1
      // with byte bit 1, then bit 0 with bit 0) leaves lane j with word
      // 2*(h*TP + 4k + (j & 3)) + (j >> 2).
      const uint32_t selA = (j & 2) ? 0x07060302u : 0x01000504u;  // halves
      const uint32_t selB = (j & 1) ? 0x07030501u : 0x02060004u;  // bytes
#pragma unroll
      for (int k = 0; k < TP / 4; ++k) {
        uint32_t w = bw[k];
        w = __builtin_amdgcn_perm(w, dpp_u<0x4E>(w), selA);
        w = __builtin_amdgcn_perm(w, dpp_u<0xB1>(w), selB);
        if (active)
          *reinterpret_cast<uint32_t *>(out + 4 * (2 * (h * TP + 4 * k + (j & 3)) + (j >> 2))) = w;
      }
      if (active && lane == grp0) {
        float *meta = reinterpret_cast<float *>(out + s.packed);
        meta[0] = norm;
        meta[1] = gamma;
        meta[2] = sigma;
      }
      return;
    }

    // Lloyd-Max index and residual (:361-367).
    uint32_t idxw[(TP + 3) / 4];
#pragma unroll
    for (int k = 0; k < (TP + 3) / 4; ++k) idxw[k] = 0u;
#pragma unroll
    for (int t = 0; t < TP; ++t) {
      int q = quantize_idx(r[t] * inv_sigma, l_bnd, s.mse_bits);
      idxw[t >> 2] |= (uint32_t)q << (8 * (t & 3));
      r[t] = fmaf(-l_cent[q], sigma, r[t]);
    }
    // gamma = ||residual|| (:369-372).
    const float gamma = sqrtf(chain_sumsq<G, TP>(r, G / 8 * TP, h, j, lane, grp0));

    // QJL signs of RHT_qjl(residual) (:374).
#pragma unroll
    for (int t = 0; t < TP; ++t) r[t] = r[t] * l_qjl[8 * (h * TP + t) + j];
    wht_group<G, TP>(r, lane);

    // Pack units (idx<<1)|sign (:375-391) and write the trailer (:393-395).
#pragma unroll
    for (int t = 0; t < TP; ++t) {
      uint32_t q = (idxw[t >> 2] >> (8 * (t & 3))) & 0xFFu;
      uint32_t bit = (r[t] * s.inv_sqrt_D >= 0.0f) ? 1u : 0u;
      uint32_t u = ((q << 1) | bit) & 0xFFu;
      if constexpr (NIB) {
        // word (h*TP + t) holds coordinates 8*(h*TP+t) .. +7, nibble j.
        uint32_t w = (u & 0xFu) << (4 * j);
        w |= dpp_u<0xB1>(w);
        w |= dpp_u<0x4E>(w);
        w |= dpp_u<0x141>(w);  // the other quad of the 8 lanes
        if (active && (t & 7) == j)
          *reinterpret_cast<uint32_t *>(out + 4 * (h * TP + t)) = w;
      } else {
        // word 2*(h*TP+t) + (j>>2) holds coordinates 4*word .. +3.
        uint32_t w = u << (8 * (j & 3));
        w |= dpp_u<0xB1>(w);
        w |= dpp_u<0x4E>(w);
        if (active && (t & 3) == (j & 3))
          *reinterpret_cast<uint32_t *>(out + 4 * (2 * (h * TP + t) + (j >> 2))) = w;
      }
    }
    if (active && lane == grp0) {
      float *meta = reinterpret_cast<float *>(out + s.packed);
      meta[0] = norm;
      meta[1] = gamma;
      meta[2] = sigma;
    }
  }
}

// D == 4 (input dims 1..4): one thread per vector, the same arithmetic
// written serially (the 8-wide chains are empty at this size).
template <int MODE>
__global__ __launch_bounds__(256) void k_encode_d4(SpaceDev s, const float *__restrict__ X,
                                                   long n, uint8_t *__restrict__ codes,
                                                   float *__restrict__ prep) {
  constexpr bool PREP = MODE == kModePrepare;
  const long vec = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (vec >= n) return;
  const float *x = X + vec * s.dim;
  float r[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int i = 0; i < s.dim; ++i) r[i] = x[i];
  float nsq = 0.0f;
  if constexpr (PREP) {
    for (int i = 0; i < s.dim; ++i) nsq = fmaf(x[i], x[i], nsq);
  } else {
    nsq = tail_sumsq(0.0f, x, 0, s.dim);
  }
  const float norm = sqrtf(nsq);
  const float inv_norm = MODE == kModeRotated ? ((nsq > 1e-20f) ? (1.0f / sqrtf(nsq)) : 0.0f)
                                              : ((norm > 1e-10f) ? (1.0f / norm) : 0.0f);
  for (int i = 0; i < 4; ++i) r[i] = (r[i] * inv_norm) * s.rot[i];
  auto wht4 = [&](float *v) {
    for (int st = 1; st < 4; st <<= 1)
      for (int i = 0; i < 4; i += 2 * st)
        for (int k = 0; k < st; ++k) {
          float a = v[i + k], b = v[i + st + k];
          v[i + k] = a + b;
          v[i + st + k] = a - b;
        }
    for (int i = 0; i < 4; ++i) v[i] = v[i] * s.inv_sqrt_D;
  };
  wht4(r);
  if constexpr (MODE == kModeRotated) {
    const float var = tail_sumsq(0.0f, r, 0, 4);
    const float inv_sigma = (var > 1e-20f) ? sqrtf(4.0f / var) : 1.0f;
    for (int i = 0; i < 4; ++i) prep[vec * 4 + i] = r[i] * inv_sigma;
  } else if constexpr (PREP) {
    float *pq = prep + vec * prep_stride(4);
    for (int i = 0; i < 4; ++i) pq[i] = r[i];
    for (int i = 0; i < 4; ++i) r[i] = r[i] * s.qjl[i];
    wht4(r);
    for (int i = 0; i < 4; ++i) pq[4 + i] = r[i];
    pq[8] = nsq;
    pq[9] = norm;
    pq[10] = 0.0f;
    pq[11] = 0.0f;
  } else {
    float var = tail_sumsq(0.0f, r, 0, 4);
    float sigma = sqrtf(var / 4.0f);
    if (sigma < 1e-10f) sigma = 1e-10f;
    const float inv_sigma = 1.0f / sigma;
    int idx[4];
    for (int i = 0; i < 4; ++i) {
      idx[i] = quantize_idx(r[i] * inv_sigma, s.bnd, s.mse_bits);
      r[i] = fmaf(-s.cent[idx[i]], sigma, r[i]);
    }
    const float gamma = sqrtf(tail_sumsq(0.0f, r, 0, 4));
    for (int i = 0; i < 4; ++i) r[i] = r[i] * s.qjl[i];
    wht4(r);
    uint8_t *out = codes + vec * s.cs;
    uint32_t u[4];
    for (int i = 0; i < 4; ++i)
      u[i] = ((uint32_t)(idx[i] << 1) | (r[i] >= 0.0f ? 1u : 0u)) & 0xFFu;
    if (s.nib) {
      out[0] = (uint8_t)(u[0] | (u[1] << 4));
      out[1] = (uint8_t)(u[2] | (u[3] << 4));
    } else {
      for (int i = 0; i < 4; ++i) out[i] = (uint8_t)u[i];
    }
    float meta[3] = {norm, gamma, sigma};
    for (int k = 0; k < 12; ++k) out[s.packed + k] = reinterpret_cast<uint8_t *>(meta)[k];
  }
}

template <int G, int TP, int MODE>
hipError_t launch_g(const SpaceDev &s, const float *X, long n, uint8_t *codes, float *prep,
                    hipStream_t st) {
  constexpr int threads = 256;
  const long per_block = (threads / 64) * (64 / G);
  const long blocks = (n + per_block - 1) / per_block;
  const size_t lds = sizeof(float) * (2 * (size_t)s.D + 2 * (size_t)s.L);
  // Past 64 KB (padded dim 8192) the kernel asks for more dynamic LDS.
  auto kern = s.nib ? k_encode<G, TP, true, MODE> : k_encode<G, TP, false, MODE>;
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(threads), lds, st, s, X, n, codes, prep);
  return hipGetLastError();
}

template <int MODE>
hipError_t dispatch(const SpaceDev &s, const float *X, long n, uint8_t *codes, float *prep,
                    hipStream_t st) {
  if (n <= 0) return hipSuccess;
  switch (s.D) {
    case 4: {
      const long blocks = (n + 255) / 256;
      hipLaunchKernelGGL((k_encode_d4<MODE>), dim3((unsigned)blocks), dim3(256), 0, st, s, X, n,
                         codes, prep);
      return hipGetLastError();
    }
    case 8: return launch_g<8, 1, MODE>(s, X, n, codes, prep, st);
    case 16: return launch_g<8, 2, MODE>(s, X, n, codes, prep, st);
    case 32: return launch_g<8, 4, MODE>(s, X, n, codes, prep, st);
    case 64: return launch_g<8, 8, MODE>(s, X, n, codes, prep, st);
    case 128: return launch_g<8, 16, MODE>(s, X, n, codes, prep, st);
    case 256: return launch_g<8, 32, MODE>(s, X, n, codes, prep, st);
    case 512: return launch_g<8, 64, MODE>(s, X, n, codes, prep, st);
    case 1024: return launch_g<16, 64, MODE>(s, X, n, codes, prep, st);
    case 2048: return launch_g<32, 64, MODE>(s, X, n, codes, prep, st);
    case 4096: return launch_g<64, 64, MODE>(s, X, n, codes, prep, st);
    case 8192: return launch_g<64, 128, MODE>(s, X, n, codes, prep, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

bool encode_supported(int D) { return D >= 4 && D <= 8192; }

hipError_t launch_encode(const SpaceDev &s, const float *X, long n, uint8_t *codes,
                         hipStream_t st) {
  return dispatch<kModeEncode>(s, X, n, codes, nullptr, st);
}

hipError_t launch_prepare(const SpaceDev &s, const float *Q, long m, float *prep,
                          hipStream_t st) {
  return dispatch<kModePrepare>(s, Q, m, nullptr, prep, st);
}

hipError_t launch_rotated(const SpaceDev &s, const float *X, long n, float *out,
                          hipStream_t st) {
  return dispatch<kModeRotated>(s, X, n, nullptr, out, st);
}

}  // namespace tq
