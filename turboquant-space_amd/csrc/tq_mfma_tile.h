// tq_mfma_tile.h -- the MFMA contraction of one (query tile, code tile), shared
// by the m x n kernel (k_mfma, tq_mfma.hip) and the fused search kernel
// (k_search, tq_search.hip), so both give the same bits for a (query, code)
// pair: the same split planes, passes, pass order and K slots, and the same
// fp32 epilogue expression.
#pragma once
#include <type_traits>

#include "tq_mfma_dev.h"

namespace tq {
namespace mfma {

constexpr int kTN = 256;  // codes per workgroup tile (64 per wavefront)
constexpr int kKC = 32;   // coordinates per chunk
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// The same at 4-byte alignment (code rows are 4-byte aligned).
typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

template <bool NIB, int QT, bool QJL>
struct MfmaShape {
  static constexpr int RB = QT / 32;                          // 32-row blocks per wavefront
  static constexpr int NPL = split_planes(NIB, QJL);          // query planes read
  static constexpr int NB = NPL * RB * 2;                     // 1 KB A blocks per chunk
  static constexpr int ALOADS = (NB * 64 + 255) / 256;        // 16-byte A loads per thread
  static constexpr int NC = NIB ? 1 : 2;                      // 16-byte code loads per thread
  static constexpr int A_BYTES = NB * 1024;
  static constexpr int STAGE = A_BYTES + 4 * NC * 1024;
  // Chunks whose loads are in the register ring: three where the registers
  // allow it, two for the 64-query tiles with QJL passes.
  static constexpr int RING = (QT == 64 && QJL) ? 2 : 3;
};

// acc_m / acc_q of queries [q0, q0 + QT) x codes [j0, j0 + 256) of one
// workgroup of 256 threads (wavefront w: codes j0 + 64 w .. + 63, as RB x 2
// blocks of 32 x 32, element v of a block at row 8 (v / 4) + 4 kg + v % 4,
// column lr). `stage` is the workgroup's two LDS stage buffers; l_tab the
// byte codes' centroid table (REP: replicated over 32 banks, decode8<true>)
// and ptab the nibble codes' perm tables. Ends with a workgroup barrier after
// the last chunk, so the caller may reuse the stage buffers.
template <bool NIB, int QT, bool QJL, bool REP>
__device__ __forceinline__ void mfma_tile(const SpaceDev &s, const _Float16 *__restrict__ planes,
                                          long m, const uint8_t *__restrict__ codes, long n,
                                          long q0, long j0,
                                          uint8_t (*stage)[MfmaShape<NIB, QT, QJL>::STAGE],
                                          const uint32_t *l_tab, const PermTab &ptab,
                                          f16v (&acc_m)[QT / 32][2],
                                          f16v (&acc_q)[QJL ? QT / 32 : 1][2]) {
  using Sh = MfmaShape<NIB, QT, QJL>;
  constexpr int RB = Sh::RB, NPL = Sh::NPL, NB = Sh::NB, NC = Sh::NC;
  constexpr int ALOADS = Sh::ALOADS, kRing = Sh::RING;
  static_assert(QT == 32 || QT == 64, "query tile");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 31, kg = lane >> 5;
  const int nchunks = s.D / kKC;
  const int nks = s.D / 16;

  // Per-thread load sources and LDS destinations, fixed for the whole walk
  // over the padded dim; a chunk only adds its offset. A chunk's A block
  // i = (row block r, k step kb, plane p), i = (r * 2 + kb) * NPL + p, is
  // 64 units of 16 bytes; thread t loads units t, t + 256, ... (a unit index
  // past the last loads the last unit again: same bytes, same place). Query
  // blocks past m read the last block, codes past n read code n - 1: every
  // load is unconditional, and those rows and columns are never written out.
  const long nrb = (m + 31) / 32;
  const _Float16 *asrc[ALOADS];
  int adst[ALOADS];
#pragma unroll
  for (int t = 0; t < ALOADS; ++t) {
    const int u = t * 256 + tid < NB * 64 ? t * 256 + tid : NB * 64 - 1;
    const int i = u >> 6, l = u & 63;
    const int r = i / (2 * NPL), kb = (i / NPL) & 1, p = i % NPL;
    long rb = q0 / 32 + r;
    if (rb >= nrb) rb = nrb - 1;
    asrc[t] = planes + ((rb * nks + kb) * NPL + p) * (long)kBlockHalves + 8 * l;
    adst[t] = u * 16;
  }
  // Code bytes: thread (w, l) loads 16-byte part h of the chunk's bytes of
  // code j0 + 64 w + l into LDS region (w * NC + h) at 16 l.
  long jl = j0 + wave * 64 + lane;
  if (jl >= n) jl = n - 1;
  const uint8_t *csrc = codes + jl * (long)s.cs;
  const long achunk = 2L * NPL * kBlockHalves;  // A halves per chunk

  u32x4 ring_a[kRing][ALOADS], ring_c[kRing][NC];
  auto load_group = [&](auto slot, int ch) __attribute__((always_inline)) {
    constexpr int k = decltype(slot)::value;
    const int cc = ch < nchunks ? ch : nchunks - 1;
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < ALOADS; ++t)
      ring_a[k][t] = *reinterpret_cast<const u32x4 *>(asrc[t] + cc * achunk);
#pragma unroll
    for (int h = 0; h < NC; ++h)
      ring_c[k][h] = *reinterpret_cast<const u32x4a4 *>(csrc + (NIB ? 16 : 32) * cc + 16 * h);
    __builtin_amdgcn_sched_barrier(0);
  };
  auto store_group = [&](auto slot, int b) __attribute__((always_inline)) {
    constexpr int k = decltype(slot)::value;
#pragma unroll
    for (int t = 0; t < ALOADS; ++t)
      *reinterpret_cast<u32x4 *>(&stage[b][adst[t]]) = ring_a[k][t];
#pragma unroll
    for (int h = 0; h < NC; ++h)
      *reinterpret_cast<u32x4 *>(&stage[b][Sh::A_BYTES + (wave * NC + h) * 1024 + 16 * lane]) =
          ring_c[k][h];
  };

#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        acc_m[r][c][v] = 0.0f;
        if constexpr (QJL) acc_q[r][c][v] = 0.0f;
      }

  // One chunk from LDS buffer b: per k step, the A fragments of every plane,
  // the decoded B fragments of the wavefront's two code blocks, then the
  // passes. Each pass runs over all blocks before the next, so consecutive
  // MFMAs never wait on the accumulator they just wrote.
  auto compute = [&](int b) __attribute__((always_inline)) {
    const uint8_t *sa = &stage[b][0];
    const uint8_t *sc = &stage[b][Sh::A_BYTES + wave * NC * 1024];
    uint32_t cw[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if constexpr (NIB) {
        const uint2 v = *reinterpret_cast<const uint2 *>(sc + (32 * c + lr) * 16 + 8 * kg);
        cw[c][0] = v.x;
        cw[c][1] = v.y;
      } else {
        const uint4 v = *reinterpret_cast<const uint4 *>(sc + kg * 1024 + (32 * c + lr) * 16);
        cw[c][0] = v.x;
        cw[c][1] = v.y;
        cw[c][2] = v.z;
        cw[c][3] = v.w;
      }
    }
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      h8 a[RB][NPL];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int p = 0; p < NPL; ++p)
          a[r][p] = *reinterpret_cast<const h8 *>(sa + ((r * 2 + kb) * NPL + p) * 1024 + 16 * lane);
      h8 bh[2], bl[2], bt[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        if constexpr (NIB) {
          decode_nib_word(cw[c][kb], ptab, bh[c], bl[c], bt[c]);
        } else {
          uint32_t u[8];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            u[v] = (cw[c][2 * kb] >> (8 * v)) & 0xFFu;
            u[4 + v] = (cw[c][2 * kb + 1] >> (8 * v)) & 0xFFu;
          }
          decode8<REP>(u, l_tab, bh[c], bl[c], bt[c], (uint32_t)lr);
        }
      }
#define TQ_PASS(ACC, A, B)                                                               \
  _Pragma("unroll") for (int r = 0; r < RB; ++r)                                         \
      _Pragma("unroll") for (int c = 0; c < 2; ++c) ACC[r][c] =                          \
          __builtin_amdgcn_mfma_f32_32x32x16_f16(a[r][A], B[c], ACC[r][c], 0, 0, 0);
      TQ_PASS(acc_m, 0, bh)
      if constexpr (QJL) { TQ_PASS(acc_q, 2, bt) }
      TQ_PASS(acc_m, 1, bh)
      if constexpr (NPL > 3) { TQ_PASS(acc_q, 3, bt) }
      TQ_PASS(acc_m, 0, bl)
#undef TQ_PASS
    }
  };

  // One chunk: the matrix work on LDS buffer ch & 1, then chunk ch + 1 from
  // ring slot `slot` into the other buffer, the slot reloaded with chunk
  // ch + 1 + kRing, one barrier. The hardware's load counter completes in
  // order, so the store of a chunk waits for its own loads only.
  int ch = 0;
  auto step = [&](auto slot) __attribute__((always_inline)) {
    const int b = ch & 1;
    compute(b);
    store_group(slot, b ^ 1);
    load_group(slot, ch + 1 + kRing);
    __syncthreads();
    ++ch;
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  using S2 = std::integral_constant<int, kRing == 3 ? 2 : 0>;
  load_group(S0{}, 0);
  load_group(S1{}, 1);
  if constexpr (kRing == 3) load_group(S2{}, 2);
  store_group(S0{}, 0);
  load_group(S0{}, kRing);
  __syncthreads();
  // Whole turns of the ring (no exit mid-turn), then the chunks of a last
  // partial turn. Chunks past the end are reloads of the last one, stored
  // into the buffer no later chunk reads.
  if constexpr (kRing == 3) {
    for (int t = nchunks / 3; t > 0; --t) {
      step(S1{});
      step(S2{});
      step(S0{});
    }
    if (nchunks % 3 >= 1) step(S1{});
    if (nchunks % 3 >= 2) step(S2{});
  } else {
    for (int t = nchunks / 2; t > 0; --t) {
      step(S1{});
      step(S0{});
    }
    if (nchunks % 2) step(S1{});
  }

}

// The reference's distance of one pair from the tile's sums
// (space_turbo_quant.h:682-685): ip = (sigma * mse + scale * gamma * dot) *
// |x| * |q|, d = max(0, |q|^2 + |x|^2 - 2 ip), with fm = sigma * mse_unscale
// and fq = scale * gamma * 2^-10 formed once per code.
template <bool QJL>
__device__ __forceinline__ float mfma_distance(float am, float aq, float fm, float fq, float xn,
                                               float xns, float qns, float qn) {
  float ip;
  if constexpr (QJL)
    ip = (am * fm + aq * fq) * xn * qn;
  else
    ip = am * fm * xn * qn;
  const float d = qns + xns - 2.0f * ip;
  return d > 0.0f ? d : 0.0f;
}

}  // namespace mfma
}  // namespace tq
