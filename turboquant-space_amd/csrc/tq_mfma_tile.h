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

// CO: the coarse pass set (mfma_coarse): no s_lo plane and no q_hi*c_lo
// pass; the caller corrects the near pairs (refine_flagged). WIDE: chunks of
// 64 coordinates (four k steps, one barrier per 64) instead of 32, for nibble
// codes; a chunk is then two 32-coordinate halves laid out as before.
template <bool NIB, int QT, bool QJL, bool CO = false, bool WIDE = false>
struct MfmaShape {
  static_assert(NIB || !WIDE, "wide chunks are for nibble codes");
  static constexpr int KB = WIDE ? 4 : 2;                     // 16-coordinate k steps per chunk
  static constexpr int RB = QT / 32;                          // 32-row blocks per wavefront
  static constexpr int NPL = split_planes(NIB, QJL, CO);      // query planes read
  static constexpr int NB = NPL * RB * KB;                    // 1 KB A blocks per chunk
  static constexpr int ALOADS = (NB * 64 + 255) / 256;        // 16-byte A loads per thread
  static constexpr int NC = NIB ? KB / 2 : 2;                 // 16-byte code loads per thread
  static constexpr int A_BYTES = NB * 1024;
  static constexpr int STAGE = A_BYTES + 4 * NC * 1024;
  // Chunks whose loads are in the register ring: three where a chunk takes
  // at most four 16-byte loads per thread, two up to six (the 64-query QJL
  // tiles of 4-bit codes with every pass, and of byte codes), one for wide
  // chunks (one chunk of lookahead = two 32-coordinate ones).
  static constexpr int RING = ALOADS + NC <= 4 ? 3 : ALOADS + NC <= 6 ? 2 : 1;
};

// acc_m / acc_q of queries [q0, q0 + QT) x codes [j0, j0 + 256) of one
// workgroup of 256 threads (wavefront w: codes j0 + 64 w .. + 63, as RB x 2
// blocks of 32 x 32, element v of a block at row 8 (v / 4) + 4 kg + v % 4,
// column lr). `stage` is the workgroup's two LDS stage buffers; l_tab the
// byte codes' centroid table (REP: replicated over 32 banks, decode8<true>)
// and ptab the nibble codes' perm tables. Ends with a workgroup barrier after
// the last chunk, so the caller may reuse the stage buffers.
template <bool NIB, int QT, bool QJL, bool REP, bool CO, bool WIDE = false, bool RUNS = false>
__device__ __forceinline__ void mfma_tile(
    const SpaceDev &s, const _Float16 *__restrict__ planes, long m,
    const uint8_t *__restrict__ codes, long n, long q0, long j0,
    uint8_t (*stage)[MfmaShape<NIB, QT, QJL, CO, WIDE>::STAGE], const uint32_t *l_tab,
    const PermTab &ptab, f16v (&acc_m)[QT / 32][2], f16v (&acc_q)[QJL ? QT / 32 : 1][2]) {
  using Sh = MfmaShape<NIB, QT, QJL, CO, WIDE>;
  constexpr int RB = Sh::RB, NPL = Sh::NPL, NB = Sh::NB, NC = Sh::NC, KB = Sh::KB;
  constexpr int ALOADS = Sh::ALOADS, kRing = Sh::RING;
  static_assert(QT == 32 || QT == 64, "query tile");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 31, kg = lane >> 5;
  const int nchunks = s.D / (16 * KB);
  const int nks = s.D / 16;

  // Per-thread load sources and LDS destinations, fixed for the whole walk
  // over the padded dim; a chunk only adds its offset. A chunk's A block
  // i = (row block r, k step kb, plane p), i = (r * KB + kb) * NPL + p, is
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
    const int r = i / (KB * NPL), kb = (i / NPL) % KB, p = i % NPL;
    long rb = q0 / 32 + r;
    if (rb >= nrb) rb = nrb - 1;
    asrc[t] = planes + ((rb * nks + kb) * NPL + p) * (long)kBlockHalves + 8 * l;
    adst[t] = u * 16;
  }
  // Code bytes: 16-byte part h of the chunk's bytes of code j0 + 64 w + c
  // goes to LDS region (w * NC + h) at 16 c. Load u of thread (w, l) takes
  // part u of code l, or with RUNS (and NC parts) part l % NC of code
  // (64 / NC) u + l / NC, so that one load instruction reads whole 16 NC-byte
  // runs of 64 / NC codes (fewer cache lines per instruction than 16 bytes of
  // each of 64 codes; 1.5% on the 64 x 10M dim-768 scan, but two more
  // pointers than the fused search's registers can spare).
  constexpr bool R = RUNS && NC > 1;
  const int cpart = R ? lane % NC : 0;
  const uint8_t *csrc[NC];
#pragma unroll
  for (int u = 0; u < NC; ++u) {
    long jl = j0 + wave * 64 + (R ? (64 / NC) * u + lane / NC : lane);
    if (jl >= n) jl = n - 1;
    csrc[u] = codes + jl * (long)s.cs + 16 * (R ? cpart : u);
  }
  const long achunk = (long)KB * NPL * kBlockHalves;  // A halves per chunk

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
      ring_c[k][h] = *reinterpret_cast<const u32x4a4 *>(csrc[h] + (NIB ? 8 : 16) * KB * cc);
    __builtin_amdgcn_sched_barrier(0);
  };
  auto store_group = [&](auto slot, int b) __attribute__((always_inline)) {
    constexpr int k = decltype(slot)::value;
#pragma unroll
    for (int t = 0; t < ALOADS; ++t)
      *reinterpret_cast<u32x4 *>(&stage[b][adst[t]]) = ring_a[k][t];
#pragma unroll
    for (int u = 0; u < NC; ++u) {
      const int c = R ? (64 / NC) * u + lane / NC : lane;
      const int part = R ? cpart : u;
      *reinterpret_cast<u32x4 *>(&stage[b][Sh::A_BYTES + (wave * NC + part) * 1024 + 16 * c]) =
          ring_c[k][u];
    }
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
    uint32_t cw[2][KB < 4 ? 4 : KB];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if constexpr (NIB) {
        // 32-coordinate half h of the chunk: its 16 code bytes in region h
#pragma unroll
        for (int h = 0; h < NC; ++h) {
          const uint2 v =
              *reinterpret_cast<const uint2 *>(sc + h * 1024 + (32 * c + lr) * 16 + 8 * kg);
          cw[c][2 * h] = v.x;
          cw[c][2 * h + 1] = v.y;
        }
      } else {
        const uint4 v = *reinterpret_cast<const uint4 *>(sc + kg * 1024 + (32 * c + lr) * 16);
        cw[c][0] = v.x;
        cw[c][1] = v.y;
        cw[c][2] = v.z;
        cw[c][3] = v.w;
      }
    }
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      h8 a[RB][NPL];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int p = 0; p < NPL; ++p)
          a[r][p] =
              *reinterpret_cast<const h8 *>(sa + ((r * KB + kb) * NPL + p) * 1024 + 16 * lane);
      h8 bh[2], bl[2], bt[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        if constexpr (NIB) {
          decode_nib_word<!CO>(cw[c][kb], ptab, bh[c], bl[c], bt[c]);
        } else {
          decode8<REP>(cw[c][2 * kb], cw[c][2 * kb + 1], l_tab, bh[c], bl[c], bt[c],
                       (uint32_t)lr);
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
      if constexpr (!CO) { TQ_PASS(acc_m, 0, bl) }
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
  if constexpr (kRing >= 2) load_group(S1{}, 1);
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
  } else if constexpr (kRing == 2) {
    for (int t = nchunks / 2; t > 0; --t) {
      step(S1{});
      step(S0{});
    }
    if (nchunks % 2) step(S1{});
  } else {
    for (int t = nchunks; t > 0; --t) step(S0{});
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

// Coarse form: a pair is corrected when its coarse distance is below
// kRefineFrac of (|q|^2 + |x|^2). The terms the coarse pass set drops
// (q_hi*c_lo, s_lo*(+-1)) move a distance by up to a few 1e-5 of that scale,
// which is a few 1e-5 relative to every distance above the bar; the pairs
// below it (a query's near codes, the ones a top-k decides on) get the
// dropped terms added (refine_flagged), which makes them the full pass set's
// value (only q_lo*c_lo, ~2^-22 relative, stays out).
constexpr float kRefineFrac = 0.5f;
// The dropped terms of each 8-coordinate group are summed in fp32 by one
// lane in a fixed order, and the group sums as 32-bit fixed-point integers,
// so the totals are the same in any lane grouping: q_hi * c_lo at 2^13
// (|q_hi| <= 2^10 |q|, |c_lo| <= 0.25, so the sum of magnitudes is below
// 2^8 sqrt(D) |q| <= 2^14.5 |q| at D = 8192: no overflow below |q| = 11, and
// prepared rows have |q| = 1), each group sum rounded once (2^-14, ~1e-9 of a
// near pair's sum); s_lo (fp16, multiples of 2^-24) at 2^22 (sum of
// magnitudes <= 2^5.5 |s| at D = 8192).
constexpr int kRefineShiftM = 13;
constexpr int kRefineShiftQ = 22;

__device__ __forceinline__ bool refine_flag(float d, float qns, float xns) {
  return d < kRefineFrac * (qns + xns);
}

// The search forms correct only the near pairs that can still reach their
// query's top-k: a pair whose coarse distance d satisfies d - E > T, with T
// an upper bound on the query's final k-th best distance (a running list's
// k-th entry) and E >= |corrected d - d| (refine_bound), cannot enter the
// top-k either way, so it keeps its coarse value and every form returns the
// same top-k. E comes from bounds on the correction sums of the pair's query
// (refine_limits): |dm| <= sum |q_hi| * max |c_lo| plus the fixed-point
// rounding (2^-14 per 8-coordinate group), |dq| <= sum |s_lo| plus 2^-23 per
// group (4-bit spaces with QJL only), each widened by 2^-10 for the fp32
// sums behind them. Rows whose sums could leave the 32-bit fixed-point range
// (never for prepared rows, |q| = 1) get an infinite bound: never pruned.
__device__ __forceinline__ float2 refine_limits(const SpaceDev &s, bool sq, float l1_qhi,
                                                float l1_slo) {
  const float ng = (float)(s.D >> 3);
  float dmb = (l1_qhi * s.clo_max + ng * 0x1p-14f) * (1.0f + 0x1p-10f);
  const float dqb = sq ? (l1_slo + ng * 0x1p-23f) * (1.0f + 0x1p-10f) : 0.0f;
  if (!(dmb < 0x1p17f && dqb < 0x1p8f)) dmb = __builtin_inff();
  return make_float2(dmb, dqb);
}

// E for one pair: mfma_distance's change when am, aq move by at most lim.x,
// lim.y (2 |dip|, dip = (dm fm + dq fq) xn qn), plus 2^-18 of the magnitudes
// its fp32 steps round (the products, the sums and qns + xns: about 22 units
// of 2^-24 in all, the final d - E comparison's included), the whole widened
// by 2^-16 for this evaluation's own rounding. NaN or inf inputs give a NaN
// or inf E, which never prunes.
__device__ __forceinline__ float refine_bound(float am, float aq, float fm, float fq, float xn,
                                              float xns, float qns, float qn, float2 lim) {
  const float f = __builtin_fabsf(fm), g = __builtin_fabsf(fq);
  const float t = (lim.x * f + lim.y * g) * xn * qn;
  const float p = (__builtin_fabsf(am) * f + __builtin_fabsf(aq) * g) * xn * qn;
  return (2.0f * t * (1.0f + 0x1p-18f) + 0x1p-18f * (p + qns + xns)) * (1.0f + 0x1p-16f);
}

// The bits of `flag` (one lane's flagged pairs) for which keep(b) holds;
// keep runs in the bit's own lane, one bit per lane per step, until every
// lane's bits are checked. Every lane of the wavefront must call it.
template <typename Keep>
__device__ __forceinline__ uint64_t prune_flagged(uint64_t flag, Keep &&keep) {
  uint64_t pend = flag, kept = 0;
  while (__any(pend != 0u)) {
    if (pend != 0u) {
      const int b = __builtin_ctzll(pend);
      pend &= pend - 1u;
      if (keep(b)) kept |= 1ull << b;
    }
  }
  return kept;
}

// The terms the coarse passes drop for one (query, code) pair, in the MFMA
// sums' units: dm = sum_i q_hi[i] * c_lo[idx_i], and for 4-bit codes with
// QJL dq = sum_i s_lo[i] * (+-1)_i (byte codes keep no s_lo term in either
// pass set), from the query's correction row (tq_mfma_dev.h) and the code's
// units, by the L lanes of one L-lane group of a wavefront (L a power of
// two): lane `sub` takes the 8-coordinate groups sub, sub + L, ...; each
// group's fp32 sum (elements 0..7 in order) becomes an integer
// (kRefineShiftM / kRefineShiftQ) before the lanes' sums meet, so every L
// and every lane gives the same bits. clo(k): the scaled centroid remainder c_lo
// of index k (from LDS). QJL = false: the symmetric MSE form (no s term).
template <bool NIB, bool QJL, typename Clo>
__device__ __forceinline__ void refine_delta(const SpaceDev &s, Clo &&clo,
                                             const _Float16 *__restrict__ crow,
                                             const uint8_t *__restrict__ code, int sub, int L,
                                             float &dm, float &dq) {
  constexpr bool SQ = QJL && NIB;  // an s_lo term to add
  // Unsigned sums: wrap-around instead of undefined behaviour should a row
  // be far outside the bound (codes with garbage trailers in the symmetric
  // form); in range they are the two's-complement sums.
  uint32_t am = 0u, aq = 0u;
  const int ng = s.D >> 3;
  // Eight groups per step, all loads issued before the first use (a group
  // past the end loads group `sub` again and is not added).
  for (int g0 = sub; g0 < ng; g0 += 8 * L) {
    uint2 w[8];
    h8 qh[8], sl[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int g = g0 + L * t < ng ? g0 + L * t : sub;
      if constexpr (NIB) {
        w[t].x = *reinterpret_cast<const uint32_t *>(code + 4 * g);
        w[t].y = 0u;
      } else {
        w[t] = *reinterpret_cast<const uint2 *>(code + 8 * g);
      }
      qh[t] = *reinterpret_cast<const h8 *>(crow + g * kCorrHalves);
      if constexpr (SQ) sl[t] = *reinterpret_cast<const h8 *>(crow + g * kCorrHalves + 8);
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (g0 + L * t >= ng) break;
      float pm = 0.0f, pq = 0.0f;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        // element e's unit: nibble perm(e) of the word, or byte e of the pair
        const uint32_t u = NIB ? (w[t].x >> (4 * split_perm(true, e))) & 15u
                               : ((e < 4 ? w[t].x : w[t].y) >> (8 * (e & 3))) & 0xFFu;
        pm = __builtin_fmaf((float)qh[t][e], clo(u >> 1), pm);
        if constexpr (SQ) pq += (u & 1u) ? (float)sl[t][e] : -(float)sl[t][e];
      }
      am += (uint32_t)__float2int_rn(__builtin_ldexpf(pm, kRefineShiftM));
      if constexpr (SQ) aq += (uint32_t)__float2int_rn(__builtin_ldexpf(pq, kRefineShiftQ));
    }
  }
  for (int o = L >> 1; o > 0; o >>= 1) {  // lane ^ o stays within the group
    am += (uint32_t)__shfl_xor((int)am, o);
    if constexpr (SQ) aq += (uint32_t)__shfl_xor((int)aq, o);
  }
  dm = __builtin_ldexpf((float)(int)am, -kRefineShiftM);
  dq = SQ ? __builtin_ldexpf((float)(int)aq, -kRefineShiftQ) : 0.0f;
}

// A wavefront's flagged pairs (bit b of `flag` in lane l: pair pos(b, l) =
// (query row, code column), b = 16 * block + element) get the dropped terms
// computed by refine_delta, up to eight pairs per round taken in lane then
// bit order (at most one per lane), the wavefront split into as many equal
// lane groups as the round has pairs (rounded up to 1, 2, 4 or 8): a lone
// pair gets all 64 lanes. The owning lane then calls finish(b, dm, dq) (b
// differs between lanes). `corr` is the batch's correction rows. Every lane
// of the wavefront must call it.
template <bool NIB, bool QJL, typename Clo, typename Pos, typename Finish>
__device__ __forceinline__ void refine_flagged(const SpaceDev &s, Clo &&clo,
                                               const _Float16 *__restrict__ corr,
                                               const uint8_t *__restrict__ codes, uint64_t flag,
                                               int lane, Pos &&pos, Finish &&finish) {
  constexpr int kMaxPairs = 8;
  uint64_t any = __ballot(flag != 0u);
  while (any) {
    int src[kMaxPairs], bit[kMaxPairs];
    bool ok[kMaxPairs];
    int np = 0;
#pragma unroll
    for (int p = 0; p < kMaxPairs; ++p) {
      ok[p] = any != 0;
      if (ok[p]) {
        src[p] = __builtin_ctzll(any);
        any &= any - 1;  // this lane is done for the round
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)flag, src[p]);
        const uint32_t hi =
            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(flag >> 32), src[p]);
        bit[p] = __builtin_ctzll(((uint64_t)hi << 32) | lo);
        np = p + 1;
      } else {  // an idle group scores pair 0 again and keeps nothing
        src[p] = src[0];
        bit[p] = bit[0];
      }
    }
    const int groups = np <= 1 ? 1 : np <= 2 ? 2 : np <= 4 ? 4 : 8;  // wavefront-uniform
    const int L = 64 / groups;
    const int grp = lane / L;
    long q = 0, j = 0;
#pragma unroll
    for (int p = 0; p < kMaxPairs; ++p) {
      long qp, jp;
      pos(bit[p], src[p], qp, jp);
      if (grp == p) {
        q = qp;
        j = jp;
      }
    }
    float dm, dq;
    refine_delta<NIB, QJL>(s, clo, corr + q * (long)(s.D / 8) * kCorrHalves,
                           codes + j * (long)s.cs, lane % L, L, dm, dq);
    int mb = -1;
    float mdm = 0.0f, mdq = 0.0f;
#pragma unroll
    for (int p = 0; p < kMaxPairs; ++p) {
      const int l0 = (L * p) & 63;  // group p's first lane (p < groups whenever ok[p])
      const float pm = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dm), l0));
      const float pq = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dq), l0));
      if (ok[p] && lane == src[p]) {
        mb = bit[p];
        mdm = pm;
        mdq = pq;
      }
    }
    if (mb >= 0) {
      flag &= flag - 1u;  // bit mb is this lane's lowest
      finish(mb, mdm, mdq);
    }
    any = __ballot(flag != 0u);
  }
}

// The scaled centroid remainder of index k (c * cscale - c_hi, the part of
// each centroid the coarse passes leave out; 0 past the L centroids), for
// l_clo.
__device__ __forceinline__ float cent_lo(const SpaceDev &s, int k) {
  _Float16 ch, cl;
  cent_h(s, k, ch, cl);
  return k < s.L ? s.cent[k] * s.cscale - (float)ch : 0.0f;
}

}  // namespace mfma
}  // namespace tq
