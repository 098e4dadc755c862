// tq_mfma_dev.h -- device helpers shared by the MFMA kernels (tq_mfma.hip,
// tq_search.hip): the fp16 hi/lo scaling constants, the layout of the split
// query planes, and the code-word decoders that turn packed code units into
// B-operand fragments.
#pragma once
#include "tq_common.h"

namespace tq {
namespace mfma {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

// Queries (q_rot, s_q) are scaled by 2^10; centroids by the space's cscale
// (SpaceDev::cscale, undone with SpaceDev::mse_unscale).
constexpr float kQScale = 1024.0f;              // 2^10 on q_rot and s_q
constexpr float kQjlUnscale = 1.0f / 1024.0f;   // 2^-10

// Split query planes. Plane 0 = q_hi, 1 = q_lo, 2 = s_hi, 3 = s_lo (fp16 of
// the 2^10-scaled q_rot and s_q, lo = fp16 of the remainder). The asymmetric
// form of a 2- or 3-bit nibble space reads all four; byte spaces and the
// coarse form of 4-bit spaces (mfma_coarse) skip s_lo (a byte space's QJL
// term carries a small gamma; the coarse form recomputes the pairs the
// dropped terms matter for); the symmetric MSE form reads q_hi and q_lo.
__host__ __device__ constexpr int split_planes(bool nib, bool qjl, bool co = false) {
  return qjl ? (nib && !co ? 4 : 3) : 2;
}
// Layout: blocks of one MFMA A fragment (32 queries x 16 K slots, 1 KB),
// [query block rb][k step ks][plane p][lane l][8 halves], ks = 0 .. D/16 - 1.
// Lane l = lr + 32 kg of block (rb, ks) holds query 32 rb + lr; its element
// e is coordinate 32 (ks / 2) + 16 kg + 8 (ks % 2) + perm(e), where perm is
// (0, 2, 4, 6, 1, 3, 5, 7) for nibble spaces (the order decode_nib_word emits
// a code word's coordinates in) and the identity for byte spaces. So within a
// 32-coordinate chunk, lane half kg covers code bytes [8 kg, 8 kg + 8)
// (nibbles) or [16 kg, 16 kg + 16) (bytes) for both k steps.
constexpr int kBlockHalves = 512;
__host__ __device__ inline long split_planes_halves(int D, long m, int npl) {
  return (m + 31) / 32 * (long)(D / 16) * npl * kBlockHalves;
}
// The coarse form (mfma_coarse) adds, after the planes of the padded rows,
// one correction row per query: for each group g of 8 coordinates, 16 halves
// = q_hi of the group's coordinates, then s_lo, both in the planes' element
// order (perm below). refine_flagged reads it for the terms the coarse passes
// drop (q_hi * c_lo, s_lo * (+-1)).
constexpr int kCorrHalves = 16;  // halves per 8-coordinate group
__host__ __device__ inline long corr_halves(int D, long m) {
  return (m + 31) / 32 * 32 * (long)(D / 8) * kCorrHalves;
}
// After the correction rows, two floats per padded row: sum |q_hi| and
// sum |s_lo| of the row (k_corr_norms), for refine_limits.
__host__ __device__ inline long corr_norm_floats(long m) { return (m + 31) / 32 * 32 * 2; }
__host__ __device__ inline int split_perm(bool nib, int e) {
  return nib ? (e < 4 ? 2 * e : 2 * (e - 4) + 1) : e;
}

// fp16 hi and lo parts of the scaled centroid k (0 past the L centroids).
__device__ __forceinline__ void cent_h(const SpaceDev &s, int k, _Float16 &ch, _Float16 &cl) {
  const float c = k < s.L ? s.cent[k] * s.cscale : 0.0f;
  ch = (_Float16)c;
  cl = (_Float16)(c - (float)ch);
}

__device__ __forceinline__ uint32_t h_bits(_Float16 h) {
  return (uint32_t)__builtin_bit_cast(uint16_t, h);
}

// Byte tables for v_perm decoding of 3-bit centroid indices: table t holds
// one byte of the 8 fp16 values as two 32-bit words (indices 0-3 | 4-7).
struct PermTab {
  uint32_t hl_lo, hl_hi;  // low bytes of c_hi
  uint32_t hh_lo, hh_hi;  // high bytes of c_hi
  uint32_t ll_lo, ll_hi;  // low bytes of c_lo
  uint32_t lh_lo, lh_hi;  // high bytes of c_lo
};

// The perm tables of a nibble space's 8 centroids.
__device__ __forceinline__ PermTab perm_tab(const SpaceDev &s) {
  uint32_t bytes[4][2] = {{0u, 0u}, {0u, 0u}, {0u, 0u}, {0u, 0u}};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    _Float16 chh, cll;
    cent_h(s, k, chh, cll);
    const uint32_t hb = h_bits(chh), lb = h_bits(cll);
    const int word = k >> 2, sh = 8 * (k & 3);
    bytes[0][word] |= (hb & 0xFFu) << sh;
    bytes[1][word] |= (hb >> 8) << sh;
    bytes[2][word] |= (lb & 0xFFu) << sh;
    bytes[3][word] |= (lb >> 8) << sh;
  }
  return PermTab{bytes[0][0], bytes[0][1], bytes[1][0], bytes[1][1],
                 bytes[2][0], bytes[2][1], bytes[3][0], bytes[3][1]};
}

// 8 nibble units in one code word -> B fragments (c_hi, c_lo unless !LO,
// +-1) in the order (0, 2, 4, 6, 1, 3, 5, 7): bytes of the even (odd)
// coordinates' values are looked up four at a time, then interleaved into
// fp16 pairs.
template <bool LO = true>
__device__ __forceinline__ void decode_nib_word(uint32_t w, const PermTab &t, h8 &bh, h8 &bl,
                                                h8 &bt) {
  const uint32_t ie = (w >> 1) & 0x07070707u;  // indices of coordinates 0, 2, 4, 6
  const uint32_t io = (w >> 5) & 0x07070707u;  // indices of coordinates 1, 3, 5, 7
  uint32_t r[4];
  {
    const uint32_t Le = __builtin_amdgcn_perm(t.hl_hi, t.hl_lo, ie);
    const uint32_t He = __builtin_amdgcn_perm(t.hh_hi, t.hh_lo, ie);
    const uint32_t Lo = __builtin_amdgcn_perm(t.hl_hi, t.hl_lo, io);
    const uint32_t Ho = __builtin_amdgcn_perm(t.hh_hi, t.hh_lo, io);
    r[0] = __builtin_amdgcn_perm(He, Le, 0x05010400u);
    r[1] = __builtin_amdgcn_perm(He, Le, 0x07030602u);
    r[2] = __builtin_amdgcn_perm(Ho, Lo, 0x05010400u);
    r[3] = __builtin_amdgcn_perm(Ho, Lo, 0x07030602u);
    bh = __builtin_bit_cast(h8, r);
  }
  if constexpr (LO) {
    const uint32_t Le = __builtin_amdgcn_perm(t.ll_hi, t.ll_lo, ie);
    const uint32_t He = __builtin_amdgcn_perm(t.lh_hi, t.lh_lo, ie);
    const uint32_t Lo = __builtin_amdgcn_perm(t.ll_hi, t.ll_lo, io);
    const uint32_t Ho = __builtin_amdgcn_perm(t.lh_hi, t.lh_lo, io);
    r[0] = __builtin_amdgcn_perm(He, Le, 0x05010400u);
    r[1] = __builtin_amdgcn_perm(He, Le, 0x07030602u);
    r[2] = __builtin_amdgcn_perm(Ho, Lo, 0x05010400u);
    r[3] = __builtin_amdgcn_perm(Ho, Lo, 0x07030602u);
    bl = __builtin_bit_cast(h8, r);
  }
  {
    // fp16 -1 / +1 have high bytes 0xBC / 0x3C and low bytes 0; selector
    // value 12 yields a zero byte.
    const uint32_t te = w & 0x01010101u;
    const uint32_t to = (w >> 4) & 0x01010101u;
    const uint32_t Te = __builtin_amdgcn_perm(0u, 0x3CBCu, te);
    const uint32_t To = __builtin_amdgcn_perm(0u, 0x3CBCu, to);
    r[0] = __builtin_amdgcn_perm(Te, 0u, 0x050C040Cu);
    r[1] = __builtin_amdgcn_perm(Te, 0u, 0x070C060Cu);
    r[2] = __builtin_amdgcn_perm(To, 0u, 0x050C040Cu);
    r[3] = __builtin_amdgcn_perm(To, 0u, 0x070C060Cu);
    bt = __builtin_bit_cast(h8, r);
  }
}

// Units of 8 consecutive coordinates -> B-operand fragments, from the two
// code words holding their bytes (w0: coordinates 0-3, w1: 4-7, low byte
// first). l_tab holds (c_hi | c_lo << 16) of centroid k at entry k
// (REP = false), or 32 copies of it at entries 32 k + c (REP = true) of
// which lane l reads copy c = l % 32, so that a half-wavefront's lookups fall
// in 32 distinct banks. Each byte's entry offset is one shift and one
// and-or of its word; the lookups are packed into fp16 pairs by byte
// permutes, and the +-1 plane is built as in decode_nib_word.
template <bool REP = false>
__device__ __forceinline__ void decode8(uint32_t w0, uint32_t w1, const uint32_t *l_tab, h8 &bh,
                                        h8 &bl, h8 &bt, uint32_t copy = 0) {
  constexpr int S = REP ? 7 : 2;  // log2 of the entry stride in bytes
  const uint8_t *tab = reinterpret_cast<const uint8_t *>(l_tab);
  const uint32_t w[2] = {w0, w1};
  uint32_t ph[4], pl[4], pt[4];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    uint32_t t[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      // centroid index = bits 8 b + 1 .. 8 b + 7 of the word, moved to bit S
      const int sh = 8 * b + 1 - S;
      const uint32_t off = (sh >= 0 ? w[h] >> sh : w[h] << -sh) & (127u << S);
      t[b] = *reinterpret_cast<const uint32_t *>(tab + (REP ? (off | (copy << 2)) : off));
    }
    ph[2 * h] = __builtin_amdgcn_perm(t[1], t[0], 0x05040100u);
    pl[2 * h] = __builtin_amdgcn_perm(t[1], t[0], 0x07060302u);
    ph[2 * h + 1] = __builtin_amdgcn_perm(t[3], t[2], 0x05040100u);
    pl[2 * h + 1] = __builtin_amdgcn_perm(t[3], t[2], 0x07060302u);
    // bit 0 of each byte: 1 -> +1 (0x3C00), 0 -> -1 (0xBC00)
    const uint32_t T = __builtin_amdgcn_perm(0u, 0x3CBCu, w[h] & 0x01010101u);
    pt[2 * h] = __builtin_amdgcn_perm(T, 0u, 0x050C040Cu);
    pt[2 * h + 1] = __builtin_amdgcn_perm(T, 0u, 0x070C060Cu);
  }
  bh = __builtin_bit_cast(h8, ph);
  bl = __builtin_bit_cast(h8, pl);
  bt = __builtin_bit_cast(h8, pt);
}

}  // namespace mfma
}  // namespace tq
