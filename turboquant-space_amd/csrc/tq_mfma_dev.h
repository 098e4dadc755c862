// tq_mfma_dev.h -- device helpers shared by the MFMA kernels (tq_mfma.hip,
// tq_search.hip): fp16 hi/lo scaling constants and the code-word decoders
// that turn packed code units into B-operand fragments.
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

// 8 nibble units in one code word -> B fragments in the order
// (0, 2, 4, 6, 1, 3, 5, 7): bytes of the even (odd) coordinates' values are
// looked up four at a time, then interleaved into fp16 pairs. The c_lo
// fragment is decoded only for spaces that need it (SpaceDev::mfma_lo).
template <bool LO>
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


}  // namespace mfma
}  // namespace tq
