// tq_lut.hip -- HBM-streaming fp32 scan for 4-bit (nibble-packed) codes.
//
// Same estimator as tq_scan.hip (distSearchScalarB4,
// include/turboquant/space_turbo_quant.h:951-983; distBuildScalarB4 :985-1019
// for the symmetric form), arranged so that a lane spends about two VALU
// operations per coordinate while the codes stream from HBM:
//
//  * A few consecutive lanes (V) share one code, each scoring its own
//    16-byte chunk (32 coordinates) of it, so one load instruction reads
//    16*V contiguous bytes per code; the V partial sums meet through DPP.
//  * MSE term through a pair table in LDS: for byte position p (coordinates
//    2p, 2p+1) the 64 entries T_p[lo + 8*hi] = q[2p]*c[lo] + q[2p+1]*c[hi]
//    cover every pair of 3-bit centroid indices. One lookup and one add per
//    code byte. The byte offset 4*(lo + 8*hi) of an entry is obtained for all
//    four bytes of a code word at once as (w & 0xEE..) + (w & 0x0E..) (no
//    carries between bytes), and one v_perm per byte splices it under the
//    table's base.
//  * QJL term as an exact integer dot product: s_q is rounded to a 19-bit
//    fixed-point integer (scale 2^e per query, |S| < 2^18) split into five
//    signed 4-bit digits, and sum_{qjl bit set} S_i is five v_dot8
//    (i4 x i4 -> i32) per code word against the word masked to its sign
//    bits; then dot = (2*T - sum_i S_i) * 2^-e. The rounding error of each
//    S_i is bounded by 2^-19 of max|s_q|, not of s_i, so it weighs most at
//    small D (few, large coordinates): at D = 128 it moves distances by
//    ~7e-8 of (||q||^2 + ||x||^2) on average (15 bits: 1.1e-6 mean,
//    1.04e-5 max over 320k distances, past the 1e-5 bar).
// The plain fp32 path and its 1e-5 bar are unchanged: the per-code epilogue
// is the reference's fp32 expression.
#include "tq_common.h"

#include <type_traits>

namespace tq {
namespace {

// Per query: [e, S_int, 0, 0 | kQjlDigits planes of D/8 words]; word wi of
// plane j holds digit j of the fixed-point s_q of coordinates 8*wi .. 8*wi+7
// (nibble v = coordinate 8wi+v).
constexpr int kQjlDigits = 5;
__host__ __device__ inline long qlimb_stride(int D) { return 4 + kQjlDigits * (D / 8); }

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
    // |s| * 2^e < 2^(4*digits - 2): balanced base-16 digits in [-8, 7] hold it
    const int e = mx > 0.0f ? 4 * kQjlDigits - 2 - ex : 0;
    int32_t *out = ql + q * lstride;
    int ssum = 0;
    for (int wi = threadIdx.x; wi < s.D / 8; wi += blockDim.x) {
      uint32_t L[kQjlDigits] = {};
      for (int v = 0; v < 8; ++v) {
        int S = (int)rintf(ldexpf(sq[8 * wi + v], e));
        ssum += S;
        for (int j = 0; j < kQjlDigits; ++j) {
          const int d = ((S + 8) & 15) - 8;
          L[j] |= (uint32_t)(d & 15) << (4 * v);
          S = (S - d) >> 4;
        }
      }
      for (int j = 0; j < kQjlDigits; ++j) out[4 + j * (s.D / 8) + wi] = (int32_t)L[j];
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

// Launch shape per padded dim D (threads, lanes per code V, register ring
// slots R; R - 1 blocks are in flight while one is computed). V consecutive
// lanes read 16*V contiguous bytes of one code, so one 16-byte load
// instruction touches 64/V codes' cache lines instead of 64. Measured on
// MI355X (10M codes of 524 B, loads only): 64 lanes on 64 codes stream at
// 3.2-3.9 TB/s (L1 tag lookups), 4 lanes per code 5.5 TB/s, 8 lanes 6.0 TB/s,
// fully coalesced 6.4 TB/s. With the table lookups added, V = 2 and 16
// wavefronts per workgroup (4 per SIMD; at D = 1024 the pair tables fill
// 128 KB of LDS, so one workgroup per CU) were fastest over D = 32..1024
// (1-to-N over 10-20M codes: 4.2-4.9 TB/s), except 8 wavefronts at D = 128.
template <int D>
struct LutCfg {
  static constexpr int THREADS = D == 128 ? 512 : 1024;
  static constexpr int V = 2;
  static constexpr int R = THREADS >= 1024 ? 2 : 3;
};

// Work split of one block: SL coordinate slices x G code groups = 8
// wavefronts. A slice is SW = 4V consecutive 32-bit words (16V bytes) of a
// code. Wavefront w takes slice w % SL of the 64 codes of group w / SL in V
// load rounds: in round i, lane l loads 16-byte chunk l % V of the slice of
// code i*64/V + l/V and scores it against that chunk's tables (a lane's
// chunk, and so its tables and digit words, is the same in every round).
// The V lanes of a code add their partial sums with DPP; the SL slice sums
// of a code meet in LDS; one thread per code applies the epilogue.
template <int D>
struct LutShape {
  static constexpr int THREADS = LutCfg<D>::THREADS;
  static constexpr int WAVES = THREADS / 64;
  static constexpr int NW = D / 8;  // 32-bit words of packed indices per code
  static constexpr int V = LutCfg<D>::V < NW / 4 ? LutCfg<D>::V : NW / 4;
  static constexpr int SW = 4 * V;                   // words per slice
  static constexpr int SL = NW / SW;                 // slices per code
  static constexpr int G = WAVES / SL;               // code groups
  static constexpr int CODES = G * 64;               // codes per block
  static_assert(SL >= 1 && SL <= WAVES && G * SL == WAVES, "every wavefront owns a (slice, group)");
  static_assert(V == 1 || V == 2 || V == 4 || V == 8, "lanes per code");
  static constexpr size_t lds_bytes = sizeof(float) * 64 * (D / 2) +  // pair tables
                                      2 * 2 * 4 * (size_t)SL * CODES +  // slice partials
                                      sizeof(int32_t) * kQjlDigits * (D / 8);  // QJL digit words
  static_assert(lds_bytes <= 160 * 1024, "LDS of one CU");
};

// Sum over the V consecutive lanes that share a code (all of them get it).
template <int V>
__device__ __forceinline__ int group_sum_i(int v) {
  if constexpr (V >= 2) v += __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);   // quad_perm [1,0,3,2]
  if constexpr (V >= 4) v += __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);   // quad_perm [2,3,0,1]
  if constexpr (V >= 8) v += __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true);  // row_half_mirror
  return v;
}
template <int V>
__device__ __forceinline__ float group_sum_f(float v) {
  if constexpr (V >= 2) v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
  if constexpr (V >= 4) v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
  if constexpr (V >= 8) v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, true));
  return v;
}

template <bool SYM, int D>
__global__ __launch_bounds__(LutShape<D>::THREADS) void k_lut_nib(SpaceDev s, const float *__restrict__ prep,
                                                         const int32_t *__restrict__ ql, long m,
                                                         const uint8_t *__restrict__ codes, long n,
                                                         float *__restrict__ out) {
  using Sh = LutShape<D>;
  constexpr int V = Sh::V, SW = Sh::SW, SL = Sh::SL, CODES = Sh::CODES;
  constexpr int R = LutCfg<D>::R, THREADS = Sh::THREADS;
  // (D/2) tables of 64 floats; fixed-size arrays keep LDS addresses
  // compile-time constants. part_*: per-slice partial sums, double-buffered.
  // ql_lds: the query's QJL digit words.
  __shared__ float tab[(D / 2) * 64];
  __shared__ float part_m[2][SL][CODES];
  __shared__ int part_q[2][SL][CODES];
  __shared__ __align__(16) int32_t ql_lds[kQjlDigits * (D / 8)];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int sl = wave % SL, grp = wave / SL;
  const int chunk = lane % V;
  const bool reducer = wave < Sh::G;  // wavefront-uniform: threads tid < CODES finish codes
  const long pstride = prep_stride(s.D);
  const long lstride = qlimb_stride(s.D);
  const long nblocks = (n + CODES - 1) / CODES;
  const long G = gridDim.x;
  const int w0 = sl * SW + 4 * chunk;  // this lane's first word in every code
  // Byte address of the first table of this lane's chunk (a multiple of 256,
  // so v_perm can splice an entry offset into its low byte).
  const uint32_t tbase = (uint32_t)(w0 * 1024);
  float cent[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) cent[k] = s.cent[k < s.L ? k : 0];

  for (long qi = 0; qi < m; ++qi) {
    const float *pq = prep + qi * pstride;
    const int32_t *L = ql + qi * lstride;
    __syncthreads();  // the previous query is done with the tables
    for (int e = tid; e < (D / 2) * 64; e += THREADS) {
      const int p = e >> 6, lo = e & 7, hi = (e >> 3) & 7;
      tab[e] = fmaf(pq[2 * p + 1], cent[hi], pq[2 * p] * cent[lo]);
    }
    if constexpr (!SYM)
      for (int e = tid; e < kQjlDigits * (D / 8); e += THREADS) ql_lds[e] = L[4 + e];
    __syncthreads();
    const float qns = pq[2 * s.D];
    const float qn = pq[2 * s.D + 1];
    const int qe = SYM ? 0 : L[0];
    const int qsum = SYM ? 0 : L[1];
    float *orow = out + qi * n;
    // This lane's digit words (4 words x 5 digits), in registers for the scan.
    int32_t lw[SYM ? 1 : 4][kQjlDigits];
    if constexpr (!SYM) {
#pragma unroll
      for (int j = 0; j < kQjlDigits; ++j) {
        const int4 v = *reinterpret_cast<const int4 *>(&ql_lds[j * (D / 8) + w0]);
        lw[0][j] = v.x;
        lw[1][j] = v.y;
        lw[2][j] = v.z;
        lw[3][j] = v.w;
      }
    }

    // Ring of register buffers: slot k holds a block's chunks (V rounds) and
    // the trailers [norm, gamma, sigma] of the codes this thread finishes.
    // Every load is unconditional (block index clamped to the last block,
    // code index to the last code), trailers first, so the in-order load
    // counter lets the newer blocks stay in flight while the oldest is used.
    uint4 ring[R][V];
    float tx[R], tg[R], tsg[R];
    auto load = [&](auto slot, long blk) {
      constexpr int k = decltype(slot)::value;
      const long bb = blk < nblocks ? blk : nblocks - 1;
      __builtin_amdgcn_sched_barrier(0);
      {
        // Reducing wavefronts take their codes' trailers; the others read one
        // shared trailer (a single cache line), so every wavefront issues the
        // same loads. D >= 32: code sizes D/2 + 12 are multiples of 4.
        long rc = bb * CODES + tid;
        rc = rc < n ? rc : n - 1;
        const float *mp = reinterpret_cast<const float *>(codes + (reducer ? rc : 0L) * (long)s.cs + s.packed);
        tx[k] = mp[0];
        tg[k] = mp[1];
        tsg[k] = mp[2];
      }
#pragma unroll
      for (int i = 0; i < V; ++i) {
        long code = bb * CODES + grp * 64 + i * (64 / V) + lane / V;
        code = code < n ? code : n - 1;
        const uint32_t *src = reinterpret_cast<const uint32_t *>(codes + code * (long)s.cs + 4 * w0);
        ring[k][i] = make_uint4(src[0], src[1], src[2], src[3]);
      }
      __builtin_amdgcn_sched_barrier(0);
    };

    long b = blockIdx.x;
    asm volatile("" : "+s"(b));  // keeps the first blocks' loads inside the query loop
    int buf = 0;
    // One block from ring slot `slot`: partial sums, slot reloaded with block
    // b + R*G, barrier, the reducing threads finish block b.
    auto step = [&](auto slot) {
      constexpr int k = decltype(slot)::value;
#pragma unroll
      for (int i = 0; i < V; ++i) {
        float am0 = 0.0f, am1 = 0.0f;
        int aq0 = 0, aq1 = 0, aq2 = 0, aq3 = 0, aq4 = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint4 q4 = ring[k][i];
          const uint32_t w = u == 0 ? q4.x : u == 1 ? q4.y : u == 2 ? q4.z : q4.w;
          const uint32_t m2 = (w & 0xEEEEEEEEu) + (w & 0x0E0E0E0Eu);
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const uint32_t addr = __builtin_amdgcn_perm(tbase, m2, 0x07060500u | (uint32_t)kk);
            const float v = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(tab) + addr +
                                                             1024 * u + 256 * kk);
            if (kk & 1)
              am1 += v;
            else
              am0 += v;
          }
          if constexpr (!SYM) {
            const int t = (int)(w & 0x11111111u);
            aq0 = __builtin_amdgcn_sdot8(t, lw[u][0], aq0, false);
            aq1 = __builtin_amdgcn_sdot8(t, lw[u][1], aq1, false);
            aq2 = __builtin_amdgcn_sdot8(t, lw[u][2], aq2, false);
            aq3 = __builtin_amdgcn_sdot8(t, lw[u][3], aq3, false);
            aq4 = __builtin_amdgcn_sdot8(t, lw[u][4], aq4, false);
          }
        }
        const float am = group_sum_f<V>(am0 + am1);
        const int aq = SYM ? 0 : group_sum_i<V>(aq0 + 16 * aq1 + 256 * aq2 + 4096 * aq3 + 65536 * aq4);
        if (chunk == i % V) {
          const int slot_i = grp * 64 + i * (64 / V) + lane / V;
          part_m[buf][sl][slot_i] = am;
          part_q[buf][sl][slot_i] = aq;
        }
      }
      const float xn = tx[k], gamma = tg[k], sigma = tsg[k];
      load(slot, b + (long)R * G);
      __syncthreads();  // partials of block b complete
      if (reducer) {
        const long rcode = b * CODES + tid;
        float am = 0.0f;
        int T = 0;
#pragma unroll
        for (int j = 0; j < SL; ++j) {
          am += part_m[buf][j][tid];
          T += part_q[buf][j][tid];
        }
        const float ip_mse = am * sigma;
        float corr = 0.0f;
        if constexpr (!SYM) {
          const float dot = ldexpf((float)(2 * T - qsum), -qe);
          corr = s.scale * gamma * dot;
        }
        const float ip = (ip_mse + corr) * xn * qn;
        const float d = qns + xn * xn - 2.0f * ip;
        if (rcode < n) orow[rcode] = d > 0.0f ? d : 0.0f;
      }
      b += G;
      buf ^= 1;
    };

    static_assert(R == 2 || R == 3, "the unrolled ring below is written for two or three slots");
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    using S2 = std::integral_constant<int, R == 3 ? 2 : 0>;
    if (b < nblocks) {
      // Blocks of this workgroup: b, b + G, ... < nblocks. The loop runs
      // whole turns of the ring and has no exit mid-turn, so the load
      // counter sees the same order of pending loads on every entry to its
      // header; the blocks of a last partial turn follow it.
      const long nb = (nblocks - 1 - b) / G + 1;
      load(S0{}, b);
      load(S1{}, b + G);
      if constexpr (R == 3) {
        load(S2{}, b + 2 * G);
        for (long t = nb / 3; t > 0; --t) {
          step(S0{});
          step(S1{});
          step(S2{});
        }
        if (nb % 3 >= 1) step(S0{});
        if (nb % 3 >= 2) step(S1{});
      } else {
        for (long t = nb / 2; t > 0; --t) {
          step(S0{});
          step(S1{});
        }
        if (nb % 2) step(S0{});
      }
    }
  }
}

// Wide form, D = 512 and 1024. In the narrow form above the lanes that score
// the same chunk of different codes read random entries of the same 64-entry
// table, so a wavefront's lookup lands on a few banks several times over.
// Here every lane of a group of LPC lanes owns its tables: LPC lanes score
// one code, lane ll = l % LPC taking the NW code words [NW ll, NW ll + NW)
// (D = 1024: 32 lanes x 16 bytes, so one load instruction reads two codes;
// D = 512: 64 lanes x 4 bytes). The tables are laid out so that every entry
// lane ll reads sits in LDS bank ll (+ 32 for odd words at D = 1024): the
// entry of indices (lo, hi) in the table of byte kk of word u of lane ll's
// slice is at byte address
//   4 ll + 128 (u & 1) + 256 (b & 0xEE) + 256 (kk & 1) + 4096 (kk >> 1) + 65536 (u >> 1),
// b & 0xEE = 2 lo + 32 hi being the code byte without its QJL bits. Within a
// half-wavefront no two lanes share a bank (SQ_LDS_BANK_CONFLICT reads 0 on
// MI355X). One v_perm puts byte kk of w & 0xEEEEEEEE into byte 1 of the
// lane's base address (the bits 8 and 12 it leaves zero take kk, with u & 1,
// as the instruction's offset).
// A wavefront scores 8 consecutive codes per round. Code bytes and trailers
// are read once, with non-temporal loads (on MI355X a plain 16-byte stream
// of 5.24 GB reads at 5.9 TB/s, a non-temporal one at 6.3-6.5 TB/s). The 8
// codes' per-lane partial sums meet through a transposed reduction (permlane
// swaps, a DPP row rotate, then DPP within 8 lanes), after which lanes
// 8c .. 8c + 7 hold the sums of one code and lane 8c writes its distance.
// The tables and QJL digits are the narrow form's, so both give the same
// terms; only the order of the fp32 MSE additions differs.
constexpr int kWideCodes = 8;  // codes per wavefront round

template <int D>
struct WideCfg {
  static_assert(D == 512 || D == 1024, "wide form: D = 512 or 1024");
  static constexpr int LPC = D == 1024 ? 32 : 64;   // lanes per code
  static constexpr int NW = D / 8 / LPC;            // code words per lane
  static constexpr int CPI = 64 / LPC;              // codes per load instruction
  static constexpr int NL = kWideCodes / CPI;       // load instructions per round
  static constexpr int THREADS = D == 1024 ? 1024 : 512;
  static constexpr int R = 3;                       // ring slots
  static constexpr size_t lds_bytes = sizeof(float) * 64 * (D / 2);
};

__device__ __forceinline__ float wide_swap_add32(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ int wide_swap_add32(int a, int b) {
  const auto r = __builtin_amdgcn_permlane32_swap((unsigned)a, (unsigned)b, false, false);
  return (int)r[0] + (int)r[1];
}
__device__ __forceinline__ float wide_swap_add16(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ int wide_swap_add16(int a, int b) {
  const auto r = __builtin_amdgcn_permlane16_swap((unsigned)a, (unsigned)b, false, false);
  return (int)r[0] + (int)r[1];
}
// x + (lane ^ 8's y), lane ^ 8 within a row of 16 being DPP row_ror:8.
__device__ __forceinline__ float wide_add_ror8(float x, float y) {
  return x + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(y), 0x128, 0xF, 0xF, true));
}
__device__ __forceinline__ int wide_add_ror8(int x, int y) {
  return x + __builtin_amdgcn_mov_dpp(y, 0x128, 0xF, 0xF, true);
}
__device__ __forceinline__ float wide_group8(float v) { return group_sum_f<8>(v); }
__device__ __forceinline__ int wide_group8(int v) { return group_sum_i<8>(v); }

// Sums over the lanes of each code of the round's per-lane values v[i] (code
// CPI i + l / LPC); the result in lane l is the total of code
// wide_code(l), the same in the 8 lanes 8c .. 8c + 7. Each step halves the
// values per lane: a permlane swap pairs value i with value i + half, lanes
// of one half (of the wavefront, or of each 32-lane row pair) keeping the
// first and the others the second, each summed over both halves; a row
// rotate by 8 does the same for lane bit 3; then the 8 lanes of a code add
// up (DPP xor 1, xor 2, half mirror).
template <int LPC, typename T>
__device__ __forceinline__ T wide_reduce(const T (&v)[kWideCodes / (64 / LPC)], int lane) {
  constexpr int NL = kWideCodes / (64 / LPC);
  T t[2];
  if constexpr (NL == 8) {  // 64 lanes per code: halves of the wavefront, then of 32-lane rows
    T s[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = wide_swap_add32(v[i], v[i + 4]);
#pragma unroll
    for (int i = 0; i < 2; ++i) t[i] = wide_swap_add16(s[i], s[i + 2]);
  } else {  // 32 lanes per code: halves of each 32-lane row pair
    static_assert(NL == 4, "values per lane");
#pragma unroll
    for (int i = 0; i < 2; ++i) t[i] = wide_swap_add16(v[i], v[i + 2]);
  }
  const bool b3 = (lane >> 3) & 1;
  return wide_group8(wide_add_ror8(b3 ? t[1] : t[0], b3 ? t[0] : t[1]));
}
// The code (0 .. 7 within the round) whose sums lane l holds after wide_reduce.
template <int LPC>
__device__ __forceinline__ int wide_code(int lane) {
  if constexpr (LPC == 64) return (lane >> 3) & 7;
  else return 2 * ((lane >> 3) & 3) + (lane >> 5);
}

template <bool SYM, int D>
__global__ __launch_bounds__(WideCfg<D>::THREADS) void k_lut_wide(
    SpaceDev s, const float *__restrict__ prep, const int32_t *__restrict__ ql, long m,
    const uint8_t *__restrict__ codes, long n, float *__restrict__ out) {
  using Cf = WideCfg<D>;
  constexpr int LPC = Cf::LPC, NW = Cf::NW, CPI = Cf::CPI, NL = Cf::NL, R = Cf::R;
  constexpr int THREADS = Cf::THREADS, WAVES = THREADS / 64, CPR = kWideCodes;
  __shared__ float tab[(D / 2) * 64];
  typedef uint32_t slice_t __attribute__((ext_vector_type(NW), aligned(4)));
  typedef float trail_t __attribute__((ext_vector_type(3), aligned(4)));

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int ll = lane % LPC, g = lane / LPC;  // slice and code of this lane in a load
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ec = wide_code<LPC>(lane);  // the code of a round whose trailer and output this lane has
  const long pstride = prep_stride(s.D);
  const long lstride = qlimb_stride(s.D);
  const long nrounds = (n + CPR - 1) / CPR;
  const long W = (long)gridDim.x * WAVES;
  const long r0 = (long)blockIdx.x * WAVES + wave;
  const uint32_t base0 = (uint32_t)(4 * ll), base1 = base0 + 65536u;
  float cent[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) cent[k] = s.cent[k < s.L ? k : 0];

  for (long qi = 0; qi < m; ++qi) {
    const float *pq = prep + qi * pstride;
    const int32_t *L = ql + qi * lstride;
    __syncthreads();  // the previous query is done with the tables
    for (int t = tid; t < (D / 2) * 64; t += THREADS) {
      // t = l + LPC (e + 64 j): lane slice l, entry e, byte j = 4u + kk of
      // the slice (consecutive threads write consecutive banks).
      const int l = t % LPC, e = (t / LPC) & 63, j = t / (64 * LPC);
      const int p = 4 * NW * l + j;  // byte position in the code
      const int lo = e & 7, hi = e >> 3, kk = j & 3, u = j >> 2;
      tab[(4 * l + 128 * (u & 1) + 65536 * (u >> 1) + 512 * lo + 8192 * hi + 256 * (kk & 1) +
           4096 * (kk >> 1)) >> 2] = fmaf(pq[2 * p + 1], cent[hi], pq[2 * p] * cent[lo]);
    }
    __syncthreads();
    const float qns = pq[2 * s.D];
    const float qn = pq[2 * s.D + 1];
    const int qe = SYM ? 0 : L[0];
    const int qsum = SYM ? 0 : L[1];
    float *orow = out + qi * n;
    int32_t lw[SYM ? 1 : NW][kQjlDigits];
    if constexpr (!SYM) {
#pragma unroll
      for (int u = 0; u < NW; ++u)
#pragma unroll
        for (int j = 0; j < kQjlDigits; ++j) lw[u][j] = L[4 + j * (D / 8) + NW * ll + u];
    }

    // Ring slot k: this lane's slices of the round's codes and the trailer
    // of code ec. Every load is unconditional (round clamped to the last one,
    // code to the last code), so the in-order load counter lets the newer
    // rounds stay in flight while the oldest is scored.
    slice_t ring[R][NL];
    trail_t trl[R];
    auto load = [&](auto slot, long r) {
      constexpr int k = decltype(slot)::value;
      const long rr = r < nrounds ? r : nrounds - 1;
      __builtin_amdgcn_sched_barrier(0);
      {
        long c = rr * CPR + ec;
        c = c < n ? c : n - 1;
        trl[k] = __builtin_nontemporal_load(
            reinterpret_cast<const trail_t *>(codes + c * (long)s.cs + s.packed));
      }
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        long c = rr * CPR + CPI * i + g;
        c = c < n ? c : n - 1;
        ring[k][i] = __builtin_nontemporal_load(
            reinterpret_cast<const slice_t *>(codes + c * (long)s.cs + 4 * NW * ll));
      }
      __builtin_amdgcn_sched_barrier(0);
    };

    long r = r0;
    asm volatile("" : "+s"(r));  // keeps the first rounds' loads inside the query loop
    auto step = [&](auto slot) {
      constexpr int k = decltype(slot)::value;
      float am[NL];
      int aq[NL];
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        uint32_t w[NW];
#pragma unroll
        for (int u = 0; u < NW; ++u) w[u] = ring[k][i][u];
        // MSE: two chains of lookups (even and odd bytes), each started by
        // its first entry.
        float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
        for (int u = 0; u < NW; ++u) {
          const uint32_t mw = w[u] & 0xEEEEEEEEu;
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const uint32_t addr =
                __builtin_amdgcn_perm(u >> 1 ? base1 : base0, mw, 0x07060004u | ((uint32_t)kk << 8));
            const float v = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(tab) + addr +
                                                             128 * (u & 1) + 256 * (kk & 1) + 4096 * (kk >> 1));
            if (u == 0 && kk < 2)
              (kk ? a1 : a0) = v;
            else if (kk & 1)
              a1 += v;
            else
              a0 += v;
          }
        }
        am[i] = a0 + a1;
        // QJL: sum_j 16^j sum_u dot8(t_u, digit j), Horner from the top digit.
        if constexpr (!SYM) {
          int t[NW];
#pragma unroll
          for (int u = 0; u < NW; ++u) t[u] = (int)(w[u] & 0x11111111u);
          int acc = 0;
#pragma unroll
          for (int j = kQjlDigits - 1; j >= 0; --j) {
            if (j < kQjlDigits - 1) acc <<= 4;
#pragma unroll
            for (int u = 0; u < NW; ++u) acc = __builtin_amdgcn_sdot8(t[u], lw[u][j], acc, false);
          }
          aq[i] = acc;
        } else {
          aq[i] = 0;
        }
      }
      const float xn = trl[k].x, gamma = trl[k].y, sigma = trl[k].z;
      const long rc = r * CPR + ec;
      load(slot, r + (long)R * W);
      const float msum = wide_reduce<LPC>(am, lane);
      const int T = SYM ? 0 : wide_reduce<LPC>(aq, lane);
      if ((lane & 7) == 0 && rc < n) {
        const float ip_mse = msum * sigma;
        float corr = 0.0f;
        if constexpr (!SYM) {
          const float dot = ldexpf((float)(2 * T - qsum), -qe);
          corr = s.scale * gamma * dot;
        }
        const float ip = (ip_mse + corr) * xn * qn;
        const float d = qns + xn * xn - 2.0f * ip;
        orow[rc] = d > 0.0f ? d : 0.0f;
      }
      r += W;
    };

    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    using S2 = std::integral_constant<int, 2>;
    static_assert(R == 3, "the unrolled ring below is written for three slots");
    if (r < nrounds) {
      // Whole turns of the ring, then the rounds of a last partial turn.
      const long nr = (nrounds - 1 - r) / W + 1;
      load(S0{}, r);
      load(S1{}, r + W);
      load(S2{}, r + 2 * W);
      for (long t = nr / 3; t > 0; --t) {
        step(S0{});
        step(S1{});
        step(S2{});
      }
      if (nr % 3 >= 1) step(S0{});
      if (nr % 3 >= 2) step(S1{});
    }
  }
}

// Compute units of the current device (asked per launch: cheap, and correct
// for every device and thread).
int cu_count() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return cus;
}

}  // namespace

bool lut_supported(const SpaceDev &s) { return s.nib && s.D >= 32 && s.D <= 1024; }

namespace {

template <bool SYM, int D>
hipError_t launch_lut_d(const SpaceDev &s, const float *prep, const int32_t *ql, long m,
                        const uint8_t *codes, long n, float *out, hipStream_t st) {
  constexpr size_t lds = LutShape<D>::lds_bytes;
  const long per_cu = (long)((160 * 1024) / lds) < 4 ? (long)((160 * 1024) / lds) : 4;
  long blocks = (long)cu_count() * (per_cu > 0 ? per_cu : 1);
  const long need = (n + LutShape<D>::CODES - 1) / LutShape<D>::CODES;
  if (blocks > need) blocks = need;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((k_lut_nib<SYM, D>), dim3((unsigned)blocks), dim3(LutShape<D>::THREADS), 0, st, s,
                     prep, ql, m, codes, n, out);
  return hipGetLastError();
}

template <bool SYM, int D>
hipError_t launch_wide_d(const SpaceDev &s, const float *prep, const int32_t *ql, long m,
                         const uint8_t *codes, long n, float *out, hipStream_t st) {
  using Cf = WideCfg<D>;
  const long per_cu = (long)((160 * 1024) / Cf::lds_bytes);
  long blocks = (long)cu_count() * (per_cu > 0 ? per_cu : 1);
  const long rounds = (n + kWideCodes - 1) / kWideCodes;
  const long need = (rounds + Cf::THREADS / 64 - 1) / (Cf::THREADS / 64);
  if (blocks > need) blocks = need;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((k_lut_wide<SYM, D>), dim3((unsigned)blocks), dim3(Cf::THREADS), 0, st, s,
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
    case 512: return launch_wide_d<SYM, 512>(s, prep, ql, m, codes, n, out, st);
    case 1024: return launch_wide_d<SYM, 1024>(s, prep, ql, m, codes, n, out, st);
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
