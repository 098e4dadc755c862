// tq_mfma.hip -- distance_m_to_n as a matrix-core contraction on gfx950.
//
// The asymmetric estimator (space_turbo_quant.h:659-686, :951-983) is, over a
// batch, two dense contractions between the queries and the decoded codes:
//   MSE[q, j] = sum_i q_rot[q, i] * c[idx_{j,i}]
//   QJL[q, j] = sum_i s_q[q, i]  * (qjl_{j,i} ? +1 : -1)
// followed by a per-pair fp32 epilogue with the code's norm/gamma/sigma and
// the query's norm (identical to the plain path's).
//
// Precision: fp16 does not hold the queries to fp32 accuracy, so they are
// split into hi + lo fp16 parts (v = hi + lo with hi = fp16(v),
// lo = fp16(v - hi), ~22 significant bits) and the MSE term is accumulated as
// q_hi*c + q_lo*c. For nibble codes the space's centroid scale makes every
// scaled centroid nearly exact in fp16 (within 2^-15; SpaceDev::cscale), so
// c is one fp16 value; byte codes (up to 128 centroids) split c as well and
// add q_hi*c_lo (the dropped q_lo*c_lo is ~2^-22 relative). The +-1 QJL
// factors are exact in fp16 and the QJL term enters the distance with a
// factor scale*gamma, so s_q is used at fp16 (one pass). Its rounding is the
// largest error of the 4-bit form: ~5e-6 std of (||q||^2 + ||x||^2) at
// D = 128, 2.3e-5 at most over 320k distances, less at larger D (smaller
// scale) and for byte codes (small gamma). Everything accumulates in fp32
// (v_mfma_f32_32x32x16_f16). Queries are scaled by 2^10 and centroids by
// cscale (2^8 for byte codes) so the lo parts stay in fp16's normal range;
// the epilogue removes the scales. Measured error against the fp64
// restatement of the estimator is held below 1e-4 of (||q||^2 + ||x||^2).
//
// Tiling: a workgroup (4 wavefronts) owns 64 queries x 256 codes; each
// wavefront owns 64 queries x 64 codes as 2 x 2 blocks of 32 x 32 MFMA
// outputs, walking the padded dim in steps of 16. The query planes (written
// once per batch by k_split_queries) and the code bytes are staged through
// LDS in chunks of 32 coordinates (see k_mfma); a lane decodes the code
// bytes it contributes per step (8 coordinates of one code) in registers:
// nibble codes through v_perm byte tables of the 8 centroids, byte codes
// through a (c_hi | c_lo << 16) table in LDS.
#include "tq_mfma.h"
#include "tq_mfma_dev.h"

namespace tq {
namespace {

using namespace mfma;

// planes: [3][m][D] fp16 = q_hi, q_lo, s_hi (scaled by 2^10). With
// `nib_order`, each group of 8 coordinates is stored in the order the nibble
// decoder emits them: 0, 2, 4, 6, 1, 3, 5, 7 (the contraction does not care
// about the order within a group as long as both operands agree).
__global__ __launch_bounds__(256) void k_split_queries(SpaceDev s, const float *__restrict__ prep,
                                                       long m, _Float16 *__restrict__ planes,
                                                       int nib_order) {
  const long stride = prep_stride(s.D);
  const long mD = m * (long)s.D;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < mD;
       e += (long)gridDim.x * blockDim.x) {
    const long q = e / s.D;
    const int slot = (int)(e - q * s.D);
    const int g = slot & ~7, k = slot & 7;
    const int i = nib_order ? g + (k < 4 ? 2 * k : 2 * (k - 4) + 1) : slot;
    const float v = prep[q * stride + i] * kQScale;
    const float w = prep[q * stride + s.D + i] * kQScale;
    const _Float16 vh = (_Float16)v, wh = (_Float16)w;  // s_q needs no lo part
    planes[e] = vh;
    planes[mD + e] = (_Float16)(v - (float)vh);
    planes[2 * mD + e] = wh;
  }
}

// One workgroup: 64 queries x 256 codes. The padded dim is walked in chunks
// of 32 coordinates; per chunk the query planes of the 64 queries
// (3 planes x 64 rows x 32 fp16) and the code bytes of the 256 codes are
// staged in LDS by all 256 threads (double-buffered LDS, loads issued
// ring_slots() chunks ahead into registers, one barrier per chunk). Rows are
// padded (A: 80 B, codes: 20 / 36 B) so the fragment reads of a wavefront
// fall on distinct LDS banks.
constexpr int kTQ = 64;            // queries per workgroup
constexpr int kTN = 256;           // codes per workgroup (64 per wavefront)
constexpr int kKC = 32;            // coordinates per chunk
constexpr int kARow = kKC + 8;     // padded A row, in halves (80 B)
constexpr int kPlanes = 3;         // q_hi, q_lo, s_hi
// Chunks whose loads are in flight: 3 for nibble codes, 2 for byte codes
// (twice the code bytes per chunk, and the register budget of two
// wavefronts per SIMD).
template <bool NIB>
constexpr int ring_slots() { return NIB ? 3 : 2; }
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// The same at 4-byte alignment (code bytes: codes are 4-byte aligned; the
// padded LDS code rows are too).
typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

template <bool NIB>
struct MfmaShape {
  static constexpr int CB = NIB ? kKC / 2 : kKC;  // code bytes per chunk
  static constexpr int CROW = CB + 4;              // padded code row, bytes
  static constexpr int CPARTS = CB / 16;           // 16-byte loads per code
};

// Two workgroups per CU (two wavefronts per SIMD) so one wavefront's MFMA
// chain overlaps the other's decode, LDS traffic and barrier waits. QT = 64
// queries per workgroup (2 x 2 blocks of 32 x 32 per wavefront) or 32 (1 x 2
// blocks) for batches of at most 32 queries: every output element depends
// only on its own query row and code column, so both give the same bits.
// QJL = false drops the s_q plane and its pass (queries with s_q = 0: the
// symmetric MSE estimator, whose "queries" are decoded codes); the QJL term
// is then exactly zero and the distances are what the three-pass form gives.
template <bool NIB, bool LO, int QT, bool QJL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mfma(SpaceDev s, const _Float16 *__restrict__ planes,
                                              const float *__restrict__ prep, long m,
                                              const uint8_t *__restrict__ codes, long n,
                                              float *__restrict__ out) {
  using Sh = MfmaShape<NIB>;
  static_assert(QT == 32 || QT == 64, "query tile");
  constexpr int RB = QT / 32;                                  // 32-row blocks per wavefront
  constexpr int CROW = Sh::CROW, CPARTS = Sh::CPARTS;
  constexpr int NPL = QJL ? kPlanes : 2;                       // query planes read
  constexpr int AUNITS = NPL * QT * (kKC / 8);                 // 16-byte A units per chunk
  constexpr int ALOADS = (AUNITS + 255) / 256;                 // per thread (QT = 32: the
                                                               // last unit is loaded twice)
  constexpr int CLOADS = kTN * CPARTS / 256;         // 16-byte code loads per thread
  __shared__ _Float16 a_lds[2][kPlanes][QT][kARow];
  __shared__ uint32_t c_lds[2][kTN * CROW / 4];
  __shared__ uint32_t l_tab[256];
  __shared__ float l_qn[QT][2];  // q_norm_sq, q_norm of the tile's queries

  const int tid = threadIdx.x;
  if (tid < QT) {
    const long q = (long)blockIdx.y * QT + tid;
    const float *pq = prep + (q < m ? q : 0L) * prep_stride(s.D);
    l_qn[tid][0] = pq[2 * s.D];
    l_qn[tid][1] = pq[2 * s.D + 1];
  }
  static_assert(NIB || LO, "byte codes carry c_lo");
  for (int i = tid; i < 256; i += blockDim.x) {
    _Float16 ch, cl;
    cent_h(s, i, ch, cl);
    l_tab[i] = h_bits(ch) | (h_bits(cl) << 16);
  }

  PermTab ptab{};
  if constexpr (NIB) {
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
    ptab = PermTab{bytes[0][0], bytes[0][1], bytes[1][0], bytes[1][1],
                   bytes[2][0], bytes[2][1], bytes[3][0], bytes[3][1]};
  }

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 31, kg = lane >> 5;
  const long q0 = (long)blockIdx.y * QT;
  const long j0 = (long)blockIdx.x * kTN;
  const long mD = m * (long)s.D;
  const int nchunks = s.D / kKC;

  // Per-thread load sources and LDS destinations, fixed for the whole walk
  // over the padded dim; a chunk only adds its coordinate offset. Rows of
  // queries past m and codes past n read row 0 / code 0 instead: every load
  // is unconditional, and those rows and columns of the accumulators are
  // never written out (an MFMA output element depends only on its own A row
  // and B column).
  const _Float16 *asrc[ALOADS];
  int adst[ALOADS];
#pragma unroll
  for (int i = 0; i < ALOADS; ++i) {
    const int idx = i * 256 + tid < AUNITS ? i * 256 + tid : AUNITS - 1;
    const int p = idx / (QT * kKC / 8);
    const int row = (idx / (kKC / 8)) % QT;
    const int part = idx % (kKC / 8);
    const long q = q0 + row;
    asrc[i] = planes + p * mD + (q < m ? q : 0L) * s.D + 8 * part;
    adst[i] = (p * QT + row) * kARow + 8 * part;
  }
  const uint8_t *csrc[CLOADS];
  int cdst[CLOADS];
#pragma unroll
  for (int i = 0; i < CLOADS; ++i) {
    const int idx = i * 256 + tid;
    const int c = idx / CPARTS, part = idx % CPARTS;
    const long j = j0 + c;
    csrc[i] = codes + (j < n ? j : 0L) * (long)s.cs + 16 * part;
    cdst[i] = (c * CROW + 16 * part) / 4;
  }

  // Register ring: slot k holds the query-plane and code loads of one chunk
  // (a group). Groups are issued three chunks before they are stored to LDS,
  // so the HBM latency of the code bytes is covered by three chunks of
  // matrix work; the hardware's load counter completes in order, so a group
  // is always issued before every group needed after it, and the store of a
  // chunk waits for its own group only.
  constexpr int kRing = ring_slots<NIB>();
  u32x4 ring_a[kRing][ALOADS], ring_c[kRing][CLOADS];
  auto load_group = [&](auto slot, int ch) __attribute__((always_inline)) {
    constexpr int k = decltype(slot)::value;
    const int cc = ch < nchunks ? ch : nchunks - 1;
    const int k0 = cc * kKC;
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < ALOADS; ++i) ring_a[k][i] = *reinterpret_cast<const u32x4 *>(asrc[i] + k0);
#pragma unroll
    for (int i = 0; i < CLOADS; ++i)
      ring_c[k][i] = *reinterpret_cast<const u32x4a4 *>(csrc[i] + (NIB ? k0 / 2 : k0));
    __builtin_amdgcn_sched_barrier(0);
  };
  auto store_group = [&](auto slot, int b) __attribute__((always_inline)) {
    constexpr int k = decltype(slot)::value;
    _Float16 *abuf = &a_lds[b][0][0][0];
#pragma unroll
    for (int i = 0; i < ALOADS; ++i) *reinterpret_cast<u32x4 *>(abuf + adst[i]) = ring_a[k][i];
#pragma unroll
    for (int i = 0; i < CLOADS; ++i) *reinterpret_cast<u32x4a4 *>(&c_lds[b][cdst[i]]) = ring_c[k][i];
  };

  f16v acc_m[RB][2], acc_q[RB][2];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        acc_m[r][c][v] = 0.0f;
        acc_q[r][c][v] = 0.0f;
      }

  // One chunk: the matrix work on LDS buffer ch & 1, then chunk ch + 1 from
  // ring slot `slot` into the other buffer, the slot reloaded with chunk
  // ch + 1 + kRing, one barrier.
  int ch = 0;
  // Inlined at each use: the ring arrays must stay in registers.
  auto step = [&](auto slot) __attribute__((always_inline)) {
    const int b = ch & 1;
#pragma unroll
    for (int kb = 0; kb < kKC / 16; ++kb) {
      h8 a[RB][kPlanes];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int p = 0; p < kPlanes; ++p)
          a[r][p] = *reinterpret_cast<const h8 *>(&a_lds[b][p][32 * r + lr][kb * 16 + 8 * kg]);
      h8 bh[2], bl[2], bt[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int code = wave * 64 + 32 * c + lr;
        if constexpr (NIB) {
          decode_nib_word<LO>(c_lds[b][(code * CROW + kb * 8 + 4 * kg) / 4], ptab, bh[c], bl[c], bt[c]);
        } else {
          uint32_t u[8];
          const uint32_t w0 = c_lds[b][(code * CROW + kb * 16 + 8 * kg) / 4];
          const uint32_t w1 = c_lds[b][(code * CROW + kb * 16 + 8 * kg) / 4 + 1];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            u[v] = (w0 >> (8 * v)) & 0xFFu;
            u[4 + v] = (w1 >> (8 * v)) & 0xFFu;
          }
          decode8(u, l_tab, bh[c], bl[c], bt[c]);
        }
      }
      // Each pass over all four 32x32 blocks before the next, so consecutive
      // MFMAs never wait on the accumulator they just wrote.
#define TQ_PASS(ACC, A, B)                                                               \
  _Pragma("unroll") for (int r = 0; r < RB; ++r)                                         \
      _Pragma("unroll") for (int c = 0; c < 2; ++c) ACC[r][c] =                          \
          __builtin_amdgcn_mfma_f32_32x32x16_f16(a[r][A], B[c], ACC[r][c], 0, 0, 0);
      TQ_PASS(acc_m, 0, bh)
      if constexpr (QJL) {
        TQ_PASS(acc_q, 2, bt)
      }
      TQ_PASS(acc_m, 1, bh)
      if constexpr (LO) {
        TQ_PASS(acc_m, 0, bl)
      }
#undef TQ_PASS
    }
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
  // partial turn. Chunks past the end are reloads of the last one, never
  // stored.
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

  // Epilogue: output element v of a 32x32 block sits at row
  // 8*(v/4) + 4*kg + (v%4), column lr. The per-code factors are formed once;
  // the expression is the reference's ip = (sigma*mse + scale*gamma*dot)*|x|*|q|,
  // d = max(0, |q|^2 + |x|^2 - 2*ip).
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const long j = j0 + wave * 64 + 32 * c + lr;
    if (j >= n) continue;
    const float *meta = reinterpret_cast<const float *>(codes + j * (long)s.cs + s.packed);
    const float xn = meta[0], gamma = meta[1], sigma = meta[2];
    const float fm = sigma * s.mse_unscale;
    const float fq = s.scale * gamma * kQjlUnscale;
    const float xns = xn * xn;
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = 32 * r + 8 * (v >> 2) + 4 * kg + (v & 3);
        const long q = q0 + row;
        if (q >= m) continue;
        const float ip = (QJL ? acc_m[r][c][v] * fm + acc_q[r][c][v] * fq : acc_m[r][c][v] * fm) *
                         xn * l_qn[row][1];
        const float d = l_qn[row][0] + xns - 2.0f * ip;
        out[q * n + j] = d > 0.0f ? d : 0.0f;
      }
  }
}

}  // namespace

hipError_t launch_split_queries(const SpaceDev &s, const float *prep, long m, _Float16 *planes,
                               hipStream_t st) {
  if (m <= 0) return hipSuccess;
  long blocks = (m * (long)s.D + 255) / 256;
  if (blocks > 65535) blocks = 65535;
  hipLaunchKernelGGL(k_split_queries, dim3((unsigned)blocks), dim3(256), 0, st, s, prep, m,
                     planes, s.nib);
  return hipGetLastError();
}

namespace {
template <int QT, bool QJL>
void launch_tile(const SpaceDev &s, dim3 grid, const _Float16 *planes, const float *prep, long m,
                 const uint8_t *codes, long n, float *out, hipStream_t st) {
  if (s.nib && s.mfma_lo)
    hipLaunchKernelGGL((k_mfma<true, true, QT, QJL>), grid, dim3(256), 0, st, s, planes, prep, m,
                       codes, n, out);
  else if (s.nib)
    hipLaunchKernelGGL((k_mfma<true, false, QT, QJL>), grid, dim3(256), 0, st, s, planes, prep, m,
                       codes, n, out);
  else
    hipLaunchKernelGGL((k_mfma<false, true, QT, QJL>), grid, dim3(256), 0, st, s, planes, prep, m,
                       codes, n, out);
}
}  // namespace

hipError_t launch_mfma_planes(const SpaceDev &s, const _Float16 *planes, const float *prep, long m,
                              const uint8_t *codes, long n, float *out, hipStream_t st, bool qjl) {
  if (s.D < kKC || m > 65535L * kTQ) return hipErrorNotSupported;
  if (m <= 0 || n <= 0) return hipSuccess;
  const unsigned gx = (unsigned)((n + kTN - 1) / kTN);
  if (m <= 32) {  // one 32-query tile: half the matrix work of a 64-query one
    const dim3 grid(gx, 1);
    if (qjl)
      launch_tile<32, true>(s, grid, planes, prep, m, codes, n, out, st);
    else
      launch_tile<32, false>(s, grid, planes, prep, m, codes, n, out, st);
    return hipGetLastError();
  }
  const dim3 grid(gx, (unsigned)((m + kTQ - 1) / kTQ));
  if (qjl)
    launch_tile<kTQ, true>(s, grid, planes, prep, m, codes, n, out, st);
  else
    launch_tile<kTQ, false>(s, grid, planes, prep, m, codes, n, out, st);
  return hipGetLastError();
}

hipError_t launch_scan_mfma(const SpaceDev &s, const float *prep, long m, const uint8_t *codes,
                            long n, float *out, hipStream_t st, bool qjl) {
  if (s.D < kKC) return hipErrorNotSupported;
  if (m <= 0 || n <= 0) return hipSuccess;
  // grid.y covers 64 queries per workgroup row; larger batches go in chunks.
  const long chunk_max = 65535L * 64;
  const long chunk = m < chunk_max ? m : chunk_max;
  _Float16 *planes = nullptr;
  const size_t bytes = kPlanes * (size_t)chunk * s.D * sizeof(_Float16);
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
                       planes, s.nib);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    e = launch_mfma_planes(s, planes, pc, mc, codes, n, oc, st, qjl);
  }
  hipError_t f = hipFreeAsync(planes, st);
  return e != hipSuccess ? e : f;
}

}  // namespace tq
