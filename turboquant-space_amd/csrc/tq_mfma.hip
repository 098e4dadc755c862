// tq_mfma.hip -- distance_m_to_n as a matrix-core contraction on gfx950.
//
// The asymmetric estimator (space_turbo_quant.h:659-686, :951-983) is, over a
// batch, two dense contractions between the queries and the decoded codes:
//   MSE[q, j] = sum_i q_rot[q, i] * c[idx_{j,i}]
//   QJL[q, j] = sum_i s_q[q, i]  * (qjl_{j,i} ? +1 : -1)
// followed by a per-pair fp32 epilogue with the code's norm/gamma/sigma and
// the query's norm (identical to the plain path's).
//
// Precision: fp16 holds neither the queries nor the centroids to fp32
// accuracy, so both are split into hi + lo fp16 parts (v = hi + lo with
// hi = fp16(v), lo = fp16(v - hi), ~22 significant bits) and
//   MSE = q_hi*c_hi + q_lo*c_hi + q_hi*c_lo      (q_lo*c_lo ~2^-22, dropped)
//   QJL = s_hi*(+-1) + s_lo*(+-1)                 (+-1 is exact in fp16)
// all accumulated in fp32 (v_mfma_f32_32x32x16_f16). Byte spaces (5-9 bits)
// drop the s_lo pass: their residual norm gamma, which scales the QJL term,
// is ~15x smaller than at 4 bits, so single-fp16 s_q stays ~1e-6 relative.
// Queries are scaled by 2^10 and centroids by 2^8 so the lo parts stay in
// fp16's normal range; the epilogue removes the scales. With every lo pass
// in, the error against the fp64 restatement of the estimator is a few 1e-6
// relative to the distance on each query's nearest codes.
// 4-bit and byte-code spaces at padded dim >= 1024 run a coarse pass set
// instead (mfma_coarse): q_hi*c_hi, q_lo*c_hi and s_hi*(+-1), three passes
// instead of five (4-bit) or four (bytes, which never take s_lo). The
// dropped terms move a distance by up to a few 1e-5 of (|q|^2 + |x|^2); so
// every pair whose coarse distance is below half of that scale (a query's
// near codes, the distances a top-k decides on) gets them added in the
// epilogue, summed in fixed point (refine_flagged, tq_mfma_tile.h), and the
// pairs above it stay within a few 1e-5 relative. Both parts are functions
// of the (query, code) pair alone.
//
// Tiling: a workgroup (4 wavefronts) owns 64 (or 32) queries x 256 codes;
// each wavefront owns the queries x 64 codes as 2 x 2 (or 1 x 2) blocks of
// 32 x 32 MFMA outputs. The padded dim is walked in chunks of 32 coordinates.
// A chunk's query-plane fragments (written by k_split_queries already in
// fragment order, 1 KB per 32 x 16 block, so every LDS fragment read is one
// contiguous conflict-free 1 KB read) and code bytes are loaded by all 256
// threads into a register ring two or three chunks ahead, stored to one of
// two LDS buffers one chunk ahead, one barrier per chunk. A lane decodes the
// code bytes of its fragment in registers: nibble codes through v_perm byte
// tables of the 8 centroids, byte codes through a (c_hi | c_lo << 16) table
// in LDS.
#include "tq_mfma.h"
#include "tq_mfma_tile.h"

namespace tq {
namespace {

using namespace mfma;

// One thread per (query block, k step, lane): the lane's 8 elements of every
// plane (layout in tq_mfma_dev.h), and with `corr` the coarse form's
// correction row group. Rows past m are zero.
__global__ __launch_bounds__(256) void k_split_queries(SpaceDev s, const float *__restrict__ prep,
                                                       long m, _Float16 *__restrict__ planes,
                                                       int npl, _Float16 *__restrict__ corr) {
  const long stride = prep_stride(s.D);
  const int nks = s.D / 16;
  const long total = (m + 31) / 32 * (long)nks * 64;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int lane = (int)(t & 63);
    const long blk = t >> 6;  // rb * nks + ks
    const int ks = (int)(blk % nks);
    const long rb = blk / nks;
    const long q = rb * 32 + (lane & 31);
    const int base = 32 * (ks >> 1) + 16 * (lane >> 5) + 8 * (ks & 1);
    float v[8], w[8];
    if (q < m) {
      const float4 *pq = reinterpret_cast<const float4 *>(prep + q * stride + base);
      const float4 *ps = reinterpret_cast<const float4 *>(prep + q * stride + s.D + base);
      const float4 a0 = pq[0], a1 = pq[1], b0 = ps[0], b1 = ps[1];
      v[0] = a0.x; v[1] = a0.y; v[2] = a0.z; v[3] = a0.w;
      v[4] = a1.x; v[5] = a1.y; v[6] = a1.z; v[7] = a1.w;
      w[0] = b0.x; w[1] = b0.y; w[2] = b0.z; w[3] = b0.w;
      w[4] = b1.x; w[5] = b1.y; w[6] = b1.z; w[7] = b1.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = w[e] = 0.0f;
    }
    h8 hq, lq, hs, ls;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int i = split_perm(s.nib != 0, e);
      const float qv = v[i] * kQScale, sv = w[i] * kQScale;
      hq[e] = (_Float16)qv;
      lq[e] = (_Float16)(qv - (float)hq[e]);
      hs[e] = (_Float16)sv;
      ls[e] = (_Float16)(sv - (float)hs[e]);
    }
    h8 *dst = reinterpret_cast<h8 *>(planes + blk * npl * kBlockHalves) + lane;
    dst[0] = hq;
    dst[64] = lq;
    if (npl > 2) dst[128] = hs;
    if (npl > 3) dst[192] = ls;
    if (corr) {
      h8 *cd = reinterpret_cast<h8 *>(corr + (q * (s.D / 8) + base / 8) * kCorrHalves);
      cd[0] = hq;
      cd[1] = ls;
    }
  }
}

// Two workgroups per CU (two wavefronts per SIMD) so one wavefront's MFMA
// chain overlaps the other's decode, LDS traffic and barrier waits. QT = 64
// queries per workgroup or 32 for batches of at most 32 queries: every output
// element depends only on its own query row and code column, so both give
// the same bits. QJL = false drops the s_q planes and their passes (queries
// with s_q = 0: the symmetric MSE estimator, whose "queries" are decoded
// codes); the QJL term is then exactly zero. thr (coarse form, blocked
// search): query q's running k-th best distance at thr[q * thr_stride], an
// upper bound on its final one; near pairs that cannot reach it keep their
// coarse value (refine_bound). thr = null: every near pair is corrected.
template <bool NIB, int QT, bool QJL, bool CO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mfma(
    SpaceDev s, const _Float16 *__restrict__ planes, const float *__restrict__ prep, long m,
    const uint8_t *__restrict__ codes, long n, float *__restrict__ out,
    const float *__restrict__ thr, long thr_stride) {
  // 64-coordinate chunks (one barrier per 64) for nibble codes where the
  // stage buffers still leave room for two workgroups per CU: the 64-query
  // coarse tiles (2 x 64 KB of the 160 KB LDS) and every 32-query tile.
  constexpr bool WIDE = NIB && (CO || QT == 32);
  using Sh = MfmaShape<NIB, QT, QJL, CO, WIDE>;
  constexpr int RB = Sh::RB;
  static_assert(QT == 32 || QT == 64, "query tile");
  __shared__ __attribute__((aligned(16))) uint8_t stage[2][Sh::STAGE];
  // Byte codes: 32 copies of each of the 128 centroid entries (copy c of
  // entry k at 32 k + c), lane l reading copy l % 32 (decode8<true>). Its
  // alignment to its own 16 KB size puts it at LDS offset 0, so an entry's
  // address is the or of the entry and copy offsets (one and-or per byte).
  __shared__ __attribute__((aligned(NIB ? 4 : 16384))) uint32_t l_tab[NIB ? 1 : 128 * 32];
  __shared__ float l_qn[QT][2];  // q_norm_sq, q_norm of the tile's queries
  // coarse form: the centroid remainders (8 or 128), for refine_flagged,
  // and with thr the tile's thresholds and refine_limits
  __shared__ float l_clo[CO ? (NIB ? 8 : 128) : 1];
  __shared__ float l_thr[CO ? QT : 1];
  __shared__ float2 l_lim[CO ? QT : 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 31, kg = lane >> 5;
  // Tile of this workgroup; tile t is (query tile t % nqt, code tile
  // t / nqt). With several query tiles, workgroup b of the one-dimensional
  // grid, which runs on XCD b % 8 (dispatch goes round the XCDs in order),
  // takes tile numbers consecutive with the other workgroups of its XCD: the
  // query tiles of a code tile run side by side on one XCD, so its code bytes
  // come from HBM into that XCD's L2 once for all of them. With one query
  // tile, workgroup b takes code tile b (consecutive code tiles spread over
  // the XCDs measured ~3% faster than contiguous ranges per XCD).
  const long nqt = (m + QT - 1) / QT;
  long tile = blockIdx.x;
  if (nqt > 1) {
    const long T = gridDim.x, b = blockIdx.x, x = b % 8;
    const long per = T / 8, rem = T % 8;
    tile = x * per + (x < rem ? x : rem) + b / 8;
  }
  const long q0 = (tile % nqt) * QT;
  const long j0 = (tile / nqt) * kTN;

  if (tid < QT) {
    const long q = q0 + tid;
    const float *pq = prep + (q < m ? q : 0L) * prep_stride(s.D);
    l_qn[tid][0] = pq[2 * s.D];
    l_qn[tid][1] = pq[2 * s.D + 1];
  }
  const _Float16 *const corr = planes + split_planes_halves(s.D, m, Sh::NPL);
  if constexpr (CO) {
    if (tid < (NIB ? 8 : 128)) l_clo[tid] = cent_lo(s, tid);
    if (thr != nullptr && tid < QT) {
      const long q = q0 + tid;
      const float *nq = reinterpret_cast<const float *>(corr + corr_halves(s.D, m)) + 2 * q;
      l_thr[tid] = q < m ? thr[q * thr_stride] : 0.0f;
      l_lim[tid] = refine_limits(s, QJL, nq[0], nq[1]);
    }
  }
  if constexpr (!NIB) {
    for (int i = tid; i < 128 * 32; i += blockDim.x) {
      _Float16 ch, cl;
      cent_h(s, i >> 5, ch, cl);
      l_tab[i] = h_bits(ch) | (h_bits(cl) << 16);
    }
  }
  PermTab ptab{};
  if constexpr (NIB) ptab = perm_tab(s);

  f16v acc_m[RB][2];
  f16v acc_q[QJL ? RB : 1][2];
  mfma_tile<NIB, QT, QJL, true, CO, WIDE, true>(s, planes, m, codes, n, q0, j0, stage, l_tab,
                                                ptab, acc_m, acc_q);

  // Epilogue: output element v of a 32x32 block sits at row
  // 8*(v/4) + 4*kg + (v%4), column lr. The per-code factors are formed once;
  // the expression is the reference's ip = (sigma*mse + scale*gamma*dot)*|x|*|q|,
  // d = max(0, |q|^2 + |x|^2 - 2*ip). The coarse form then corrects its
  // flagged pairs (bit 16 * (2 r + c) + v) and writes them again; every lane
  // takes part, so no lane leaves early.
  float fm[2], fq[2], xnv[2], xns[2];
  bool jin[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const long j = j0 + wave * 64 + 32 * c + lr;
    jin[c] = j < n;
    const float *meta =
        reinterpret_cast<const float *>(codes + (jin[c] ? j : n - 1) * (long)s.cs + s.packed);
    const float xn = meta[0], gamma = meta[1], sigma = meta[2];
    fm[c] = sigma * s.mse_unscale;
    fq[c] = s.scale * gamma * kQjlUnscale;
    xnv[c] = xn;
    xns[c] = xn * xn;
  }
  // Tiles inside both bounds (all but the last of each dimension) store
  // without per-element checks; each lane's element rows are 4 kg plus a
  // compile-time offset rr, so every store address is the lane's base plus a
  // wavefront-uniform rr * n.
  uint64_t flag = 0;
  float *const obase = out + (q0 + 4 * kg) * n + j0 + wave * 64 + lr;
  auto epilogue = [&](auto full) {
    constexpr bool FULL = decltype(full)::value;
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rr = 32 * r + 8 * (v >> 2) + (v & 3);
        const int row = rr + 4 * kg;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          float aq = 0.0f;
          if constexpr (QJL) aq = acc_q[r][c][v];
          const float d = mfma_distance<QJL>(acc_m[r][c][v], aq, fm[c], fq[c], xnv[c], xns[c],
                                             l_qn[row][0], l_qn[row][1]);
          if constexpr (FULL) {
            obase[(long)rr * n + 32 * c] = d;
            if constexpr (CO)
              flag |= (uint64_t)refine_flag(d, l_qn[row][0], xns[c]) << (16 * (2 * r + c) + v);
          } else if (jin[c] && q0 + row < m) {
            obase[(long)rr * n + 32 * c] = d;
            if constexpr (CO)
              if (refine_flag(d, l_qn[row][0], xns[c])) flag |= 1ull << (16 * (2 * r + c) + v);
          }
        }
      }
  };
  if (q0 + QT <= m && j0 + kTN <= n)
    epilogue(std::true_type{});
  else
    epilogue(std::false_type{});
  if constexpr (CO) {
    // this lane's sums of element b (a select chain: no indexing by b)
    auto sums = [&](int b, float &am, float &aq) {
      am = 0.0f;
      aq = 0.0f;
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int v = 0; v < 16; ++v)
            if (b == 16 * (2 * r + c) + v) {
              am = acc_m[r][c][v];
              if constexpr (QJL) aq = acc_q[r][c][v];
            }
    };
    if (thr != nullptr && __any(flag != 0u))
      flag = prune_flagged(flag, [&](int b) {
        float am, aq;
        sums(b, am, aq);
        const int v = b & 15, c = (b >> 4) & 1;
        const int row = 32 * (b >> 5) + 8 * (v >> 2) + 4 * kg + (v & 3);
        const float fmc = c ? fm[1] : fm[0], fqc = c ? fq[1] : fq[0];
        const float xnc = c ? xnv[1] : xnv[0], xsc = c ? xns[1] : xns[0];
        const float qns = l_qn[row][0], qn = l_qn[row][1];
        const float d = mfma_distance<QJL>(am, aq, fmc, fqc, xnc, xsc, qns, qn);
        const float e = refine_bound(am, aq, fmc, fqc, xnc, xsc, qns, qn, l_lim[row]);
        return !(d - e > l_thr[row]);
      });
    refine_flagged<NIB, QJL>(
        s, [&](uint32_t k) { return l_clo[k]; }, corr, codes, flag, lane,
        [&](int b, int src, long &q, long &j) {
          const int v = b & 15;
          q = q0 + 32 * (b >> 5) + 8 * (v >> 2) + 4 * (src >> 5) + (v & 3);
          j = j0 + wave * 64 + 32 * ((b >> 4) & 1) + (src & 31);
        },
        [&](int b, float dm, float dq) {
          float am, aq;
          sums(b, am, aq);
          const int v = b & 15, c = (b >> 4) & 1;
          const int row = 32 * (b >> 5) + 8 * (v >> 2) + 4 * kg + (v & 3);
          out[(q0 + row) * n + j0 + wave * 64 + 32 * c + lr] = mfma_distance<QJL>(
              am + dm, aq + dq, c ? fm[1] : fm[0], c ? fq[1] : fq[0], c ? xnv[1] : xnv[0],
              c ? xns[1] : xns[0], l_qn[row][0], l_qn[row][1]);
        });
  }
}

// One wavefront per query row (m rounded up to 32; rows past m give 0):
// sum |q_hi| and sum |s_lo| over the row's correction groups, each lane
// summing its groups in order and the lanes folding in a fixed butterfly,
// into norms[2 q], norms[2 q + 1].
__global__ __launch_bounds__(256) void k_corr_norms(int D, long m,
                                                    const _Float16 *__restrict__ corr,
                                                    float *__restrict__ norms) {
  const long q = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (q >= (m + 31) / 32 * 32) return;  // no barriers below
  const int ng = D >> 3;
  float a = 0.0f, b = 0.0f;
  if (q < m)
    for (int g = lane; g < ng; g += 64) {
      const h8 qh = *reinterpret_cast<const h8 *>(corr + (q * ng + g) * kCorrHalves);
      const h8 sl = *reinterpret_cast<const h8 *>(corr + (q * ng + g) * kCorrHalves + 8);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        a += __builtin_fabsf((float)qh[e]);
        b += __builtin_fabsf((float)sl[e]);
      }
    }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  if (lane == 0) {
    norms[2 * q] = a;
    norms[2 * q + 1] = b;
  }
}

}  // namespace

size_t split_planes_bytes(const SpaceDev &s, long m, bool qjl) {
  const bool co = mfma_coarse(s);
  long halves = split_planes_halves(s.D, m, split_planes(s.nib != 0, qjl, co));
  if (co) halves += corr_halves(s.D, m) + 2 * corr_norm_floats(m);
  return (size_t)halves * sizeof(_Float16);
}

hipError_t launch_split_queries(const SpaceDev &s, const float *prep, long m, _Float16 *planes,
                                hipStream_t st, bool qjl) {
  if (m <= 0) return hipSuccess;
  const long threads = (m + 31) / 32 * (long)(s.D / 16) * 64;
  long blocks = (threads + 255) / 256;
  if (blocks > 65535) blocks = 65535;
  const bool co = mfma_coarse(s);
  const int npl = split_planes(s.nib != 0, qjl, co);
  _Float16 *corr = co ? planes + split_planes_halves(s.D, m, npl) : nullptr;
  hipLaunchKernelGGL(k_split_queries, dim3((unsigned)blocks), dim3(256), 0, st, s, prep, m,
                     planes, npl, corr);
  if (co) {
    const long rows = (m + 31) / 32 * 32;
    hipLaunchKernelGGL(k_corr_norms, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, st, s.D, m,
                       corr, reinterpret_cast<float *>(corr + corr_halves(s.D, m)));
  }
  return hipGetLastError();
}

namespace {
template <int QT, bool QJL>
void launch_tile(const SpaceDev &s, dim3 grid, const _Float16 *planes, const float *prep, long m,
                 const uint8_t *codes, long n, float *out, const float *thr, long thr_stride,
                 hipStream_t st) {
  if (mfma_coarse(s) && s.nib)
    hipLaunchKernelGGL((k_mfma<true, QT, QJL, true>), grid, dim3(256), 0, st, s, planes, prep, m,
                       codes, n, out, thr, thr_stride);
  else if (s.nib)
    hipLaunchKernelGGL((k_mfma<true, QT, QJL, false>), grid, dim3(256), 0, st, s, planes, prep, m,
                       codes, n, out, thr, thr_stride);
  else if (mfma_coarse(s))
    hipLaunchKernelGGL((k_mfma<false, QT, QJL, true>), grid, dim3(256), 0, st, s, planes, prep,
                       m, codes, n, out, thr, thr_stride);
  else
    hipLaunchKernelGGL((k_mfma<false, QT, QJL, false>), grid, dim3(256), 0, st, s, planes, prep,
                       m, codes, n, out, thr, thr_stride);
}
}  // namespace

hipError_t launch_mfma_planes(const SpaceDev &s, const _Float16 *planes, const float *prep, long m,
                              const uint8_t *codes, long n, float *out, hipStream_t st, bool qjl,
                              int qt, const float *thr, long thr_stride) {
  if (s.D < kKC || m > 65535L * 64) return hipErrorNotSupported;
  if (m <= 0 || n <= 0) return hipSuccess;
  const long gx = (n + kTN - 1) / kTN;
  if (qt == 0) qt = m <= 32 ? 32 : 64;  // one 32-query tile: half the matrix work of a 64-query one
  if (qt == 32) {
    const long tiles = gx * ((m + 31) / 32);
    if (tiles > 0xFFFFFFFFL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)tiles);
    if (qjl)
      launch_tile<32, true>(s, grid, planes, prep, m, codes, n, out, thr, thr_stride, st);
    else
      launch_tile<32, false>(s, grid, planes, prep, m, codes, n, out, thr, thr_stride, st);
    return hipGetLastError();
  }
  const long tiles = gx * ((m + 63) / 64);
  if (tiles > 0xFFFFFFFFL) return hipErrorInvalidValue;
  const dim3 grid((unsigned)tiles);
  if (qjl)
    launch_tile<64, true>(s, grid, planes, prep, m, codes, n, out, thr, thr_stride, st);
  else
    launch_tile<64, false>(s, grid, planes, prep, m, codes, n, out, thr, thr_stride, st);
  return hipGetLastError();
}

hipError_t launch_scan_mfma(const SpaceDev &s, const float *prep, long m, const uint8_t *codes,
                            long n, float *out, hipStream_t st, bool qjl) {
  if (m <= 0 || n <= 0) return hipSuccess;
  // Below one 32-coordinate chunk there is no MFMA tile: the plain fp32
  // kernels serve the call (same estimator, fp32 accuracy).
  if (s.D < kKC) return launch_scan_fp32(s, prep, m, codes, n, out, !qjl, st);
  // grid.y covers 64 queries per workgroup row; larger batches go in chunks.
  const long chunk_max = 65535L * 64;
  const long chunk = m < chunk_max ? m : chunk_max;
  _Float16 *planes = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void **>(&planes),
                                split_planes_bytes(s, chunk, qjl), st);
  if (e != hipSuccess) return e;
  const long pstride = prep_stride(s.D);
  for (long q0 = 0; q0 < m && e == hipSuccess; q0 += chunk) {
    const long mc = (m - q0) < chunk ? (m - q0) : chunk;
    const float *pc = prep + q0 * pstride;
    float *oc = out + q0 * n;
    e = launch_split_queries(s, pc, mc, planes, st, qjl);
    if (e != hipSuccess) break;
    e = launch_mfma_planes(s, planes, pc, mc, codes, n, oc, st, qjl);
  }
  hipError_t f = hipFreeAsync(planes, st);
  return e != hipSuccess ? e : f;
}

}  // namespace tq
