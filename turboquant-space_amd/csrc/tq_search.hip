// tq_search.hip -- top-k search: the MFMA m x n scan with a fused top-k, run
// as one persistent launch over a queue of (job, query tile, code range)
// items. Jobs may come from different collections, dims and bit widths.
//
// Per item a workgroup walks its code tiles with k_mfma's tile contraction
// (mfma_tile, tq_mfma_tile.h: 256 codes per tile, 32-coordinate chunks
// through a register ring and LDS, the fp16 hi+lo MFMA passes) and the
// reference's fp32 epilogue (space_turbo_quant.h:682-685), over a query tile
// of 32 (QR = 1) or 64 queries (QR = 2), two workgroups per CU. Instead of
// writing the distances, each wavefront keeps the running best K (distance,
// code id) entries of the codes it has seen in 64 per-lane lists in LDS:
// with QR = 2 lane q holds query q's list; with QR = 1 lanes q and q + 32
// hold query q's lists for the two halves of the wavefront's 64 codes. A
// tile's distances reach the lists through an LDS transpose; a lane reads
// its list into registers only when one of the wavefront's distances beats
// a list's K-th entry. Every (query, code) distance is the same arithmetic
// whatever else is in the launch (and the same as k_mfma's), and ties break
// by id, so a query's result does not depend on the batch it was scored in.
//
// k_topk_merge then takes, per query, the best k of the lists of the items
// that covered it. This is the fused form of a search; k_topk_rows and
// k_topk_rows_fold (further down) give the row-wise top-k of dense distance
// blocks, which the blocked form (k_mfma blocks over chunks of codes, the
// plans' default) and the plain fp32 plans use. The reference has no top-k (its users run numpy
// argpartition on the full distance matrix, python/benchmarks/
// run_benchmark.py:251-255); ties there are unordered, here they are ordered
// by id.
#include "tq_mfma.h"
#include "tq_mfma_tile.h"
#include "tq_search.h"

namespace tq {
namespace {

using namespace mfma;

constexpr int kSN = kSearchNTile;  // codes per tile (64 per wavefront)
constexpr int kTRow = 65;          // padded transpose row, in floats

__device__ __forceinline__ bool lex_less(float a, uint32_t ia, float b, uint32_t ib) {
  return a < b || (a == b && ia < ib);
}

// Sorted insertion into a lane's ascending list (no-op when the entry is not
// below the last one). Branch-free, so the wavefront runs it in lockstep.
template <int K>
__device__ __forceinline__ void topk_insert(float (&td)[K], uint32_t (&ti)[K], float d, uint32_t id) {
  float cd = d;
  uint32_t ci = id;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool less = lex_less(cd, ci, td[k], ti[k]);
    const float nd = less ? td[k] : cd;
    const uint32_t ni = less ? ti[k] : ci;
    td[k] = less ? cd : td[k];
    ti[k] = less ? ci : ti[k];
    cd = nd;
    ci = ni;
  }
}

// Lexicographic minimum of (d, i) with the lane DPP control CTRL selects.
template <int CTRL>
__device__ __forceinline__ void lex_min_dpp(float &d, uint32_t &i) {
  const float od = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(d), CTRL, 0xF, 0xF, true));
  const uint32_t oi = (uint32_t)__builtin_amdgcn_mov_dpp((int)i, CTRL, 0xF, 0xF, true);
  if (lex_less(od, oi, d, i)) {
    d = od;
    i = oi;
  }
}

// Wavefront-wide lexicographic minimum of (d, i), returned to every lane:
// DPP within each row of 16 lanes (quad_perm lane^1, lane^2, row_half_mirror
// 7-i, row_mirror 15-i), then the four rows' results through readlane. All
// lanes must be active. No LDS round trips (a shuffle tree costs six).
__device__ __forceinline__ void wave_lex_min(float &d, uint32_t &i) {
  lex_min_dpp<0xB1>(d, i);
  lex_min_dpp<0x4E>(d, i);
  lex_min_dpp<0x141>(d, i);
  lex_min_dpp<0x140>(d, i);
  float bd = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d), 0));
  uint32_t bi = (uint32_t)__builtin_amdgcn_readlane((int)i, 0);
#pragma unroll
  for (int r = 1; r < 4; ++r) {
    const float rd = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d), 16 * r));
    const uint32_t ri = (uint32_t)__builtin_amdgcn_readlane((int)i, 16 * r);
    if (lex_less(rd, ri, bd, bi)) {
      bd = rd;
      bi = ri;
    }
  }
  d = bd;
  i = bi;
}

template <int K>
__device__ __forceinline__ void topk_clear(float (&td)[K], uint32_t (&ti)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    td[k] = __builtin_inff();
    ti[k] = 0xFFFFFFFFu;
  }
}

__global__ void k_reset_counter(int *counter) { *counter = 0; }

__global__ __launch_bounds__(256) void k_fill_f32(float *__restrict__ p, long n, float v) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = v;
}

template <int QR>
struct SearchShape {
  static constexpr int QT = 32 * QR;  // queries per item
  using ShN = MfmaShape<true, QT, true>;
  using ShB = MfmaShape<false, QT, true>;
  static constexpr int STAGE = ShN::STAGE > ShB::STAGE ? ShN::STAGE : ShB::STAGE;
  static constexpr int TRANS = 4 * 32 * kTRow * 4;  // one 32-row distance block per wavefront
  static constexpr int LDS = 2 * STAGE > TRANS ? 2 * STAGE : TRANS;
};

// One code tile of an item: the tile's contraction (mfma_tile, the same as
// k_mfma's), the reference epilogue, and the distances into the
// wavefront's lists, one 32-query block at a time through a transpose.
// QR = 2: lanes 32 r .. 32 r + 31 own the queries of block r and read all 64
// codes of the wavefront; QR = 1: lanes q and q + 32 take codes 0-31 and
// 32-63 of query q. A lane's list (lane l of wavefront w: l_ld / l_li
// [w][0..K-1][l], ascending, empty entries (inf, ~0) last) is read into
// registers only when one of the wavefront's distances beats a list's kk-th
// entry (kk = the k asked for), and written back after.
template <bool NIB, bool CO, int QR, int K>
__device__ __forceinline__ void search_tile(const SearchJobDev &job, long q0, long j0, int kk,
                                            uint8_t *lds, const uint32_t *l_tab,
                                            const PermTab &ptab, const float (*l_qn)[2],
                                            const float *l_clo, const float2 *l_lim,
                                            float (*l_gthr)[64], float (*l_ld)[K][64],
                                            uint32_t (*l_li)[K][64]) {
  constexpr int QT = 32 * QR;
  using Sh = MfmaShape<NIB, QT, true, CO>;
  const SpaceDev &s = job.s;
  const long m = job.m, n = job.n;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 31, kg = lane >> 5;
  auto stage = reinterpret_cast<uint8_t(*)[Sh::STAGE]>(lds);
  auto t_lds = reinterpret_cast<float(*)[32][kTRow]>(lds);

  f16v acc_m[QR][2], acc_q[QR][2];
  // Coarse form: the job's running bound on the k-th best of query row `grow`
  // (lane l holds row l, or l % 32 for 32-query tiles), loaded now and
  // consumed after the contraction.
  const int grow = QR == 2 ? lane : (lane & 31);
  float gth = __builtin_inff();
  if constexpr (CO)
    if (q0 + grow < m) gth = job.thr[q0 + grow];
  __syncthreads();  // the previous tile's transpose reads are done with the stage area
  mfma_tile<NIB, QT, true, false, CO>(s, job.planes, m, job.codes, n, q0, j0, stage, l_tab, ptab,
                                      acc_m, acc_q);

  const long jbase = j0 + wave * 64;
  float fm[2], fq[2], xnv[2], xns[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const long j = jbase + 32 * c + lr;
    const float *meta =
        reinterpret_cast<const float *>(job.codes + (j < n ? j : n - 1) * (long)s.cs + s.packed);
    const float xn = meta[0], gamma = meta[1], sigma = meta[2];
    fm[c] = sigma * s.mse_unscale;
    fq[c] = s.scale * gamma * kQjlUnscale;
    xnv[c] = xn;
    xns[c] = xn * xn;
  }
  // k_mfma's epilogue: the same distances, and the coarse form's flagged pairs
  // (bit 32 r + 16 c + v) corrected the same way, here by adding the dropped
  // terms to the sums before the distances are formed (the same fp32 sum
  // k_mfma forms), all of the wavefront's blocks in one pass. Near pairs
  // that cannot beat their list's kk-th entry (a bound on the query's final
  // kk-th best) stay uncorrected, as in the blocked form (refine_bound).
  if constexpr (CO) {
    l_gthr[wave][lane] = gth;  // the wavefront's own copy: no barrier needed
    uint64_t flag = 0;
#pragma unroll
    for (int r = 0; r < QR; ++r)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const bool jin = jbase + 32 * c + lr < n;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int qrow = 32 * r + 8 * (v >> 2) + 4 * kg + (v & 3);
          const float d = mfma_distance<true>(acc_m[r][c][v], acc_q[r][c][v], fm[c], fq[c],
                                              xnv[c], xns[c], l_qn[qrow][0], l_qn[qrow][1]);
          if (jin && q0 + qrow < m && refine_flag(d, l_qn[qrow][0], xns[c]))
            flag |= 1ull << (32 * r + 16 * c + v);
        }
      }
    if (__any(flag != 0u))
      flag = prune_flagged(flag, [&](int b) {
        const int r = b >> 5, c = (b >> 4) & 1, v = b & 15;
        float am = 0.0f, aq = 0.0f;  // a select chain: no indexing by b
#pragma unroll
        for (int rr = 0; rr < QR; ++rr)
#pragma unroll
          for (int cc = 0; cc < 2; ++cc)
#pragma unroll
            for (int vv = 0; vv < 16; ++vv)
              if (b == 32 * rr + 16 * cc + vv) {
                am = acc_m[rr][cc][vv];
                aq = acc_q[rr][cc][vv];
              }
        const int qrow = 32 * r + 8 * (v >> 2) + 4 * kg + (v & 3);
        // the list this pair would enter: lane qrow (QR = 2) or lane
        // qrow + 32 c (QR = 1, lanes q and q + 32 split the codes)
        const int owner = QR == 2 ? qrow : qrow + 32 * c;
        const float fmc = c ? fm[1] : fm[0], fqc = c ? fq[1] : fq[0];
        const float xnc = c ? xnv[1] : xnv[0], xsc = c ? xns[1] : xns[0];
        const float qns = l_qn[qrow][0], qn = l_qn[qrow][1];
        const float d = mfma_distance<true>(am, aq, fmc, fqc, xnc, xsc, qns, qn);
        const float e = refine_bound(am, aq, fmc, fqc, xnc, xsc, qns, qn, l_lim[qrow]);
        return !(d - e > __builtin_fminf(l_ld[wave][kk - 1][owner], l_gthr[wave][qrow]));
      });
    refine_flagged<NIB, true>(
        s, [&](uint32_t k) { return l_clo[k]; }, job.planes + split_planes_halves(s.D, m, Sh::NPL),
        job.codes, flag, lane,
        [&](int b, int src, long &q, long &j) {
          const int v = b & 15;
          q = q0 + 32 * (b >> 5) + 8 * (v >> 2) + 4 * (src >> 5) + (v & 3);
          j = jbase + 32 * ((b >> 4) & 1) + (src & 31);
        },
        [&](int b, float dm, float dq) {
#pragma unroll
          for (int r = 0; r < QR; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
              for (int v = 0; v < 16; ++v)
                if (b == 32 * r + 16 * c + v) {
                  acc_m[r][c][v] = acc_m[r][c][v] + dm;
                  acc_q[r][c][v] = acc_q[r][c][v] + dq;
                }
        });
  }
  const int h = QR == 1 ? (lane >> 5) : 0;
  const long cbase = jbase + 32 * h;
#pragma unroll
  for (int r = 0; r < QR; ++r) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = 8 * (v >> 2) + 4 * kg + (v & 3);
        const int qrow = 32 * r + row;
        t_lds[wave][row][32 * c + lr] = mfma_distance<true>(acc_m[r][c][v], acc_q[r][c][v], fm[c],
                                                            fq[c], xnv[c], xns[c], l_qn[qrow][0],
                                                            l_qn[qrow][1]);
      }
    __syncthreads();
    const bool mine = QR == 1 || (lane >> 5) == r;
    const float *rowp = &t_lds[wave][lane & 31][32 * h];
    // Candidates: the lane's codes below its list's kk-th entry (only a
    // query's best kk can reach its result), as a bit mask over the block.
    const float thd = l_ld[wave][kk - 1][lane];
    const uint32_t thi = l_li[wave][kk - 1][lane];
    constexpr int NK = 64 / (3 - QR);  // codes a lane scores in this block
    uint64_t cm = 0;
    if (mine && jbase < n) {
      for (int k = 0; k < NK; ++k)
        if (cbase + k < n && lex_less(rowp[k], (uint32_t)(cbase + k), thd, thi)) cm |= 1ull << k;
    }
    if (__any(cm != 0)) {
      float td[K];
      uint32_t ti[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        td[k] = l_ld[wave][k][lane];
        ti[k] = l_li[wave][k][lane];
      }
      // One candidate per lane per round (lowest code first, so ids enter in
      // ascending order as before); lanes without one insert the empty entry
      // (a no-op). Rounds = the most candidates any lane has.
      while (__any(cm != 0)) {
        float d = __builtin_inff();
        uint32_t id = 0xFFFFFFFFu;
        if (cm != 0) {
          const int k = __builtin_ctzll(cm);
          cm &= cm - 1;
          d = rowp[k];
          id = (uint32_t)(cbase + k);
        }
        topk_insert<K>(td, ti, d, id);
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        l_ld[wave][k][lane] = td[k];
        l_li[wave][k][lane] = ti[k];
      }
    }
    __syncthreads();  // block r's transpose reads are done before block r + 1 is written
  }
  // Publish a bound on each query's k-th best from the workgroup's lists of
  // it (4 lists, one per wavefront, or 8 with 32-query tiles, over disjoint
  // codes): with r = ceil(kk / lists), the largest of the lists' r-th
  // entries has at least kk distinct codes at or below it. One thread per
  // query row; distances are never negative or NaN, so their bits order as
  // unsigned integers.
  if constexpr (CO) {
    if (tid < QT) {
      constexpr int NL = QR == 2 ? 4 : 8;
      const int r = (kk + NL - 1) / NL;
      float bnd = 0.0f;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        bnd = __builtin_fmaxf(bnd, l_ld[w][r - 1][tid]);
        if constexpr (QR == 1) bnd = __builtin_fmaxf(bnd, l_ld[w][r - 1][tid + 32]);
      }
      if (q0 + tid < m && bnd < gth)
        atomicMin(reinterpret_cast<unsigned int *>(job.thr) + q0 + tid, __float_as_uint(bnd));
    }
  }
}

// Persistent fused search: workgroups pull (job, query tile, code-tile
// range) items from a queue; per item each wavefront keeps 64 lists of K
// (distance, id) entries in LDS (lane l's list: its query's best entries
// among the codes it scored) and writes them to the candidate buffer at the
// end: entries ((item * 4 + wavefront) * 64 + lane) * K. Two workgroups per
// CU for both query tiles (the lists live in LDS, the tile's registers are
// k_mfma's).
template <int QR, int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
k_search(const SearchJobDev *__restrict__ jobs, const SearchItem *__restrict__ items,
         const int *__restrict__ order, int nitems, int *__restrict__ counter, int kk,
         float *__restrict__ cand_d, uint32_t *__restrict__ cand_i) {
  using Sh = SearchShape<QR>;
  constexpr int QT = Sh::QT;
  // The stage buffers of the tile's contraction and the distance transpose
  // share one LDS area.
  __shared__ __attribute__((aligned(16))) uint8_t lds[Sh::LDS];
  __shared__ uint32_t l_tab[128];
  __shared__ float l_qn[QT][2];
  __shared__ float l_clo[128];  // coarse jobs: the centroid remainders, for refine_flagged
  __shared__ float2 l_lim[QT];  // coarse jobs: the tile's refine_limits
  __shared__ float l_gthr[4][64];  // coarse jobs: each wavefront's copy of the job-wide bounds
  __shared__ float l_ld[4][K][64];
  __shared__ uint32_t l_li[4][K][64];
  __shared__ int s_item;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;

  for (;;) {
    __syncthreads();  // the previous item is done with every shared buffer
    if (tid == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    // The same value in every lane: read into scalar registers, so the
    // item and job loads below are scalar and stay in SGPRs.
    const int pos = __builtin_amdgcn_readfirstlane(s_item);
    if (pos >= nitems) break;  // every wavefront reads the same value and leaves together
    const int it = __builtin_amdgcn_readfirstlane(order[pos]);  // queue position -> item
    const SearchItem item = items[it];
    const SearchJobDev &job = jobs[item.job];
    const SpaceDev &s = job.s;
    const bool nib = s.nib != 0;
    const bool coarse = mfma_coarse(s);
    const long m = job.m;
    const long q0 = (long)item.qtile * QT;

    if (!nib)
      for (int i = tid; i < 128; i += blockDim.x) {
        _Float16 ch, cl;
        cent_h(s, i, ch, cl);
        l_tab[i] = h_bits(ch) | (h_bits(cl) << 16);
      }
    if (coarse && tid < (nib ? 8 : 128)) l_clo[tid] = cent_lo(s, tid);
    if (coarse && tid < QT) {
      const _Float16 *corr =
          job.planes + split_planes_halves(s.D, m, split_planes(nib, true, true));
      const float *nq = reinterpret_cast<const float *>(corr + corr_halves(s.D, m));
      const long q = q0 + tid < m ? q0 + tid : 0L;
      l_lim[tid] = refine_limits(s, nib, nq[2 * q], nq[2 * q + 1]);
    }
    if (tid < QT) {
      const long q = q0 + tid;
      const float *pq = job.prep + (q < m ? q : 0L) * prep_stride(s.D);
      l_qn[tid][0] = pq[2 * s.D];
      l_qn[tid][1] = pq[2 * s.D + 1];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      l_ld[wave][k][lane] = __builtin_inff();
      l_li[wave][k][lane] = 0xFFFFFFFFu;
    }
    PermTab ptab{};
    if (nib) ptab = perm_tab(s);
    // (mfma_tile starts with the first chunk's loads; search_tile's barrier
    // orders the tables above before its first use of them.)
    for (int tile = item.t0; tile < item.t1; ++tile) {
      const long j0 = (long)tile * kSN;
      if (coarse && nib)
        search_tile<true, true, QR, K>(job, q0, j0, kk, lds, l_tab, ptab, l_qn, l_clo, l_lim,
                                       l_gthr, l_ld, l_li);
      else if (nib)
        search_tile<true, false, QR, K>(job, q0, j0, kk, lds, l_tab, ptab, l_qn, l_clo, l_lim,
                                        l_gthr, l_ld, l_li);
      else if (coarse)
        search_tile<false, true, QR, K>(job, q0, j0, kk, lds, l_tab, ptab, l_qn, l_clo, l_lim,
                                        l_gthr, l_ld, l_li);
      else
        search_tile<false, false, QR, K>(job, q0, j0, kk, lds, l_tab, ptab, l_qn, l_clo, l_lim,
                                         l_gthr, l_ld, l_li);
    }

    // This wavefront's lists for the item.
    const long base = (((long)it * 4 + wave) * 64 + lane) * K;
    for (int k = 0; k < K; ++k) {
      cand_d[base + k] = l_ld[wave][k][lane];
      cand_i[base + k] = l_li[wave][k][lane];
    }
  }
}

// Lane-held sorted lists -> the wavefront's overall best k_out, ascending:
// k_out rounds of a wavefront-wide (distance, id) minimum; the lane holding
// the minimum pops its head. Lane 0 writes the entries (id ~0 -> -1).
template <int K>
__device__ void wave_pop_topk(float (&td)[K], uint32_t (&ti)[K], int k_out, int lane,
                              float *out_d, int64_t *out_i) {
  for (int k = 0; k < k_out; ++k) {
    float bd = td[0];
    uint32_t bi = ti[0];
    wave_lex_min(bd, bi);
    if (lane == 0) {
      out_d[k] = bd;
      out_i[k] = bi == 0xFFFFFFFFu ? -1 : (int64_t)bi;
    }
    // Ids are unique within a query's lists, so at most one lane holds a
    // real minimum; empty entries popping together is harmless.
    if (td[0] == bd && ti[0] == bi) {
#pragma unroll
      for (int t = 0; t + 1 < K; ++t) {
        td[t] = td[t + 1];
        ti[t] = ti[t + 1];
      }
      td[K - 1] = __builtin_inff();
      ti[K - 1] = 0xFFFFFFFFu;
    }
  }
}

// Sorted list of K entries at (d, i) into the lane's list; stops at the first
// entry the list does not take.
template <int K>
__device__ __forceinline__ void take_sorted(float (&td)[K], uint32_t (&ti)[K], const float *d,
                                            const uint32_t *i) {
  for (int k = 0; k < K; ++k) {
    const float dk = d[k];
    const uint32_t ik = i[k];
    if (!lex_less(dk, ik, td[K - 1], ti[K - 1])) break;
    topk_insert<K>(td, ti, dk, ik);
  }
}

// One wavefront per query: lanes take the query's candidate lists (4 or 8
// per item that covered it), then pop the best k_out. Waves are laid out as
// 64 per group; a group with 32-query tiles leaves half of them idle.
template <int K>
__global__ __launch_bounds__(256) void k_topk_merge(const SearchJobDev *__restrict__ jobs,
                                                    const SearchGroup *__restrict__ groups,
                                                    int ngroups, int k_out,
                                                    const float *__restrict__ cand_d,
                                                    const uint32_t *__restrict__ cand_i) {
  const int lane = threadIdx.x & 63;
  const long wq = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long g = wq >> 6;
  if (g >= ngroups) return;
  const int qi = (int)(wq & 63);
  const SearchGroup grp = groups[g];
  const int qt = 32 * grp.qr;
  if (qi >= qt) return;
  const SearchJobDev &job = jobs[grp.job];
  const long q = (long)grp.qtile * qt + qi;
  if (q >= job.m) return;
  float td[K];
  uint32_t ti[K];
  topk_clear<K>(td, ti);
  const int per_item = grp.qr == 1 ? 8 : 4;  // lists per item for this query
  const int nlists = (grp.iend - grp.ibeg) * per_item;
  for (int l = lane; l < nlists; l += 64) {
    const long it = grp.ibeg + l / per_item;
    const int r = l % per_item;
    const int w = grp.qr == 1 ? r >> 1 : r;
    const int src_lane = grp.qr == 1 ? qi + 32 * (r & 1) : qi;
    const long base = ((it * 4 + w) * 64 + src_lane) * K;
    take_sorted<K>(td, ti, cand_d + base, cand_i + base);
  }
  wave_pop_topk<K>(td, ti, k_out, lane, job.out_d + q * k_out, job.out_i + q * k_out);
}

// Row-wise top-k of dense distance blocks (the blocked search form and the
// plain fp32 plans), for k up to kTopkMaxK. Lists are sorted (distance, id)
// entries, `lcap` per list (lcap = the k asked for), empty entries (+inf, ~0)
// at the tail.
//
// Both stages are one wavefront selecting the kk smallest (distance, id)
// entries of a set (a segment of a row's columns, or a row's lists):
//   1. A threshold T from a sample (the set's first 64 x kTopkPL entries,
//      kTopkPL per lane in registers): with r = 1 (kk <= 64) or 2, each lane's
//      r-th smallest sample entry, and T = the ceil(kk / r)-th smallest of
//      those 64 values. ceil(kk / r) lanes then hold r entries each at or below
//      T, so at least kk entries of the set lie at or below T and the set's kk
//      best all do. A caller's bound (the row's running kk-th entry) can lower
//      T further: entries above it cannot be among the row's kk best.
//   2. The real entries at or below T are packed into a per-wavefront LDS list
//      by ballot and lane prefix counts, in set order.
//   3. If more than kTopkCap were packed, T becomes the kk-th smallest of the
//      first kTopkCap packed entries (all at or below T, so T only drops and
//      at least kk entries stay at or below it) and the set is packed again,
//      until the list holds them all (exact either way; typical rows never
//      come here, adversarial orders a few times).
//   4. Lanes take the packed entries (up to 4 or 16 each) into sorted
//      lists, and min(count, kk) wavefront-minimum pops give the result in
//      order; result r stays with lane r % 64 (slot r / 64) for one
//      coalesced store.
// Only the wavefront's own lanes touch its LDS list: no barrier.
constexpr int kTopkSub = 2048;             // sample / pass: 64 lanes x kTopkPL
constexpr int kTopkPL = kTopkSub / 64;
constexpr int kBufferWord3 = 0x00020000;  // raw 32-bit buffer resource (gfx9 DATA_FORMAT)
constexpr int kTopkSeg = 4 * kTopkSub;     // columns per (row, segment) of stage 1
constexpr int kTopkCap = 1024;             // packed entries per wavefront

// kq (1..64, wavefront-uniform) pops of the 64 lanes' values: the kq-th
// smallest, returned to every lane.
__device__ __forceinline__ void wave_kth(float d, uint32_t i, int kq, float &td, uint32_t &ti) {
  td = __builtin_inff();
  ti = 0xFFFFFFFFu;
  for (int r = 0; r < kq; ++r) {
    float bd = d;
    uint32_t bi = i;
    wave_lex_min(bd, bi);
    td = bd;
    ti = bi;
    if (d == bd && i == bi) {  // ids are unique; empty entries popping together is harmless
      d = __builtin_inff();
      i = 0xFFFFFFFFu;
    }
  }
}

// A lane's r-th smallest (r = 1 or 2) of its entries (d[u], id[u]).
struct LaneRth {
  float d0 = __builtin_inff(), d1 = __builtin_inff();
  uint32_t i0 = 0xFFFFFFFFu, i1 = 0xFFFFFFFFu;
  __device__ __forceinline__ void add(float d, uint32_t i) {
    if (lex_less(d, i, d1, i1)) {
      if (lex_less(d, i, d0, i0)) {
        d1 = d0;
        i1 = i0;
        d0 = d;
        i0 = i;
      } else {
        d1 = d;
        i1 = i;
      }
    }
  }
};

// Step 4: the min(n, kk) smallest of the n <= kTopkCap packed entries, in
// order; result r into lane r % 64, slot r / 64 (empty past the last).
// Returns the last result (wavefront-uniform).
template <int L>
__device__ __forceinline__ void topk_pop_packed(const float *cd, const uint32_t *ci, int n, int kk,
                                                int lane, float (&rd)[2], uint32_t (&ri)[2],
                                                float &last_d, uint32_t &last_i) {
  float td[L];
  uint32_t ti[L];
  topk_clear<L>(td, ti);
  const int per = (n + 63) / 64;
#pragma unroll
  for (int t = 0; t < L; ++t) {
    if (t < per) {  // wavefront-uniform
      const int q = lane + 64 * t;
      if (q < n) topk_insert<L>(td, ti, cd[q], ci[q]);
    }
  }
  rd[0] = rd[1] = __builtin_inff();
  ri[0] = ri[1] = 0xFFFFFFFFu;
  last_d = __builtin_inff();
  last_i = 0xFFFFFFFFu;
  const int np = n < kk ? n : kk;
  for (int r = 0; r < np; ++r) {
    float bd = td[0];
    uint32_t bi = ti[0];
    wave_lex_min(bd, bi);
    if ((r & 63) == lane) {
      if (r < 64) {
        rd[0] = bd;
        ri[0] = bi;
      } else {
        rd[1] = bd;
        ri[1] = bi;
      }
    }
    last_d = bd;
    last_i = bi;
    if (td[0] == bd && ti[0] == bi) {
#pragma unroll
      for (int t = 0; t + 1 < L; ++t) {
        td[t] = td[t + 1];
        ti[t] = ti[t + 1];
      }
      td[L - 1] = __builtin_inff();
      ti[L - 1] = 0xFFFFFFFFu;
    }
  }
}

__device__ __forceinline__ void topk_pop_any(const float *cd, const uint32_t *ci, int n, int kk,
                                             int lane, float (&rd)[2], uint32_t (&ri)[2],
                                             float &last_d, uint32_t &last_i) {
  if (n <= 4 * 64)
    topk_pop_packed<4>(cd, ci, n, kk, lane, rd, ri, last_d, last_i);
  else
    topk_pop_packed<16>(cd, ci, n, kk, lane, rd, ri, last_d, last_i);
}

// Steps 1-4 over a set of E entries that SRC loads kTopkPL per lane per
// pass: src.pass(p, d, id) fills entries e = p * kTopkSub + u * 64 + lane
// (empty past E). (hint_d, hint_i) bounds T from above when real.
template <class SRC>
__device__ __forceinline__ void wave_select(const SRC &src, long E, int kk, float hint_d,
                                            uint32_t hint_i, float *cd, uint32_t *ci, int lane,
                                            float (&rd)[2], uint32_t (&ri)[2], int &np) {
  float v[kTopkPL];
  uint32_t vi[kTopkPL];
  src.pass(0, v, vi);
  // 1. T from the sample.
  const int r = kk > 64 ? 2 : 1;
  const int kq = (kk + r - 1) / r;
  LaneRth lr;
#pragma unroll
  for (int u = 0; u < kTopkPL; ++u) lr.add(v[u], vi[u]);
  float T_d;
  uint32_t T_i;
  wave_kth(r == 1 ? lr.d0 : lr.d1, r == 1 ? lr.i0 : lr.i1, kq, T_d, T_i);
  if (hint_i != 0xFFFFFFFFu && lex_less(hint_d, hint_i, T_d, T_i)) {
    T_d = hint_d;
    T_i = hint_i;
  }
  const long npass = (E + kTopkSub - 1) / kTopkSub;
  // 2. pack the real entries at or below T.
  auto pack = [&](bool reload0) -> long {
    long cnt = 0;
    for (long p = 0; p < npass; ++p) {
      if (p > 0 || reload0) src.pass(p, v, vi);
#pragma unroll
      for (int u = 0; u < kTopkPL; ++u) {
        const bool cand = vi[u] != 0xFFFFFFFFu && !lex_less(T_d, T_i, v[u], vi[u]);
        const uint64_t b = __ballot(cand);
        const long pos = cnt + (long)__builtin_amdgcn_mbcnt_hi(
                                   (uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
        if (cand && pos < kTopkCap) {
          cd[pos] = v[u];
          ci[pos] = vi[u];
        }
        cnt += __popcll(b);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return cnt;
  };
  long cnt = E > 0 ? pack(false) : 0;
  // 3. tighten while the packed entries do not all fit.
  while (cnt > kTopkCap) {
    topk_pop_packed<16>(cd, ci, kTopkCap, kk, lane, rd, ri, T_d, T_i);
    // the next pack's stores to the list must follow these reads
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    cnt = pack(true);
  }
  // 4. the result, in order.
  float ld;
  uint32_t li;
  topk_pop_any(cd, ci, (int)cnt, kk, lane, rd, ri, ld, li);
  np = (int)(cnt < kk ? cnt : kk);
  // the next set's stores to this LDS list must follow these reads
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Stores a selection as a sorted list of lcap entries (the rest empty).
__device__ __forceinline__ void store_list(float *od, uint32_t *oi, int lcap, int np, int lane,
                                           const float (&rd)[2], const uint32_t (&ri)[2]) {
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int idx = 64 * s + lane;
    if (idx < lcap) {
      od[idx] = idx < np ? rd[s] : __builtin_inff();
      oi[idx] = idx < np ? ri[s] : 0xFFFFFFFFu;
    }
  }
}

// Stage 1 source: columns [0, nl) of one row segment through a buffer
// resource (a 32-bit byte offset per lane, clamped to the segment's last
// column in the vector offset itself, so no load leaves the segment whatever
// part of the offset the hardware's range check covers); id = idb + column.
struct SegSource {
  __amdgpu_buffer_rsrc_t rs;
  int nl;
  uint32_t idb;
  int lane;
  __device__ __forceinline__ void pass(long p, float (&v)[kTopkPL], uint32_t (&vi)[kTopkPL]) const {
#pragma unroll
    for (int u = 0; u < kTopkPL; ++u) {
      const int jl = (int)p * kTopkSub + u * 64 + lane;
      const int jc = jl < nl ? jl : nl - 1;
      v[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, 4 * jc, 0, 0));
    }
#pragma unroll
    for (int u = 0; u < kTopkPL; ++u) {
      const int jl = (int)p * kTopkSub + u * 64 + lane;
      vi[u] = jl < nl ? idb + (uint32_t)jl : 0xFFFFFFFFu;
      if (jl >= nl) v[u] = __builtin_inff();
    }
  }
};

// Stage 2 source: E entries of a row's consecutive lists (flat).
struct ListSource {
  const float *d;
  const uint32_t *i;
  long E;
  int lane;
  __device__ __forceinline__ void pass(long p, float (&v)[kTopkPL], uint32_t (&vi)[kTopkPL]) const {
#pragma unroll
    for (int u = 0; u < kTopkPL; ++u) {
      const long e = p * kTopkSub + u * 64 + lane;
      const long ec = e < E ? e : E - 1;
      v[u] = d[ec];
      vi[u] = i[ec];
    }
#pragma unroll
    for (int u = 0; u < kTopkPL; ++u) {
      const long e = p * kTopkSub + u * 64 + lane;
      if (e >= E) {
        v[u] = __builtin_inff();
        vi[u] = 0xFFFFFFFFu;
      }
    }
  }
};

// Stage 1: the m x nseg (row, segment) pairs of kTopkSeg columns, row-major,
// go to the wavefronts in turn (wavefront w of workgroup x takes pairs
// 4x + w, + 4 * gridDim.x, ...); a segment's best kk go to list seg0 + seg of
// its row's `stride` lists of lcap entries (ids = column + id0). Only a row's
// kk best are ever asked for, and they are among its segments' kk best. With
// use_run, list 0 of the row holds the row's kk best over the columns before
// this block (ids below every id here), and its kk-th entry bounds T.
__global__ __launch_bounds__(256) void k_topk_rows(const float *__restrict__ dist, long m, long n,
                                                   long nseg, long stride, long seg0, uint32_t id0,
                                                   int kk, int lcap, int use_run,
                                                   float *__restrict__ seg_d,
                                                   uint32_t *__restrict__ seg_i) {
  __shared__ float c_d[4][kTopkCap];
  __shared__ uint32_t c_i[4][kTopkCap];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const long pairs = m * nseg;
  for (long t = (long)blockIdx.x * 4 + w; t < pairs; t += (long)gridDim.x * 4) {
    const long row = t / nseg, seg = t - row * nseg;
    float *od = seg_d + (row * stride + seg0 + seg) * lcap;
    uint32_t *oi = seg_i + (row * stride + seg0 + seg) * lcap;
    float rd[2];
    uint32_t ri[2];
    int np = 0;
    if (n > 0) {  // no columns (dist may be null): an empty list
      const long c0 = seg * kTopkSeg;
      const int nl = (int)(n - c0 < kTopkSeg ? n - c0 : kTopkSeg);  // >= 1
      SegSource src{__builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(dist + row * n + c0), 0,
                                                      4 * nl, kBufferWord3),
                    nl, (uint32_t)c0 + id0, lane};
      float hd = __builtin_inff();
      uint32_t hi = 0xFFFFFFFFu;
      if (use_run) {
        hd = seg_d[row * stride * lcap + kk - 1];
        hi = seg_i[row * stride * lcap + kk - 1];
      }
      wave_select(src, nl, kk, hd, hi, c_d[w], c_i[w], lane, rd, ri, np);
    }
    store_list(od, oi, lcap, np, lane, rd, ri);
  }
}

// Stage 2: merges lists [l0, l1) of each row (`stride` sorted lists of lcap
// per row), one wavefront per row, into the row's best k_out: into out_d /
// out_i (when out_d is set; ids ~0 -> -1) or back into list 0 as a sorted
// list. A row whose lists past list 0 are all empty, when list 0 is one of
// the inputs and the output, keeps list 0 as it is. Every read of the lists
// precedes the first write.
__global__ __launch_bounds__(256) void k_topk_rows_fold(long m, long stride, long l0, long l1,
                                                        int k_out, int lcap,
                                                        float *__restrict__ seg_d,
                                                        uint32_t *__restrict__ seg_i,
                                                        float *__restrict__ out_d,
                                                        int64_t *__restrict__ out_i) {
  __shared__ float c_d[4][kTopkCap];
  __shared__ uint32_t c_i[4][kTopkCap];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const long row = (long)blockIdx.x * 4 + w;
  if (row >= m) return;  // no barrier below
  float *rd0 = seg_d + row * stride * lcap;
  uint32_t *ri0 = seg_i + row * stride * lcap;
  if (!out_d && l0 == 0) {
    bool fresh = false;  // a real entry in a list other than list 0
    for (long l = 1 + lane; l < l1; l += 64) fresh = fresh || ri0[l * lcap] != 0xFFFFFFFFu;
    if (!__any(fresh)) return;  // wavefront-uniform
  }
  const ListSource src{rd0 + l0 * lcap, ri0 + l0 * lcap, (l1 - l0) * lcap, lane};
  float rd[2];
  uint32_t ri[2];
  int np = 0;
  wave_select(src, src.E, k_out, __builtin_inff(), 0xFFFFFFFFu, c_d[w], c_i[w], lane, rd, ri, np);
  if (out_d) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int idx = 64 * s + lane;
      if (idx < k_out) {
        out_d[row * k_out + idx] = idx < np ? rd[s] : __builtin_inff();
        out_i[row * k_out + idx] = idx < np && ri[s] != 0xFFFFFFFFu ? (int64_t)ri[s] : -1;
      }
    }
  } else {
    store_list(rd0, ri0, lcap, np, lane, rd, ri);
  }
}

}  // namespace

int search_max_k() { return kSearchMaxK; }

hipError_t launch_fill_f32(float *p, long n, float v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const long blocks = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_search(const SearchJobDev *d_jobs, const SearchItem *d_items,
                         const int *d_order, const int *nitems_qr, const SearchGroup *d_groups,
                         int ngroups, int k, int *d_counter, float *d_cand_d,
                         uint32_t *d_cand_i, const int *grid_qr, hipStream_t st) {
  if (k < 1 || k > kSearchMaxK) return hipErrorInvalidValue;
  // Items with 32-query tiles come first in `order`, then the 64-query ones;
  // each kind drains in its own persistent launch.
  int off = 0;
  for (int qr = 1; qr <= 2; ++qr) {
    const int cnt = nitems_qr[qr - 1];
    if (cnt <= 0) continue;
    if (grid_qr[qr - 1] < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, d_counter + (qr - 1));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (qr == 1)
      hipLaunchKernelGGL((k_search<1, kSearchMaxK>), dim3((unsigned)grid_qr[0]), dim3(256), 0, st,
                         d_jobs, d_items, d_order + off, cnt, d_counter, k, d_cand_d, d_cand_i);
    else
      hipLaunchKernelGGL((k_search<2, kSearchMaxK>), dim3((unsigned)grid_qr[1]), dim3(256), 0, st,
                         d_jobs, d_items, d_order + off, cnt, d_counter + 1, k, d_cand_d,
                         d_cand_i);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    off += cnt;
  }
  if (ngroups <= 0) return hipSuccess;
  const long waves = (long)ngroups * 64;
  hipLaunchKernelGGL(k_topk_merge<kSearchMaxK>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0,
                     st, d_jobs, d_groups, ngroups, k, d_cand_d, d_cand_i);
  return hipGetLastError();
}

// Workgroups for m x nseg (row, segment) pairs, four per workgroup; past
// 2^20 workgroups the wavefronts loop.
unsigned topk_rows_grid(long m, long nseg) {
  const long wg = (m * nseg + 3) / 4;
  return (unsigned)(wg < (1L << 20) ? wg : (1L << 20));
}

long topk_rows_segments(long m, long n) {
  (void)m;
  const long seg = (n + kTopkSeg - 1) / kTopkSeg;  // one wavefront per kTopkSeg columns
  return seg > 1 ? seg : 1;
}

hipError_t launch_topk_rows(const float *dist, long m, long n, int k, float *out_d,
                            int64_t *out_i, float *seg_d, uint32_t *seg_i, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  if (k < 1 || k > kTopkMaxK || m > 65535 || n >= 0xFFFFFFFFL) return hipErrorInvalidValue;
  const long nseg = topk_rows_segments(m, n);
  hipLaunchKernelGGL(k_topk_rows, dim3(topk_rows_grid(m, nseg)), dim3(256), 0, st, dist, m, n,
                     nseg, nseg, 0L, 0u, k, k, 0, seg_d, seg_i);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_topk_rows_fold, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, st, m, nseg,
                     0L, nseg, k, k, seg_d, seg_i, out_d, out_i);
  return hipGetLastError();
}

hipError_t launch_topk_rows_chunk(const float *dist, long m, long n, long stride, long seg0,
                                  uint32_t id0, int k, bool use_run, float *seg_d,
                                  uint32_t *seg_i, hipStream_t st) {
  if (m <= 0 || n <= 0) return hipSuccess;
  const long nseg = topk_rows_segments(m, n);
  if (k < 1 || k > kTopkMaxK || m > 65535 || n >= 0xFFFFFFFFL || seg0 + nseg > stride ||
      (use_run && seg0 < 1))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_topk_rows, dim3(topk_rows_grid(m, nseg)), dim3(256), 0, st, dist, m, n,
                     nseg, stride, seg0, id0, k, k, use_run ? 1 : 0, seg_d, seg_i);
  return hipGetLastError();
}

hipError_t launch_topk_rows_fold(long m, long stride, long l0, long l1, int k, float *seg_d,
                                 uint32_t *seg_i, float *out_d, int64_t *out_i, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  if (k < 1 || k > kTopkMaxK || l0 < 0 || l1 > stride || l0 >= l1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_topk_rows_fold, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, st, m, stride,
                     l0, l1, k, k, seg_d, seg_i, out_d, out_i);
  return hipGetLastError();
}

}  // namespace tq
