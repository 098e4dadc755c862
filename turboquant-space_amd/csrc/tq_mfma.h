// tq_mfma.h -- fp16 hi+lo MFMA form of the m x n asymmetric scan.
#pragma once
#include "tq_common.h"

namespace tq {
// out[q*n + j] for q < m, j < n from prepared queries (prep_stride floats
// each) and codes. Padded dims below 32 (no MFMA tile) run the plain fp32
// kernels. qjl = false: the queries' s_q are all zero (symmetric MSE
// estimator) and the QJL passes are skipped.
hipError_t launch_scan_mfma(const SpaceDev &s, const float *prep, long m,
                            const uint8_t *codes, long n, float *out,
                            hipStream_t st, bool qjl = true);
// Bytes of the split query planes of m queries (layout in tq_mfma_dev.h).
size_t split_planes_bytes(const SpaceDev &s, long m, bool qjl = true);
// Writes the split query planes of m prepared queries.
hipError_t launch_split_queries(const SpaceDev &s, const float *prep, long m,
                                _Float16 *planes, hipStream_t st, bool qjl = true);
// The MFMA distances [m][n] of queries already split into `planes` (written
// by launch_split_queries with the same qjl, m <= 65535 * 64), no allocation.
// qt: queries per workgroup tile, 32 or 64 (0: 32 for m <= 32, else 64);
// every element is the same whichever tile computes it. thr (coarse form
// only, for a top-k over the block): query q's running k-th best distance at
// thr[q * thr_stride]; near pairs that provably cannot beat it are left
// uncorrected (their entries are then only good for that top-k).
hipError_t launch_mfma_planes(const SpaceDev &s, const _Float16 *planes, const float *prep, long m,
                              const uint8_t *codes, long n, float *out, hipStream_t st,
                              bool qjl = true, int qt = 0, const float *thr = nullptr,
                              long thr_stride = 0);
}  // namespace tq
