// tq_mfma.h -- fp16 hi+lo MFMA form of the m x n asymmetric scan.
#pragma once
#include "tq_common.h"

namespace tq {
// out[q*n + j] for q < m, j < n from prepared queries (prep_stride floats
// each) and codes; returns hipErrorNotSupported for padded dims it does not
// handle. qjl = false: the queries' s_q are all zero (symmetric MSE
// estimator) and the QJL pass is skipped.
hipError_t launch_scan_mfma(const SpaceDev &s, const float *prep, long m,
                            const uint8_t *codes, long n, float *out,
                            hipStream_t st, bool qjl = true);
// Query hi/lo planes [3][m][D] fp16 for the MFMA kernels (4-bit spaces store
// each 8-coordinate group in the nibble decoder's order).
hipError_t launch_split_queries(const SpaceDev &s, const float *prep, long m,
                                _Float16 *planes, hipStream_t st);
// The MFMA distances [m][n] of queries already split into `planes` (the
// layout launch_split_queries writes, m <= 65535 * 64), no allocation.
hipError_t launch_mfma_planes(const SpaceDev &s, const _Float16 *planes, const float *prep, long m,
                              const uint8_t *codes, long n, float *out, hipStream_t st,
                              bool qjl = true);
}  // namespace tq
