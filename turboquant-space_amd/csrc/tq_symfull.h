// tq_symfull.h -- symmetric distances with QJL correction (full / light).
#pragma once
#include "tq_common.h"

namespace tq {

// Prepared symmetric row: [r (D) | G (D) | h' (D) | n^2 | n | 0 | 0].
__host__ __device__ inline long sym_prep_stride(int D) { return 3L * D + 4; }

bool sym_qjl_supported(const SpaceDev &s);
// Rows of `sp` for m codes (light: the light estimator's row).
hipError_t launch_prepare_sym(const SpaceDev &s, const uint8_t *codes, long m, bool light,
                              float *sp, hipStream_t st);
// out[q*n + j] = full / light symmetric distance of prepared row q and code j.
hipError_t launch_sym_qjl(const SpaceDev &s, const float *sp, long m, const uint8_t *codes,
                          long n, float *out, hipStream_t st);

}  // namespace tq
