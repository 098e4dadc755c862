// ref_harness_sym.cpp -- extern "C" entry to the reference's prepared
// full-symmetric batch (space_turbo_quant.h:561-572). Test infrastructure.
//
// Kept in its own translation unit: adding it to ref_harness.cpp changes
// GCC's inlining there, and prepareQuery's rounding (its ||q||^2 loop may or
// may not be vectorised, depending on whether the compiler can see that the
// returned struct does not alias the query) changes with it.
#include <cstddef>
#include <cstdint>

#include "turboquant/space_turbo_quant.h"

using turboquant::TurboQuantSpace;

extern "C" void tqref_distance_m_to_n_symmetric_full_prepared(void *h, const uint8_t *A, size_t m,
                                                              const uint8_t *B, size_t n,
                                                              float *out) {
  static_cast<TurboQuantSpace *>(h)->distanceBatchMToNSymmetricFullPrepared(A, m, B, n, out);
}
