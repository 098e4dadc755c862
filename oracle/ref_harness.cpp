// ref_harness.cpp -- test infrastructure only.
//
// A thin extern "C" shim over the reference's own header-only library,
// compiled in place from include/turboquant/space_turbo_quant.h with the
// reference's Release flags (CMakeLists.txt:122-135, TURBOQUANT_PORTABLE=ON:
// -O3 -ffast-math -fno-finite-math-only -march=x86-64-v3). The built library
// lands in oracle/_ref/ and is used only to (1) validate the C restatement in
// oracle/tq_oracle.c, (2) write the golden fixtures under tests/golden/ and
// (3) serve as the CPU baseline timed beside the GPU. It is never part of the
// product path.
#include "turboquant/space_turbo_quant.h"

#include <cstdint>
#include <cstring>

using turboquant::TurboQuantSpace;

extern "C" {

void *tqref_create(size_t dim, int bits, uint64_t rot_seed, uint64_t qjl_seed,
                   int num_threads) {
  return new TurboQuantSpace(dim, bits, rot_seed, qjl_seed, num_threads);
}

void tqref_destroy(void *h) { delete static_cast<TurboQuantSpace *>(h); }

size_t tqref_code_size(void *h) {
  return static_cast<TurboQuantSpace *>(h)->codeSizeBytes();
}

size_t tqref_padded_dim(void *h) {
  return static_cast<TurboQuantSpace *>(h)->paddedDim();
}

int tqref_num_threads(void *h) {
  return static_cast<TurboQuantSpace *>(h)->numThreads();
}

float tqref_scale(void *h) { return static_cast<TurboQuantSpace *>(h)->scale(); }

void tqref_encode_batch(void *h, const float *X, size_t n, uint8_t *out) {
  static_cast<TurboQuantSpace *>(h)->encodeBatch(X, n, out);
}

void tqref_rotated_coords_batch(void *h, const float *X, size_t n, float *out) {
  static_cast<TurboQuantSpace *>(h)->rotatedCoordsBatch(X, n, out);
}

void tqref_prepare_query(void *h, const float *q, float *q_rot, float *s_q,
                         float *q_norm_sq, float *q_norm) {
  auto *s = static_cast<TurboQuantSpace *>(h);
  auto pq = s->prepareQuery(q);
  std::memcpy(q_rot, pq.q_rot.data(), pq.q_rot.size() * sizeof(float));
  std::memcpy(s_q, pq.s_q.data(), pq.s_q.size() * sizeof(float));
  *q_norm_sq = pq.q_norm_sq;
  *q_norm = pq.q_norm;
}

float tqref_distance(void *h, const float *q, const uint8_t *code) {
  return static_cast<TurboQuantSpace *>(h)->distance(q, code);
}

float tqref_distance_symmetric(void *h, const uint8_t *a, const uint8_t *b) {
  return static_cast<TurboQuantSpace *>(h)->distanceSymmetric(a, b);
}

float tqref_distance_symmetric_full(void *h, const uint8_t *a, const uint8_t *b) {
  return static_cast<TurboQuantSpace *>(h)->distanceSymmetricFull(a, b);
}

float tqref_distance_symmetric_light(void *h, const uint8_t *a, const uint8_t *b) {
  return static_cast<TurboQuantSpace *>(h)->distanceSymmetricLight(a, b);
}

void tqref_distance_1_to_n(void *h, const float *q, const uint8_t *codes,
                           size_t n, float *out) {
  static_cast<TurboQuantSpace *>(h)->distanceBatch1ToN(q, codes, n, out);
}

void tqref_distance_m_to_n(void *h, const float *Q, size_t m,
                           const uint8_t *codes, size_t n, float *out) {
  static_cast<TurboQuantSpace *>(h)->distanceBatchMToN(Q, m, codes, n, out);
}

void tqref_distance_m_to_n_symmetric(void *h, const uint8_t *A, size_t m,
                                     const uint8_t *B, size_t n, float *out) {
  static_cast<TurboQuantSpace *>(h)->distanceBatchMToNSymmetric(A, m, B, n, out);
}

void tqref_distance_m_to_n_symmetric_full(void *h, const uint8_t *A, size_t m,
                                          const uint8_t *B, size_t n,
                                          float *out) {
  static_cast<TurboQuantSpace *>(h)->distanceBatchMToNSymmetricFull(A, m, B, n,
                                                                    out);
}

void tqref_distance_m_to_n_symmetric_light(void *h, const uint8_t *A, size_t m,
                                           const uint8_t *B, size_t n,
                                           float *out) {
  static_cast<TurboQuantSpace *>(h)->distanceBatchMToNSymmetricLight(A, m, B, n,
                                                                     out);
}

// Lloyd-Max table exactly as the space builds it (bits_per_coord - 1 MSE bits).
void tqref_codebook(int mse_bits, float *centroids, float *boundaries) {
  auto t = turboquant::computeLloydMax(mse_bits);
  std::memcpy(centroids, t.centroids.data(), t.centroids.size() * sizeof(float));
  std::memcpy(boundaries, t.boundaries.data(),
              t.boundaries.size() * sizeof(float));
}

void tqref_signs(size_t d, uint64_t seed, float *out) {
  auto s = turboquant::generateSigns(d, seed);
  std::memcpy(out, s.data(), d * sizeof(float));
}

uint64_t tqref_splitmix64(uint64_t seed, int k) {
  turboquant::RndGen64 g(seed);
  uint64_t v = 0;
  for (int i = 0; i <= k; ++i) v = g.next();
  return v;
}

}  // extern "C"
