/* Sanitizer driver for the CPU oracle (test infrastructure): built together
 * with oracle/tq_oracle.c under -fsanitize=address,undefined by
 * tests/test_sanitize_cpu.py, it runs every oracle entry point over a few
 * spaces (tiny, odd, non-power-of-two and large padded dims; nibble and byte
 * layouts; edge rows) so out-of-bounds accesses, leaks and undefined
 * behaviour in the restatement are reported. Exit status 0 and a checksum
 * line on success. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../oracle/tq_oracle.h"

static uint64_t rng = 0x1234567ull;
static float frand(void) {
  rng = rng * 6364136223846793005ull + 1442695040888963407ull;
  return (float)((int64_t)(rng >> 11) % 2000001 - 1000000) * 1e-6f;
}

static double run(size_t dim, int bits) {
  tqo_space *s = tqo_create(dim, bits, 42, 137);
  if (!s) return -1.0;
  const size_t D = tqo_padded_dim(s), cs = tqo_code_size(s);
  const size_t n = 9, m = 3;
  float *X = malloc(n * dim * sizeof(float)), *Q = malloc(m * dim * sizeof(float));
  for (size_t i = 0; i < n * dim; ++i) X[i] = frand();
  for (size_t i = 0; i < m * dim; ++i) Q[i] = frand();
  memset(X, 0, dim * sizeof(float));                            /* zero row */
  for (size_t i = 0; i < dim; ++i) X[dim + i] = i == dim / 2 ? 3.0f : 0.0f; /* basis row */
  for (size_t i = 0; i < dim; ++i) X[2 * dim + i] = 1e30f;      /* huge row */
  uint8_t *codes = malloc(n * cs);
  tqo_encode_batch(s, X, n, codes);
  uint8_t *one = malloc(cs);
  tqo_encode(s, X + 3 * dim, one);
  float *rot = malloc(n * D * sizeof(float));
  tqo_rotated_coords_batch(s, X, n, rot);
  float *qr = malloc(D * sizeof(float)), *sq = malloc(D * sizeof(float)), qns, qn;
  tqo_prepare_query(s, Q, qr, sq, &qns, &qn);
  double acc = tqo_distance_prepared(s, qr, sq, qns, qn, codes) + tqo_distance(s, Q, one);
  float *d1 = malloc(n * sizeof(float)), *dm = malloc(m * n * sizeof(float));
  tqo_distance_1_to_n(s, Q, codes, n, d1);
  tqo_distance_m_to_n(s, Q, m, codes, n, dm);
  tqo_distance_m_to_n_symmetric(s, codes, m, codes, n, dm);
  double *d64 = malloc(m * n * sizeof(double));
  tqo_distance_m_to_n_f64(s, Q, m, codes, n, d64);
  for (size_t i = 0; i < m * n; ++i) acc += d64[i] * 1e-9 + dm[i] * 1e-9;
  tqo_distance_m_to_n_symmetric_f64(s, codes, m, codes, n, d64);
  tqo_distance_m_to_n_symmetric_full_f64(s, codes, m, codes, n, 0, d64);
  tqo_distance_m_to_n_symmetric_full_f64(s, codes, m, codes, n, 1, d64);
  for (size_t j = 0; j < n; ++j) {
    acc += tqo_distance_symmetric(s, codes, codes + j * cs) * 1e-9;
    acc += tqo_distance_symmetric_full(s, codes + cs, codes + j * cs) * 1e-9;
    acc += tqo_distance_symmetric_light(s, codes + cs, codes + j * cs) * 1e-9;
    acc += tqo_distance_symmetric_full_prepared(s, codes + cs, codes + j * cs) * 1e-9;
    acc += d1[j] * 1e-9 + rot[j * D] * 1e-9;
  }
  free(X); free(Q); free(codes); free(one); free(rot); free(qr); free(sq);
  free(d1); free(dm); free(d64);
  tqo_destroy(s);
  return acc;
}

int main(void) {
  static const int shapes[][2] = {{1, 4}, {3, 3}, {4, 2}, {7, 9}, {13, 5}, {128, 4},
                                  {200, 8}, {768, 4}, {1000, 6}, {5000, 4}, {8192, 8}};
  double total = 0.0;
  for (size_t k = 0; k < sizeof(shapes) / sizeof(shapes[0]); ++k) {
    const double r = run((size_t)shapes[k][0], shapes[k][1]);
    if (!(r == r)) total += 1.0; else total += r;
  }
  float c[128], b[127];
  for (int mb = 1; mb <= 7; ++mb) tqo_lloyd_max(mb, c, b);
  float sig[64];
  tqo_signs(64, 7, sig);
  uint64_t st = 0;
  const uint64_t k0 = tqo_splitmix64_next(&st);
  float w[16];
  for (int i = 0; i < 16; ++i) w[i] = (float)i;
  tqo_wht(w, 16);
  tqo_rht(w, sig, 16);
  printf("asan-ok checksum %.6e %llx %.3f\n", total, (unsigned long long)k0, tqo_sumsq(w, 16));
  return 0;
}
