/* tq_oracle.c -- CPU restatement of the TurboQuant hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see tq_oracle.h). Compiled with -ffp-contract=off
 * and no fast-math: every fused multiply-add the pinned reference build
 * performs is written here as an explicit fmaf(), every other operation is a
 * single IEEE-754 binary32 operation, so the results do not depend on how this
 * file is compiled.
 *
 * Parity pins (tests/test_oracle.py):
 *   - splitmix64: published known-answer values for seed 0;
 *   - signs, codebook, codes, trailers: golden fixtures written by the
 *     reference compiled in place (tests/golden/make_golden.py), and a live
 *     comparison against that build where it is present;
 *   - WHT: closed form H[i][j] = (-1)^popcount(i&j) at small D;
 *   - Lloyd-Max: symmetry, midpoint boundaries, boundary[half-1] = 0;
 *   - encode: zero vector and scaled basis vectors in closed form;
 *   - distances: golden fixtures, and the FP64 restatement of the same
 *     estimator.
 */
#include "tq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_SQRT1_2
#define M_SQRT1_2 0.70710678118654752440
#endif

struct tqo_space {
  size_t input_dim; /* caller's dim */
  size_t dim;       /* padded: next pow2 >= max(dim, 4) */
  int bits;
  int nibble;       /* bits <= 4: two units per byte */
  size_t packed_bytes;
  size_t code_size;
  int num_levels;   /* 2^(bits-1) centroids */
  float scale;      /* sqrt(pi/2)/sqrt(D), fp32 as the reference computes it */
  float *centroids;
  float *boundaries;
  float *rot_signs;
  float *qjl_signs;
};

/* ---- splitmix64 + signs: turbo_quant.h:27-39, :74-86 ------------------- */

uint64_t tqo_splitmix64_next(uint64_t *state) {
  uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

void tqo_signs(size_t d, uint64_t seed, float *out) {
  uint64_t st = seed;
  for (size_t i = 0; i < d; ++i)
    out[i] = (tqo_splitmix64_next(&st) & 1ULL) ? 1.0f : -1.0f;
}

/* ---- Lloyd-Max for N(0,1): space_turbo_quant.h:76-131 ------------------- */

static double lm_phi(double x) { return exp(-0.5 * x * x) / sqrt(2.0 * M_PI); }
static double lm_Phi(double x) { return 0.5 * erfc(-x * M_SQRT1_2); }
static double lm_cond_mean(double a, double b) {
  double denom = lm_Phi(b) - lm_Phi(a);
  if (denom < 1e-15) return 0.5 * (a + b);
  return (lm_phi(a) - lm_phi(b)) / denom;
}

void tqo_lloyd_max(int mse_bits, float *centroids, float *boundaries) {
  const int levels = 1 << mse_bits;
  const int half = levels / 2;
  double *pos_c = (double *)malloc(sizeof(double) * (size_t)half);
  double *new_c = (double *)malloc(sizeof(double) * (size_t)half);
  double *pos_b = (double *)malloc(sizeof(double) * (size_t)(half > 1 ? half - 1 : 1));
  for (int i = 0; i < half; ++i) pos_c[i] = (i + 0.5) * 3.5 / half;
  for (int iter = 0; iter < 1000; ++iter) {
    for (int i = 0; i < half - 1; ++i) pos_b[i] = 0.5 * (pos_c[i] + pos_c[i + 1]);
    double max_delta = 0.0;
    for (int i = 0; i < half; ++i) {
      double lo = (i == 0) ? 0.0 : pos_b[i - 1];
      double hi = (i == half - 1) ? 1e10 : pos_b[i];
      new_c[i] = lm_cond_mean(lo, hi);
      double dlt = fabs(new_c[i] - pos_c[i]);
      if (dlt > max_delta) max_delta = dlt;
    }
    memcpy(pos_c, new_c, sizeof(double) * (size_t)half);
    if (max_delta < 1e-12) break;
  }
  for (int i = 0; i < half; ++i) {
    centroids[half + i] = (float)pos_c[i];
    centroids[half - 1 - i] = -(float)pos_c[i];
  }
  boundaries[half - 1] = 0.0f;
  for (int i = 0; i < half - 1; ++i) {
    double b = 0.5 * (pos_c[i] + pos_c[i + 1]);
    boundaries[half + i] = (float)b;
    boundaries[half - 2 - i] = -(float)b;
  }
  free(pos_c);
  free(new_c);
  free(pos_b);
}

/* ---- reductions ----------------------------------------------------------
 * The pinned build vectorises `s += x[i]*x[i]` (space_turbo_quant.h:342-344,
 * :353-355, :369-371) as: eight lane accumulators a[k] updated with FMA over
 * x[8t+k] for every full block of 8; the lanes folded as
 * ((a0+a4)+(a2+a6)) + ((a1+a5)+(a3+a7)); then, if at least 4 elements remain,
 * four plain products m[k] folded as (m0+m2)+(m1+m3) and added; then the last
 * 0..3 elements by scalar FMA. */
float tqo_sumsq(const float *x, size_t n) {
  float s = 0.0f;
  size_t i = 0;
  size_t n8 = n & ~(size_t)7;
  if (n8) {
    float a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (; i < n8; i += 8)
      for (int k = 0; k < 8; ++k) a[k] = fmaf(x[i + k], x[i + k], a[k]);
    float h0 = a[0] + a[4], h1 = a[1] + a[5], h2 = a[2] + a[6], h3 = a[3] + a[7];
    s = (h0 + h2) + (h1 + h3);
  }
  if (n - i >= 4) {
    float m0 = x[i] * x[i], m1 = x[i + 1] * x[i + 1];
    float m2 = x[i + 2] * x[i + 2], m3 = x[i + 3] * x[i + 3];
    s = s + ((m0 + m2) + (m1 + m3));
    i += 4;
  }
  for (; i < n; ++i) s = fmaf(x[i], x[i], s);
  return s;
}

/* ---- WHT / RHT: turbo_quant.h:46-99 -------------------------------------- */

void tqo_wht(float *data, size_t d) {
  for (size_t step = 1; step < d; step <<= 1) {
    size_t jump = step << 1;
    for (size_t i = 0; i < d; i += jump)
      for (size_t j = 0; j < step; ++j) {
        float a = data[i + j], b = data[i + step + j];
        data[i + j] = a + b;
        data[i + step + j] = a - b;
      }
  }
  float norm = 1.0f / sqrtf((float)d);
  for (size_t i = 0; i < d; ++i) data[i] *= norm;
}

void tqo_rht(float *data, const float *signs, size_t d) {
  for (size_t i = 0; i < d; ++i) data[i] *= signs[i];
  tqo_wht(data, d);
}

/* ---- space: space_turbo_quant.h:233-311 --------------------------------- */

tqo_space *tqo_create(size_t dim, int bits, uint64_t rot_seed, uint64_t qjl_seed) {
  if (dim < 1 || bits < 2 || bits > 9) return NULL;
  tqo_space *s = (tqo_space *)calloc(1, sizeof(tqo_space));
  size_t d = dim < 4 ? 4 : dim, p = 1;
  while (p < d) p <<= 1;
  s->input_dim = dim;
  s->dim = p;
  s->bits = bits;
  s->nibble = bits <= 4;
  s->packed_bytes = s->nibble ? (p + 1) / 2 : p;
  s->code_size = s->packed_bytes + 12;
  s->num_levels = 1 << (bits - 1);
  s->scale = sqrtf((float)M_PI / 2.0f) / sqrtf((float)p);
  s->centroids = (float *)malloc(sizeof(float) * (size_t)s->num_levels);
  s->boundaries = (float *)malloc(sizeof(float) * (size_t)(s->num_levels - 1 > 0 ? s->num_levels - 1 : 1));
  tqo_lloyd_max(bits - 1, s->centroids, s->boundaries);
  s->rot_signs = (float *)malloc(sizeof(float) * p);
  s->qjl_signs = (float *)malloc(sizeof(float) * p);
  tqo_signs(p, rot_seed, s->rot_signs);
  tqo_signs(p, qjl_seed, s->qjl_signs);
  return s;
}

void tqo_destroy(tqo_space *s) {
  if (!s) return;
  free(s->centroids);
  free(s->boundaries);
  free(s->rot_signs);
  free(s->qjl_signs);
  free(s);
}

size_t tqo_padded_dim(const tqo_space *s) { return s->dim; }
size_t tqo_code_size(const tqo_space *s) { return s->code_size; }
size_t tqo_packed_bytes(const tqo_space *s) { return s->packed_bytes; }
float tqo_scale(const tqo_space *s) { return s->scale; }
const float *tqo_centroids(const tqo_space *s) { return s->centroids; }
const float *tqo_boundaries(const tqo_space *s) { return s->boundaries; }
const float *tqo_rot_signs(const tqo_space *s) { return s->rot_signs; }
const float *tqo_qjl_signs(const tqo_space *s) { return s->qjl_signs; }

/* ---- encode: space_turbo_quant.h:332-407 -------------------------------- */

static unsigned quantize(const tqo_space *s, float v) {
  unsigned idx = 0;
  for (int i = 0; i < s->num_levels - 1; ++i) idx += (v > s->boundaries[i]);
  return idx;
}

void tqo_encode(const tqo_space *s, const float *raw, uint8_t *out) {
  const size_t D = s->dim, n_in = s->input_dim;
  float *rot = (float *)calloc(D, sizeof(float));
  uint8_t *idx = (uint8_t *)malloc(D);

  float norm = sqrtf(tqo_sumsq(raw, n_in));
  float inv_norm = (norm > 1e-10f) ? (1.0f / norm) : 0.0f;
  for (size_t i = 0; i < n_in; ++i) rot[i] = raw[i] * inv_norm;
  tqo_rht(rot, s->rot_signs, D);

  float sigma = sqrtf(tqo_sumsq(rot, D) / (float)D);
  if (sigma < 1e-10f) sigma = 1e-10f;
  float inv_sigma = 1.0f / sigma;

  for (size_t i = 0; i < D; ++i) {
    unsigned q = quantize(s, rot[i] * inv_sigma);
    idx[i] = (uint8_t)q;
    /* rotated - centroid*sigma, contracted to one FMA by the pinned build */
    rot[i] = fmaf(-s->centroids[q], sigma, rot[i]);
  }
  float gamma = sqrtf(tqo_sumsq(rot, D));
  tqo_rht(rot, s->qjl_signs, D);

  if (s->nibble) {
    memset(out, 0, s->packed_bytes);
    for (size_t i = 0; i < D; ++i) {
      uint8_t u = (uint8_t)((idx[i] << 1) | (rot[i] >= 0.0f ? 1u : 0u));
      if ((i & 1) == 0)
        out[i >> 1] = u;
      else
        out[i >> 1] |= (uint8_t)(u << 4);
    }
  } else {
    for (size_t i = 0; i < D; ++i)
      out[i] = (uint8_t)((idx[i] << 1) | (rot[i] >= 0.0f ? 1u : 0u));
  }
  float meta[3] = {norm, gamma, sigma};
  memcpy(out + s->packed_bytes, meta, sizeof(meta));
  free(rot);
  free(idx);
}

void tqo_encode_batch(const tqo_space *s, const float *X, size_t n, uint8_t *out) {
#pragma omp parallel for schedule(static) if (n > 64)
  for (long long i = 0; i < (long long)n; ++i)
    tqo_encode(s, X + (size_t)i * s->input_dim, out + (size_t)i * s->code_size);
}

/* space_turbo_quant.h:411-435 */
void tqo_rotated_coords_batch(const tqo_space *s, const float *X, size_t n, float *out) {
  const size_t D = s->dim, n_in = s->input_dim;
  for (size_t r = 0; r < n; ++r) {
    const float *raw = X + r * n_in;
    float *dst = out + r * D;
    float nsq = tqo_sumsq(raw, n_in);
    float inv_norm = (nsq > 1e-20f) ? (1.0f / sqrtf(nsq)) : 0.0f;
    memset(dst, 0, D * sizeof(float));
    for (size_t j = 0; j < n_in; ++j) dst[j] = raw[j] * inv_norm;
    tqo_rht(dst, s->rot_signs, D);
    float var = tqo_sumsq(dst, D);
    float inv_sigma = (var > 1e-20f) ? sqrtf((float)D / var) : 1.0f;
    for (size_t j = 0; j < D; ++j) dst[j] *= inv_sigma;
  }
}

/* ---- query preparation: space_turbo_quant.h:439-460 -------------------- */

void tqo_prepare_query(const tqo_space *s, const float *q, float *q_rot, float *s_q,
                       float *q_norm_sq, float *q_norm) {
  const size_t D = s->dim, n_in = s->input_dim;
  /* q_norm_sq accumulates in the returned struct (:443-445), which may alias
   * the query, so the pinned build keeps this loop scalar: one FMA chain. */
  float nsq = 0.0f;
  for (size_t i = 0; i < n_in; ++i) nsq = fmaf(q[i], q[i], nsq);
  float qn = sqrtf(nsq);
  float q_inv = (qn > 1e-10f) ? (1.0f / qn) : 0.0f;
  memset(q_rot, 0, D * sizeof(float));
  for (size_t i = 0; i < n_in; ++i) q_rot[i] = q[i] * q_inv;
  tqo_rht(q_rot, s->rot_signs, D);
  memcpy(s_q, q_rot, D * sizeof(float));
  tqo_rht(s_q, s->qjl_signs, D);
  *q_norm_sq = nsq;
  *q_norm = qn;
}

/* ---- asymmetric estimator: space_turbo_quant.h:659-686, :951-983 -------- */

static inline unsigned code_unit(const tqo_space *s, const uint8_t *packed, size_t i) {
  if (s->nibble) {
    uint8_t b = packed[i >> 1];
    return (i & 1) ? (b >> 4) : (b & 0x0F);
  }
  return packed[i];
}

static inline void code_meta(const tqo_space *s, const uint8_t *code, float meta[3]) {
  memcpy(meta, code + s->packed_bytes, 3 * sizeof(float));
}

float tqo_distance_prepared(const tqo_space *s, const float *q_rot, const float *s_q,
                            float q_norm_sq, float q_norm, const uint8_t *code) {
  float meta[3];
  code_meta(s, code, meta);
  const float x_norm = meta[0], gamma = meta[1], sigma = meta[2];
  float ip_mse = 0.0f, dot_qjl = 0.0f;
  for (size_t i = 0; i < s->dim; ++i) {
    unsigned u = code_unit(s, code, i);
    ip_mse += q_rot[i] * s->centroids[u >> 1];
    dot_qjl += s_q[i] * ((u & 1) ? 1.0f : -1.0f);
  }
  ip_mse *= sigma;
  const float correction = s->scale * gamma * dot_qjl;
  const float ip = (ip_mse + correction) * x_norm * q_norm;
  float d = q_norm_sq + x_norm * x_norm - 2.0f * ip;
  return d > 0.0f ? d : 0.0f;
}

float tqo_distance(const tqo_space *s, const float *q, const uint8_t *code) {
  float *qr = (float *)malloc(2 * s->dim * sizeof(float));
  float nsq, qn;
  tqo_prepare_query(s, q, qr, qr + s->dim, &nsq, &qn);
  float d = tqo_distance_prepared(s, qr, qr + s->dim, nsq, qn, code);
  free(qr);
  return d;
}

void tqo_distance_1_to_n(const tqo_space *s, const float *q, const uint8_t *codes,
                         size_t n, float *out) {
  float *qr = (float *)malloc(2 * s->dim * sizeof(float));
  float nsq, qn;
  tqo_prepare_query(s, q, qr, qr + s->dim, &nsq, &qn);
#pragma omp parallel for schedule(static) if (n > 64)
  for (long long i = 0; i < (long long)n; ++i)
    out[i] = tqo_distance_prepared(s, qr, qr + s->dim, nsq, qn,
                                   codes + (size_t)i * s->code_size);
  free(qr);
}

void tqo_distance_m_to_n(const tqo_space *s, const float *Q, size_t m,
                         const uint8_t *codes, size_t n, float *out) {
#pragma omp parallel for schedule(static) if (m > 64)
  for (long long qi = 0; qi < (long long)m; ++qi) {
    float *qr = (float *)malloc(2 * s->dim * sizeof(float));
    float nsq, qn;
    tqo_prepare_query(s, Q + (size_t)qi * s->input_dim, qr, qr + s->dim, &nsq, &qn);
    for (size_t i = 0; i < n; ++i)
      out[(size_t)qi * n + i] =
          tqo_distance_prepared(s, qr, qr + s->dim, nsq, qn, codes + i * s->code_size);
    free(qr);
  }
}

/* ---- symmetric MSE-only: space_turbo_quant.h:688-714, :985-1019 --------- */

float tqo_distance_symmetric(const tqo_space *s, const uint8_t *a, const uint8_t *b) {
  float ma[3], mb[3];
  code_meta(s, a, ma);
  code_meta(s, b, mb);
  const float na = ma[0], sa = ma[2], nb = mb[0], sb = mb[2];
  float ip_rot = 0.0f;
  for (size_t i = 0; i < s->dim; ++i) {
    float ca = s->centroids[code_unit(s, a, i) >> 1] * sa;
    float cb = s->centroids[code_unit(s, b, i) >> 1] * sb;
    ip_rot += ca * cb;
  }
  const float ip = ip_rot * na * nb;
  float d = na * na + nb * nb - 2.0f * ip;
  return d > 0.0f ? d : 0.0f;
}

void tqo_distance_m_to_n_symmetric(const tqo_space *s, const uint8_t *A, size_t m,
                                   const uint8_t *B, size_t n, float *out) {
#pragma omp parallel for schedule(static) if (m > 64)
  for (long long i = 0; i < (long long)m; ++i)
    for (size_t j = 0; j < n; ++j)
      out[(size_t)i * n + j] = tqo_distance_symmetric(
          s, A + (size_t)i * s->code_size, B + j * s->code_size);
}

/* ---- symmetric with QJL correction: space_turbo_quant.h:731-800 -------- */

float tqo_distance_symmetric_full(const tqo_space *s, const uint8_t *a, const uint8_t *b) {
  const size_t D = s->dim;
  float ma[3], mb[3];
  code_meta(s, a, ma);
  code_meta(s, b, mb);
  float *ra = (float *)malloc(D * sizeof(float)), *rb = (float *)malloc(D * sizeof(float));
  float ip_mse = 0.0f;
  for (size_t i = 0; i < D; ++i) {
    float ca = s->centroids[code_unit(s, a, i) >> 1] * ma[2];
    float cb = s->centroids[code_unit(s, b, i) >> 1] * mb[2];
    ip_mse += ca * cb;
    ra[i] = ca;
    rb[i] = cb;
  }
  tqo_rht(ra, s->qjl_signs, D);
  tqo_rht(rb, s->qjl_signs, D);
  float d_ab = 0.0f, d_ba = 0.0f, d_ss = 0.0f;
  for (size_t i = 0; i < D; ++i) {
    float sa = (code_unit(s, a, i) & 1) ? 1.0f : -1.0f;
    float sb = (code_unit(s, b, i) & 1) ? 1.0f : -1.0f;
    d_ab += ra[i] * sb;
    d_ba += sa * rb[i];
    d_ss += sa * sb;
  }
  float t2 = s->scale * mb[1] * d_ab;
  float t3 = s->scale * ma[1] * d_ba;
  const float pi_over_2d = (float)M_PI / (2.0f * (float)D);
  float t4 = pi_over_2d * ma[1] * mb[1] * d_ss;
  float corr = t2 + t3 + t4;
  float mc = fabsf(ip_mse);
  corr = corr < -mc ? -mc : (corr > mc ? mc : corr);
  float ip = (ip_mse + corr) * ma[0] * mb[0];
  float d = ma[0] * ma[0] + mb[0] * mb[0] - 2.0f * ip;
  free(ra);
  free(rb);
  return d > 0.0f ? d : 0.0f;
}

/* prepareSymmetricQuery (space_turbo_quant.h:498-532) + the prepared full
 * distance (distBuildFullPreparedScalar, :818-891): the query side is
 * reconstructed and Hadamard-transformed once; all four dot products run in
 * one loop over the post-Hadamard vectors. */
float tqo_distance_symmetric_full_prepared(const tqo_space *s, const uint8_t *a,
                                           const uint8_t *b) {
  const size_t D = s->dim;
  float ma[3], mb[3];
  code_meta(s, a, ma);
  code_meta(s, b, mb);
  float *rq = (float *)malloc(D * sizeof(float)), *sq = (float *)malloc(D * sizeof(float));
  float *rb = (float *)malloc(D * sizeof(float));
  for (size_t i = 0; i < D; ++i) {
    unsigned u = code_unit(s, a, i);
    rq[i] = s->centroids[u >> 1] * ma[2];
    sq[i] = (u & 1) ? 1.0f : -1.0f;
    rb[i] = s->centroids[code_unit(s, b, i) >> 1] * mb[2];
  }
  tqo_rht(rq, s->qjl_signs, D);
  tqo_rht(rb, s->qjl_signs, D);
  const float scale_gamma_q = s->scale * ma[1];
  const float pi_over_2d = (float)M_PI / (2.0f * (float)D);
  const float pi_over_2d_gamma_q = pi_over_2d * ma[1];
  float d_rr = 0.0f, d_rs = 0.0f, d_sr = 0.0f, d_ss = 0.0f;
  for (size_t i = 0; i < D; ++i) {
    float sb = (code_unit(s, b, i) & 1) ? 1.0f : -1.0f;
    d_rr += rq[i] * rb[i];
    d_rs += rq[i] * sb;
    d_sr += sq[i] * rb[i];
    d_ss += sq[i] * sb;
  }
  float t2 = s->scale * mb[1] * d_rs;
  float t3 = scale_gamma_q * d_sr;
  float t4 = pi_over_2d_gamma_q * mb[1] * d_ss;
  float corr = t2 + t3 + t4;
  float mc = fabsf(d_rr);
  corr = corr < -mc ? -mc : (corr > mc ? mc : corr);
  float ip = (d_rr + corr) * ma[0] * mb[0];
  float d = ma[0] * ma[0] + mb[0] * mb[0] - 2.0f * ip;
  free(rq);
  free(sq);
  free(rb);
  return d > 0.0f ? d : 0.0f;
}

/* space_turbo_quant.h:876-924 */
float tqo_distance_symmetric_light(const tqo_space *s, const uint8_t *a, const uint8_t *b) {
  const size_t D = s->dim;
  float ma[3], mb[3];
  code_meta(s, a, ma);
  code_meta(s, b, mb);
  float ip_mse = 0.0f;
  int agree = 0;
  for (size_t i = 0; i < D; ++i) {
    unsigned ua = code_unit(s, a, i), ub = code_unit(s, b, i);
    ip_mse += (s->centroids[ua >> 1] * ma[2]) * (s->centroids[ub >> 1] * mb[2]);
    agree += ((ua ^ ub) & 1) == 0 ? 1 : 0;
  }
  float dot_signs = (float)(2 * agree - (int)D);
  const float pi_over_2d = (float)M_PI / (2.0f * (float)D);
  float corr = pi_over_2d * ma[1] * mb[1] * dot_signs;
  float mc = fabsf(ip_mse);
  corr = corr < -mc ? -mc : (corr > mc ? mc : corr);
  float ip = (ip_mse + corr) * ma[0] * mb[0];
  float d = ma[0] * ma[0] + mb[0] * mb[0] - 2.0f * ip;
  return d > 0.0f ? d : 0.0f;
}

/* ---- FP64 restatement of the same estimators ---------------------------- */

static void wht64(double *x, size_t d) {
  for (size_t step = 1; step < d; step <<= 1)
    for (size_t i = 0; i < d; i += step << 1)
      for (size_t j = 0; j < step; ++j) {
        double a = x[i + j], b = x[i + step + j];
        x[i + j] = a + b;
        x[i + step + j] = a - b;
      }
  double nrm = 1.0 / sqrt((double)d);
  for (size_t i = 0; i < d; ++i) x[i] *= nrm;
}

void tqo_distance_m_to_n_f64(const tqo_space *s, const float *Q, size_t m,
                             const uint8_t *codes, size_t n, double *out) {
  const size_t D = s->dim, n_in = s->input_dim;
  const double scale = sqrt(M_PI / 2.0) / sqrt((double)D);
#pragma omp parallel for schedule(static)
  for (long long qi = 0; qi < (long long)m; ++qi) {
    const float *q = Q + (size_t)qi * n_in;
    double *qr = (double *)calloc(D, sizeof(double));
    double *sq = (double *)calloc(D, sizeof(double));
    double nsq = 0.0;
    for (size_t i = 0; i < n_in; ++i) nsq += (double)q[i] * (double)q[i];
    double qn = sqrt(nsq);
    double inv = qn > 1e-10 ? 1.0 / qn : 0.0;
    for (size_t i = 0; i < n_in; ++i) qr[i] = (double)q[i] * inv * s->rot_signs[i];
    wht64(qr, D);
    for (size_t i = 0; i < D; ++i) sq[i] = qr[i] * s->qjl_signs[i];
    wht64(sq, D);
    for (size_t j = 0; j < n; ++j) {
      const uint8_t *c = codes + j * s->code_size;
      float meta[3];
      code_meta(s, c, meta);
      double mse = 0.0, dot = 0.0;
      for (size_t i = 0; i < D; ++i) {
        unsigned u = code_unit(s, c, i);
        mse += qr[i] * (double)s->centroids[u >> 1];
        dot += (u & 1) ? sq[i] : -sq[i];
      }
      double ip = ((double)meta[2] * mse + scale * (double)meta[1] * dot) *
                  (double)meta[0] * qn;
      double d = nsq + (double)meta[0] * (double)meta[0] - 2.0 * ip;
      out[(size_t)qi * n + j] = d > 0.0 ? d : 0.0;
    }
    free(qr);
    free(sq);
  }
}

void tqo_distance_m_to_n_symmetric_f64(const tqo_space *s, const uint8_t *A, size_t m,
                                       const uint8_t *B, size_t n, double *out) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)m; ++i) {
    const uint8_t *a = A + (size_t)i * s->code_size;
    float ma[3];
    code_meta(s, a, ma);
    for (size_t j = 0; j < n; ++j) {
      const uint8_t *b = B + j * s->code_size;
      float mb[3];
      code_meta(s, b, mb);
      double acc = 0.0;
      for (size_t k = 0; k < s->dim; ++k)
        acc += (double)s->centroids[code_unit(s, a, k) >> 1] *
               (double)s->centroids[code_unit(s, b, k) >> 1];
      double ip = acc * (double)ma[2] * (double)mb[2] * (double)ma[0] * (double)mb[0];
      double d = (double)ma[0] * ma[0] + (double)mb[0] * mb[0] - 2.0 * ip;
      out[(size_t)i * n + j] = d > 0.0 ? d : 0.0;
    }
  }
}

/* FP64 restatements of the symmetric QJL-corrected estimators (full,
 * space_turbo_quant.h:731-800, whose prepared form :818-891 is the same
 * quantity; light, :876-924), over m x n code pairs. */
void tqo_distance_m_to_n_symmetric_full_f64(const tqo_space *s, const uint8_t *A, size_t m,
                                            const uint8_t *B, size_t n, int light,
                                            double *out) {
  const size_t D = s->dim;
  const double scale = sqrt(M_PI / 2.0) / sqrt((double)D);
  const double pi_over_2d = M_PI / (2.0 * (double)D);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)m; ++i) {
    const uint8_t *a = A + (size_t)i * s->code_size;
    float ma[3];
    code_meta(s, a, ma);
    double *ra = (double *)malloc(D * sizeof(double)), *rb = (double *)malloc(D * sizeof(double));
    double *ha = (double *)malloc(D * sizeof(double)), *hb = (double *)malloc(D * sizeof(double));
    for (size_t k = 0; k < D; ++k) {
      ra[k] = (double)s->centroids[code_unit(s, a, k) >> 1] * (double)ma[2];
      ha[k] = ra[k] * s->qjl_signs[k];
    }
    wht64(ha, D);
    for (size_t j = 0; j < n; ++j) {
      const uint8_t *b = B + j * s->code_size;
      float mb[3];
      code_meta(s, b, mb);
      for (size_t k = 0; k < D; ++k) {
        rb[k] = (double)s->centroids[code_unit(s, b, k) >> 1] * (double)mb[2];
        hb[k] = rb[k] * s->qjl_signs[k];
      }
      wht64(hb, D);
      double ip_mse = 0.0, d_ab = 0.0, d_ba = 0.0, d_ss = 0.0;
      for (size_t k = 0; k < D; ++k) {
        const double sa = (code_unit(s, a, k) & 1) ? 1.0 : -1.0;
        const double sb = (code_unit(s, b, k) & 1) ? 1.0 : -1.0;
        ip_mse += ra[k] * rb[k];
        d_ab += ha[k] * sb;
        d_ba += sa * hb[k];
        d_ss += sa * sb;
      }
      const double t4 = pi_over_2d * (double)ma[1] * (double)mb[1] * d_ss;
      double corr = light ? t4
                          : scale * (double)mb[1] * d_ab + scale * (double)ma[1] * d_ba + t4;
      const double mc = fabs(ip_mse);
      corr = corr < -mc ? -mc : (corr > mc ? mc : corr);
      const double ip = (ip_mse + corr) * (double)ma[0] * (double)mb[0];
      const double d = (double)ma[0] * ma[0] + (double)mb[0] * mb[0] - 2.0 * ip;
      out[(size_t)i * n + j] = d > 0.0 ? d : 0.0;
    }
    free(ra);
    free(rb);
    free(ha);
    free(hb);
  }
}
