/* tq_oracle.h -- CPU restatement of the TurboQuant encode + distance hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the GPU path is compared
 * against; the product never links or calls it. Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it.
 *
 * Every function follows a cited reference location (paths relative to the
 * reference repository root). The fp32 encoder reproduces the pinned reference
 * build bit-for-bit (g++ 11.4, -O3 -ffast-math -fno-finite-math-only
 * -march=x86-64-v3): see tqo_sumsq() for the reduction order that build uses.
 */
#ifndef TQ_ORACLE_H
#define TQ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tqo_space tqo_space;

/* splitmix64, include/turboquant/turbo_quant.h:27-39 */
uint64_t tqo_splitmix64_next(uint64_t *state);
/* include/turboquant/turbo_quant.h:74-86 */
void tqo_signs(size_t d, uint64_t seed, float *out);
/* include/turboquant/space_turbo_quant.h:76-131 (mse_bits = bits_per_coord-1) */
void tqo_lloyd_max(int mse_bits, float *centroids, float *boundaries);
/* Reduction order of the pinned build for `s += x[i]*x[i]` loops
 * (space_turbo_quant.h:342-344, :353-355, :369-371). */
float tqo_sumsq(const float *x, size_t n);
/* include/turboquant/turbo_quant.h:46-72 */
void tqo_wht(float *data, size_t d);
/* include/turboquant/turbo_quant.h:89-99 */
void tqo_rht(float *data, const float *signs, size_t d);

/* include/turboquant/space_turbo_quant.h:250-311 */
tqo_space *tqo_create(size_t dim, int bits, uint64_t rot_seed, uint64_t qjl_seed);
void tqo_destroy(tqo_space *s);
size_t tqo_padded_dim(const tqo_space *s);
size_t tqo_code_size(const tqo_space *s);
size_t tqo_packed_bytes(const tqo_space *s);
float tqo_scale(const tqo_space *s);
const float *tqo_centroids(const tqo_space *s);
const float *tqo_boundaries(const tqo_space *s);
const float *tqo_rot_signs(const tqo_space *s);
const float *tqo_qjl_signs(const tqo_space *s);

/* space_turbo_quant.h:339-396 and :400-407 */
void tqo_encode(const tqo_space *s, const float *x, uint8_t *out);
void tqo_encode_batch(const tqo_space *s, const float *X, size_t n, uint8_t *out);
/* space_turbo_quant.h:411-435 */
void tqo_rotated_coords_batch(const tqo_space *s, const float *X, size_t n, float *out);
/* space_turbo_quant.h:439-460 */
void tqo_prepare_query(const tqo_space *s, const float *q, float *q_rot, float *s_q,
                       float *q_norm_sq, float *q_norm);

/* Asymmetric estimator, fp32 in the scalar kernels' order
 * (space_turbo_quant.h:659-686 for b>=5, :951-983 for b<=4). */
float tqo_distance_prepared(const tqo_space *s, const float *q_rot, const float *s_q,
                            float q_norm_sq, float q_norm, const uint8_t *code);
float tqo_distance(const tqo_space *s, const float *q, const uint8_t *code);
void tqo_distance_1_to_n(const tqo_space *s, const float *q, const uint8_t *codes,
                         size_t n, float *out);
void tqo_distance_m_to_n(const tqo_space *s, const float *Q, size_t m,
                         const uint8_t *codes, size_t n, float *out);
/* Symmetric MSE-only estimator (space_turbo_quant.h:688-714, :985-1019). */
float tqo_distance_symmetric(const tqo_space *s, const uint8_t *a, const uint8_t *b);
void tqo_distance_m_to_n_symmetric(const tqo_space *s, const uint8_t *A, size_t m,
                                   const uint8_t *B, size_t n, float *out);
/* Symmetric with QJL correction (space_turbo_quant.h:731-800, :876-924). */
float tqo_distance_symmetric_full(const tqo_space *s, const uint8_t *a, const uint8_t *b);
float tqo_distance_symmetric_light(const tqo_space *s, const uint8_t *a, const uint8_t *b);
float tqo_distance_symmetric_full_prepared(const tqo_space *s, const uint8_t *a,
                                           const uint8_t *b);

/* The same estimators evaluated in double precision from the same codes
 * (query preparation in double as well). */
void tqo_distance_m_to_n_f64(const tqo_space *s, const float *Q, size_t m,
                             const uint8_t *codes, size_t n, double *out);
void tqo_distance_m_to_n_symmetric_f64(const tqo_space *s, const uint8_t *A, size_t m,
                                       const uint8_t *B, size_t n, double *out);
/* FP64 symmetric full (light = 0) / light (light = 1) estimators, [m, n]. */
void tqo_distance_m_to_n_symmetric_full_f64(const tqo_space *s, const uint8_t *A, size_t m,
                                            const uint8_t *B, size_t n, int light,
                                            double *out);

#ifdef __cplusplus
}
#endif
#endif
