/* tq_hip.h -- C ABI of the MI355X (gfx950) TurboQuant encode + distance path.
 *
 * The reference has no FFI: its only boundary is the pybind11 module
 * python/turboquant/bindings.cpp over class TurboQuantSpace
 * (include/turboquant/space_turbo_quant.h:207-648). Each entry point below
 * replaces one method of that class; the citation names the method it
 * stands in for. The Python drop-in (module turboquant._turboquant) and the
 * benchmarks call only these functions.
 *
 * Conventions
 *   - Every function returns TQ_OK (0) or a negative TQ_E* status; the
 *     message of the last failure on the calling thread is tq_last_error().
 *   - `mem` says where the caller's buffers live: TQ_MEM_HOST (pageable host
 *     memory; the call stages through device memory and returns when the
 *     results are back on the host) or TQ_MEM_DEVICE (device pointers; the
 *     call only enqueues work on `stream` and returns).
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *   - Layouts are the reference's: vectors row-major float32 [n, dim];
 *     codes row-major uint8 [n, code_size] with the nibble/byte layout and the
 *     12-byte float32 [norm, gamma, sigma] trailer of
 *     include/turboquant/turbo_quant.h:101-185; distance matrices row-major
 *     float32 [m, n].
 *   - A space handle is immutable after creation and may be shared by
 *     threads and streams.
 */
#ifndef TQ_HIP_H
#define TQ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TQ_OK 0
#define TQ_EINVAL -1      /* bad argument (dim, bits, sizes, null pointer) */
#define TQ_EHIP -2        /* HIP runtime error (message has the HIP text) */
#define TQ_ENOMEM -3      /* device allocation failed */
#define TQ_EUNSUPPORTED -4

#define TQ_MEM_HOST 0
#define TQ_MEM_DEVICE 1

/* distance_m_to_n arithmetic: plain fp32 VALU, or fp16 hi+lo MFMA with an
 * fp32 epilogue (error <= 1e-4 of the fp64 estimator). */
#define TQ_MODE_FP32 0
#define TQ_MODE_MFMA 1

typedef struct tq_space_s *tq_space;

typedef struct tq_space_info {
  uint64_t dim;          /* input dim, TurboQuantSpace::dim()            :316 */
  uint64_t padded_dim;   /* TurboQuantSpace::paddedDim()                 :317 */
  uint64_t code_size;    /* TurboQuantSpace::codeSizeBytes()             :320 */
  uint64_t packed_bytes; /* TurboQuantSpace::packedBytes()               :321 */
  int32_t bits_per_coord;
  int32_t num_levels;    /* 2^(bits-1) Lloyd-Max centroids               :324 */
  float scale;           /* sqrt(pi/2)/sqrt(padded_dim)                  :325 */
  int32_t padded;        /* dim != padded_dim                            :318 */
  uint64_t prepared_stride; /* floats per prepared query (tq_prepare_queries) */
} tq_space_info;

/* Thread-local text of the last error ("" if none). */
const char *tq_last_error(void);

/* Number of visible GPUs (0 if none); never fails. */
int tq_device_count(void);

/* TurboQuantSpace(dim, bits_per_coord, rot_seed, qjl_seed)
 * -- space_turbo_quant.h:250-311; bindings.cpp:60-65.
 * Builds the Lloyd-Max table and both sign vectors on the host and uploads
 * them to the current device. bits_per_coord must be in [2, 9]. */
int tq_space_create(uint64_t dim, int bits_per_coord, uint64_t rot_seed,
                    uint64_t qjl_seed, tq_space *out);
int tq_space_destroy(tq_space s);
int tq_space_info_get(tq_space s, tq_space_info *out);
/* Host copies of the tables the space uses (sizes: num_levels,
 * num_levels-1, padded_dim, padded_dim); any pointer may be NULL. */
int tq_space_tables(tq_space s, float *centroids, float *boundaries,
                    float *rot_signs, float *qjl_signs);

/* encodeBatch -- space_turbo_quant.h:400-407 (encodeVector :339-396);
 * bindings.cpp:93-116 (encode_batch / encode_batch_into), :75-90 (encode).
 * X: [n, dim] float32; codes: [n, code_size] uint8. Codes are byte-identical
 * to the reference's. */
int tq_encode_batch(tq_space s, const float *X, uint64_t n, uint8_t *codes,
                    int mem, void *stream);

/* prepareQuery -- space_turbo_quant.h:439-460 (TurboQuantPreparedQuery
 * :138-144). Writes m prepared queries, each prepared_stride floats:
 * [q_rot (padded_dim) | s_q (padded_dim) | q_norm_sq | q_norm | 0 | 0]. */
int tq_prepare_queries(tq_space s, const float *Q, uint64_t m, float *prepared,
                       int mem, void *stream);

/* distanceBatch1ToN -- space_turbo_quant.h:577-585; bindings.cpp:156-172. */
int tq_distance_1_to_n(tq_space s, const float *query, const uint8_t *codes,
                       uint64_t n, float *out, int mem, void *stream);

/* distanceBatchMToN -- space_turbo_quant.h:588-599; bindings.cpp:175-194.
 * mode: TQ_MODE_FP32 or TQ_MODE_MFMA. out: [m, n]. */
int tq_distance_m_to_n(tq_space s, const float *Q, uint64_t m,
                       const uint8_t *codes, uint64_t n, float *out, int mode,
                       int mem, void *stream);

/* The scan of distanceBatchMToN without the per-query preparation (the loop
 * body of space_turbo_quant.h:592-598): prepared queries from
 * tq_prepare_queries, device memory only. */
int tq_distance_prepared(tq_space s, const float *prepared, uint64_t m,
                         const uint8_t *codes, uint64_t n, float *out,
                         int mode, void *stream);

/* distanceBatchMToNSymmetric -- space_turbo_quant.h:602-615;
 * bindings.cpp:197-218 (and distance_symmetric :145-153 with m = n = 1).
 * MSE-only estimator (the QJL bits and gamma are not used). */
int tq_distance_m_to_n_symmetric(tq_space s, const uint8_t *codes_a,
                                 uint64_t m, const uint8_t *codes_b,
                                 uint64_t n, float *out, int mem,
                                 void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TQ_HIP_H */
