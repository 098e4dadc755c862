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
  uint64_t sym_prepared_stride; /* floats per prepared symmetric code
                                   (tq_prepare_symmetric) */
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

/* No counterpart in the reference: tq_distance_m_to_n for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Device buffers only; the call
 * enqueues the same kernels on `stream` and returns without waiting. Its
 * scratch (prepared queries, then the MFMA query planes or the fp32 scan's
 * QJL digit words) is one fixed area of TQ_NOALLOC_SCRATCH_BYTES per device
 * inside the library, held until the process ends: calls needing more
 * (about 16 KB per query at padded dim 1024, i.e. up to ~4,000 queries) fail
 * with TQ_ENOMEM. Every call on a device shares that area, so calls that
 * overlap on one device (different streams, other threads) race: order them,
 * e.g. on one stream. No allocation, host synchronisation or library
 * allocation of any kind happens in the call. */
#define TQ_NOALLOC_SCRATCH_BYTES (64u << 20)
int tq_distance_m_to_n_noalloc(tq_space s, const float *Q, uint64_t m,
                               const uint8_t *codes, uint64_t n, float *out,
                               int mode, void *stream);

/* The scan of distanceBatchMToN without the per-query preparation (the loop
 * body of space_turbo_quant.h:592-598): prepared queries from
 * tq_prepare_queries, device memory only. */
int tq_distance_prepared(tq_space s, const float *prepared, uint64_t m,
                         const uint8_t *codes, uint64_t n, float *out,
                         int mode, void *stream);

/* distanceBatchMToNSymmetric -- space_turbo_quant.h:602-615;
 * bindings.cpp:197-218 (and distance_symmetric :145-153 with m = n = 1).
 * MSE-only estimator (the QJL bits and gamma are not used); fp32. */
int tq_distance_m_to_n_symmetric(tq_space s, const uint8_t *codes_a,
                                 uint64_t m, const uint8_t *codes_b,
                                 uint64_t n, float *out, int mem,
                                 void *stream);

/* Symmetric estimator variants. */
#define TQ_SYM_MSE 0   /* distBuildScalar(B4)           space_turbo_quant.h:688-714, :985-1019 */
#define TQ_SYM_FULL 1  /* distBuildFullScalar           :731-799 */
#define TQ_SYM_LIGHT 2 /* distBuildLightScalar          :902-945 */

/* distanceBatchMToNSymmetric / ...SymmetricFull / ...SymmetricLight
 * (:602-647; bindings.cpp:197-266). mode TQ_MODE_MFMA is available for
 * TQ_SYM_MSE (the decoded a codes as fp16 hi+lo queries, error <= 1e-4 of
 * the fp64 estimator); the QJL-corrected variants run in fp32. */
int tq_distance_m_to_n_symmetric_ex(tq_space s, const uint8_t *codes_a,
                                    uint64_t m, const uint8_t *codes_b,
                                    uint64_t n, float *out, int variant,
                                    int mode, int mem, void *stream);

/* prepareSymmetricQuery -- space_turbo_quant.h:499-532 (bindings.cpp:276-282).
 * m rows of sym_prepared_stride floats: the a side of the full estimator
 * (reconstruction, transposed-Hadamard signs, Hadamard of the reconstruction
 * with the sign term folded in, norm^2, norm). */
int tq_prepare_symmetric(tq_space s, const uint8_t *codes, uint64_t m,
                         float *prepared, int mem, void *stream);

/* distanceSymmetricFullPrepared / distanceBatch1ToNSymmetricFull /
 * distanceBatchMToNSymmetricFullPrepared -- :537-572 (bindings.cpp:284-334):
 * full symmetric distances of m prepared codes against n codes, [m, n]. */
int tq_distance_symmetric_prepared(tq_space s, const float *prepared,
                                   uint64_t m, const uint8_t *codes,
                                   uint64_t n, float *out, int mem,
                                   void *stream);

/* rotatedCoordsBatch -- space_turbo_quant.h:411-435 (bindings.cpp:119-132):
 * [n, padded_dim] float32 sigma-normalised rotated coordinates, bit-identical
 * to the reference. */
int tq_rotated_coords_batch(tq_space s, const float *X, uint64_t n, float *out,
                            int mem, void *stream);

/* Encode graphs -- no counterpart in the reference (encodeBatch,
 * space_turbo_quant.h:400-407, is an OpenMP loop on the host). The encode of
 * a fixed batch of n rows at X (device memory, n x dim float32) into codes
 * (device memory, n x code_size), captured once into a HIP graph; each
 * tq_encode_graph_run replays it on `stream` with one graph launch and reads
 * X when it runs. For ingest loops and serving appends that refill the same
 * staging buffer. X and codes must stay allocated for the graph's life; host
 * memory is rejected (TQ_EINVAL). */
typedef struct tq_encode_graph_s *tq_encode_graph;
int tq_encode_graph_create(tq_space s, const float *X, uint64_t n, uint8_t *codes,
                           tq_encode_graph *out);
int tq_encode_graph_run(tq_encode_graph g, void *stream);
int tq_encode_graph_destroy(tq_encode_graph g);

/* ---- top-k search ---------------------------------------------------------
 * No reference counterpart: the reference returns full distance matrices and
 * its users rank rows with numpy (argpartition + argsort,
 * python/benchmarks/run_benchmark.py:251-255; tests/test_turboquant.py:
 * 668-674). Here a search returns, per query, the k codes of smallest
 * distance_m_to_n distance, ascending, ties broken by the smaller code index.
 *
 * A job is one query batch against one collection; jobs of one call may use
 * different spaces (dims, bit widths) on the same device. All job buffers are
 * device memory. k is at most TQ_SEARCH_MAX_K (the reference's recall
 * harness asks for up to 100, python/benchmarks/run_benchmark.py:329-337);
 * the fused form keeps its lists in registers and takes k <= TQ_FUSED_MAX_K,
 * so larger k always runs blocked. */
#define TQ_SEARCH_MAX_K 128
#define TQ_FUSED_MAX_K 16

typedef struct tq_search_job {
  tq_space space;
  const float *queries; /* [m, dim] float32 */
  uint64_t m;
  const uint8_t *codes; /* [n, code_size], n < 2^32 - 1 */
  uint64_t n;
  float *out_dist;      /* [m, k] float32, ascending */
  int64_t *out_ids;     /* [m, k] row index into codes; -1 (dist +inf) past n */
} tq_search_job;

/* Forms of an MFMA search over one collection. */
#define TQ_FORM_AUTO 0
#define TQ_FORM_FUSED 1   /* the persistent launch with the top-k fused into the scan */
#define TQ_FORM_BLOCKED 2 /* MFMA distance blocks over chunks of codes, row top-k, fold */

/* Knobs of TQ_MODE_MFMA plans. 0 = automatic: the autotuner's table for the
 * job's (padded dim, bits, m) bucket if it has an entry
 * (tq_search_plan_tune), else the built-in rule (blocked; fused where a batch
 * has more than 65535 queries or no codes). query_tile, items_per_wg and
 * wg_per_cu shape the fused form's work queue and persistent grid;
 * blocked_query_tile and block_mb the blocked form's MFMA tiles and distance
 * blocks. No knob changes a result: every distance is a function of its
 * (query, code) pair and ties go by id. */
typedef struct tq_search_tuning {
  int32_t query_tile;   /* fused form: 32 or 64 queries per work item, or 0 */
  int32_t items_per_wg; /* fused form: work items per persistent workgroup (load balance), or 0 */
  int32_t form;         /* TQ_FORM_*, or 0 */
  int32_t wg_per_cu;    /* fused form: persistent workgroups per compute unit, 1 or 2, or 0 */
  int32_t blocked_query_tile; /* blocked form: MFMA query tile, 32 or 64, or 0 (32 up to
                                 32 queries, else 64) */
  int32_t block_mb;     /* blocked form: distance block budget in MiB, 16..256, or 0 (256) */
} tq_search_tuning;

typedef struct tq_search_plan_s *tq_search_plan;

/* Builds a reusable search over `jobs` (copied): checks every job, sizes the
 * work queue and allocates all device scratch once. Jobs aimed at one
 * collection are scored as one batch. mode TQ_MODE_MFMA, per batch in the
 * form the knobs give (k <= TQ_SEARCH_MAX_K): blocked (MFMA distances of
 * the batch against a chunk of codes into a scratch block of at most 256 MB
 * or 4,096 columns, a row top-k per chunk, candidates folded per query), or
 * fused (every fused batch of the plan in one persistent launch over a
 * (query tile, code range) queue, the top-k lists kept in LDS); both give
 * the top-k of the batch's full MFMA distance matrix (tq_distance_m_to_n),
 * bit for bit. k > TQ_FUSED_MAX_K always runs blocked (a fused
 * `tuning` is then refused, as are MFMA plans of more than 65535 queries).
 * TQ_MODE_FP32: per job the plain fp32 distance matrix in row blocks, then a
 * row top-k. `tuning` may be NULL. */
int tq_search_plan_create(const tq_search_job *jobs, uint64_t njobs, int k, int mode,
                          const tq_search_tuning *tuning, tq_search_plan *out);
/* Enqueues the plan on `stream`: query preparation, the scan and the top-k.
 * No allocation and no host synchronisation, so it may be captured into a
 * HIP graph; the job buffers are read when the work runs. */
int tq_search_plan_run(tq_search_plan p, void *stream);
/* As tq_search_plan_run, replaying a HIP graph of the plan's work (captured
 * on the first call). */
int tq_search_plan_run_graph(tq_search_plan p, void *stream);
/* Times the candidate knobs for each (padded dim, bits, m) bucket of the
 * plan's jobs on the plan's own buffers (overwriting the outputs): the
 * fused form with query_tile in {32, 64} x items_per_wg in {1, 2, 4, 8} x
 * wg_per_cu in {1, 2}, and the blocked form with blocked_query_tile in
 * {32, 64} x block_mb in {64, 128, 256}. Records the fastest form and the
 * fastest knobs of each form in the process-wide table, and rebuilds the
 * plan with them. Synchronous. `chosen` (may be NULL) receives the plan's
 * resulting knobs (those of its costliest batch). */
int tq_search_plan_tune(tq_search_plan p, void *stream, tq_search_tuning *chosen);
/* The autotuner's table entry for (padded dim, bits, m bucket of m), if any:
 * TQ_OK and *out filled, or TQ_EINVAL when the bucket has not been tuned.
 * No counterpart in the reference (it has no search or tuner). */
int tq_search_tuning_lookup(int32_t padded_dim, int32_t bits, uint64_t m, tq_search_tuning *out);
/* Work items and persistent workgroups of the plan (either may be NULL). */
int tq_search_plan_stats(tq_search_plan p, uint64_t *items, uint64_t *workgroups);
int tq_search_plan_destroy(tq_search_plan p);

/* One-shot: create, run on `stream`, wait, destroy. */
int tq_search(const tq_search_job *jobs, uint64_t njobs, int k, int mode, void *stream);

/* Row-wise top-k of a device float32 [m, n] matrix (ties by column index):
 * out_dist / out_ids [m, k] on the device; m <= 65535, k <= TQ_SEARCH_MAX_K. */
int tq_topk_rows(const float *dist, uint64_t m, uint64_t n, int k, float *out_dist,
                 int64_t *out_ids, void *stream);

/* ---- serving ----------------------------------------------------------------
 * No reference counterpart (the reference is a library of synchronous calls).
 * A server keeps code collections resident in device memory and runs search
 * and append requests submitted from any thread. Its scheduler batches
 * continuously: requests that arrive while earlier batches are on the device
 * are admitted as soon as a batch slot and its workspace allow, and all
 * searches of a batch run as one search plan (tq_search_plan_create: one
 * batch per collection). Requests take effect in submission order: a search sees the
 * codes of exactly the appends submitted before it. */
typedef struct tq_server_s *tq_server;

typedef struct tq_server_config {
  int32_t mode;               /* TQ_MODE_MFMA or TQ_MODE_FP32 */
  int32_t k;                  /* top-k of every search, 1..TQ_SEARCH_MAX_K */
  uint64_t hbm_budget;        /* device bytes: collections + batch slots */
  uint64_t slot_bytes;        /* device workspace per batch slot (0: 1 GiB) */
  uint32_t max_inflight;      /* batch slots, i.e. batches on the device at once (0: 2) */
  uint32_t max_batch_queries; /* queries per batch (0: 4096) */
} tq_server_config;

typedef struct tq_server_stats {
  uint64_t batches;
  uint64_t searches_done;
  uint64_t appends_done;
  uint64_t queries_done;
  uint64_t failed;
  uint64_t resident_bytes;  /* collection capacity held in device memory */
  uint64_t reserved_bytes;  /* batch slot workspaces */
  uint64_t max_batch_requests;
} tq_server_stats;

/* Starts the scheduler; batch slots are charged to hbm_budget up front. */
int tq_server_create(const tq_server_config *cfg, tq_server *out);
/* Completes every submitted request, then stops. */
int tq_server_destroy(tq_server s);
/* An empty resident collection of up to `capacity` codes of `space`
 * (capacity * code_size device bytes from the budget). */
int tq_server_add_collection(tq_server s, tq_space space, uint64_t capacity,
                             uint32_t *coll);
/* Encodes n vectors (host float32 [n, dim]) and appends their codes. X must
 * stay valid until the ticket completes. Fails at submission (TQ_ENOMEM)
 * when the collection would exceed its capacity. */
int tq_server_submit_append(tq_server s, uint32_t coll, const float *X, uint64_t n,
                            uint64_t *ticket);
/* Top-k search of m queries (host float32 [m, dim]); results are written to
 * host out_dist / out_ids [m, k] before the ticket completes. Q must stay
 * valid until then. */
int tq_server_submit_search(tq_server s, uint32_t coll, const float *Q, uint64_t m,
                            float *out_dist, int64_t *out_ids, uint64_t *ticket);
/* Blocks until the ticket completes and releases it. Returns the request's
 * status; latency_s (may be NULL) receives submission-to-completion time. */
int tq_server_wait(tq_server s, uint64_t ticket, double *latency_s);
/* Codes in the collection counting every submitted append, and the device
 * address of its code rows. */
int tq_server_collection(tq_server s, uint32_t coll, uint64_t *n, const uint8_t **codes);
/* Appends n existing codes (the reference's byte layout; host or device
 * memory per `mem`) to the collection, e.g. to load a persisted collection.
 * Synchronous: searches submitted after it returns see the rows; none may be
 * submitted on this collection while it runs. */
int tq_server_load_codes(tq_server s, uint32_t coll, const uint8_t *codes, uint64_t n,
                         int mem);
/* Copies code rows [start, start + count) to host memory (the reference's
 * byte layout, e.g. to persist the collection). The rows must come from
 * appends whose tickets have completed. */
int tq_server_read_codes(tq_server s, uint32_t coll, uint64_t start, uint64_t count,
                         uint8_t *host);
int tq_server_stats_get(tq_server s, tq_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif /* TQ_HIP_H */
