/*
 * svdb.h — batch entry points of the MI355X engine behind simple-vector-db.
 *
 * These have no counterpart in the reference; they are the batched forms of
 * its per-request calls, for callers that can hand over many queries or
 * pairs at once:
 *   svdb_search        batched top-k find-nearest with a metric per query —
 *                      the reference answers one query, k = 1, Euclidean
 *                      only (kdtree_nearest, kdtree.c:171-178, called from
 *                      compare_handler.c:403)
 *   svdb_pair_metrics  batched FP64 cosine / Euclidean / dot over stored rows
 *                      (the reference computes one float pair per request,
 *                      vector_database.c:301-352 via compare_handler.c:152-170)
 *   svdb_bulk_insert   bulk load (the reference inserts row by row,
 *                      vector_database.c:81-119, 277-279)
 *
 * Ranking (the "FP64 oracle" contract): Euclidean ascending on the squared
 * distance Σ fma((x_i - q_i),(x_i - q_i),·) accumulated in index order;
 * dot descending on Σ fma(q_i, x_i, ·); cosine descending on
 * dot / (sqrt(‖q‖²) * sqrt(‖x‖²)) with the norms accumulated the same way.
 * Equal keys go to the lowest index; NaN ranks after every number.  Returned
 * scores: Euclidean distance (sqrt of the key), dot product, cosine.  Every
 * precision mode returns exactly these ids and FP64 scores; the low-precision
 * modes only change how candidates are found (a certified BF16/FP8 MFMA
 * first pass, then an FP64 rerank; uncertified queries are rescanned in
 * FP64).
 *
 * All functions return 0 on success or a negative SVDB_E* code;
 * svdb_last_error() describes the last failure on the calling thread.
 */
#ifndef SVDB_H
#define SVDB_H

#include <stddef.h>
#include <stdint.h>

#include "vector_database.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { SVDB_METRIC_L2 = 0, SVDB_METRIC_COS = 1, SVDB_METRIC_DOT = 2 };

enum {
    SVDB_PREC_FP64 = 0, /* plain FP64 exhaustive scan */
    SVDB_PREC_BF16 = 1, /* BF16 MFMA first pass + shortlist + FP64 rerank */
    SVDB_PREC_FP8 = 2,  /* FP8-e4m3 MFMA first pass + shortlist + FP64 rerank
                           (needs svdb_enable_shadow(db, SVDB_PREC_FP8); d <= 1536) */
    SVDB_PREC_AUTO = 3  /* BF16 when the store has a shadow, else FP64 */
};

enum {
    SVDB_OK = 0,
    SVDB_EINVAL = -1,  /* bad argument */
    SVDB_ENOMEM = -2,  /* device allocation failed or HBM budget exceeded */
    SVDB_EDEVICE = -3, /* HIP runtime error / no usable GPU */
    SVDB_EEMPTY = -4,  /* store has no rows */
    SVDB_EBUSY = -5    /* server queue full */
};

/* Largest k accepted by the search entry points. */
#define SVDB_MAX_K 128

/* Top-k for nq host queries (row i at queries + i*dim).  k[i] in
 * [1, SVDB_MAX_K] and metric[i] per query; out_stride >= max k[i].
 * ids_out[i*out_stride + j], scores_out[...]: j-th best; slots past the
 * number of stored rows get (size_t)-1 and NaN.  dim is the store's vector
 * dimension, or, for Euclidean only, less: the first dim coordinates are
 * compared, as the reference's kd-tree does (kdtree.c:134-137). */
int svdb_search(VectorDatabase* db, const double* queries, size_t nq, size_t dim,
                const size_t* k, const int* metric, int precision, size_t out_stride,
                size_t* ids_out, double* scores_out);

/* Same, with queries, ids and scores already in device memory (HBM):
 * d_queries[i*ldq + c], d_ids[i*k + j], d_scores[i*k + j]; one k and metric
 * for the batch.  Work is enqueued on the engine's stream and the call
 * returns when it has completed. */
int svdb_search_device(VectorDatabase* db, const double* d_queries, size_t nq, size_t ldq,
                       size_t k, int metric, int precision, uint64_t* d_ids, double* d_scores);

/* The FP64 form of svdb_search_device (the exact scan) that allocates
 * nothing, host or device: every intermediate (k per query, query norms, the
 * row slices' partial top-k lists) lives in a 64 MiB device array that the
 * library holds on each GPU for the life of the process.  No counterpart in
 * the reference: it is for callers that cannot allocate — code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Limits: nq <= 8192 and nq * k small enough for the array (the row slices
 * shrink to fit; at 8 slices nq * k <= about 170,000); larger requests get
 * SVDB_EINVAL ("... allocation-free scratch") and nothing runs.  Calls from
 * any thread are serialised (they share the array); ids and scores equal
 * svdb_search_device(..., SVDB_PREC_FP64, ...) bit for bit. */
int svdb_search_device_noalloc(VectorDatabase* db, const double* d_queries, size_t nq,
                               size_t ldq, size_t k, int metric, uint64_t* d_ids,
                               double* d_scores);

/* FP64 metric for npairs pairs of stored rows (logical indices), one 64-lane
 * wavefront per pair: lane l sums elements i ≡ l (mod 64) with fma, then an
 * xor butterfly.  Euclidean returns the distance. */
int svdb_pair_metrics(VectorDatabase* db, const size_t* index_a, const size_t* index_b,
                      size_t npairs, int metric, double* out);

/* Append n rows (row i at rows + i*dim) in one device pass.  uuids: n
 * NUL-padded 37-byte records, or NULL for empty uuids.  The host mirror gets
 * malloc'ed copies.  *first_index receives the first new row's index. */
int svdb_bulk_insert(VectorDatabase* db, const double* rows, size_t n, size_t dim,
                     const char* uuids, size_t* first_index);

/* Engine knobs. */
int svdb_set_shortlist(VectorDatabase* db, size_t width); /* first-pass shortlist L (>= k) */
int svdb_set_budget_bytes(VectorDatabase* db, size_t bytes); /* device-memory budget */
int svdb_enable_shadow(VectorDatabase* db, int precision); /* build the BF16 or FP8 shadow
                                                               (d <= 1536) */

typedef struct svdb_stats {
    size_t rows;            /* logical rows */
    size_t dim;             /* vector dimension (0 before the first insert) */
    size_t phys_rows;       /* device slots in use (rows + tombstones) */
    size_t device_bytes;    /* device memory held by the engine */
    size_t budget_bytes;    /* device-memory budget */
    uint64_t searches;      /* search calls */
    uint64_t queries;       /* queries answered */
    uint64_t certified;     /* low-precision queries certified exact */
    uint64_t fallbacks;     /* low-precision queries rescanned in FP64 */
    double last_max_rel_err;/* max |first-pass score - FP64 score| / |FP64| of the last call */
    uint64_t widened;       /* of the certified: certified only by a wider retry (1024, 4096) */
    uint64_t second_pass;   /* FP8 queries handed to the BF16 first pass (counted again there) */
    uint64_t compactions;   /* tombstone compactions (live rows moved to the front, shadows
                               rebuilt) */
    uint64_t graph_launches;/* exact searches run as one launch of a captured HIP graph */
    uint64_t graph_captures;/* HIP graphs captured (a launch sequence seen a second time) */
} svdb_stats;
int svdb_get_stats(VectorDatabase* db, svdb_stats* out);

/* Kernel timing, measured with HIP events on the engine's stream. */
typedef struct svdb_kernel_time {
    uint64_t launches;
    double total_ms;
} svdb_kernel_time;
enum { SVDB_K_SCAN_FP64 = 0, SVDB_K_SCAN_LOWP = 1, SVDB_K_SHORTLIST = 2, SVDB_K_RERANK = 3,
       SVDB_K_MERGE = 4, SVDB_K_PREPASS = 5 /* sample scan + k-th key */,
       SVDB_K_SECOND_PASS = 6 /* FP8 queries rerun through the BF16 pass, whole */,
       SVDB_K_COUNT = 7 };
int svdb_set_profiling(VectorDatabase* db, int on); /* also resets the counters */
int svdb_get_kernel_times(VectorDatabase* db, svdb_kernel_time* out /* [SVDB_K_COUNT] */);

/* Measurement helper (no counterpart in the reference): the sustained
 * matrix-core throughput of one data type on this device — a loop of
 * independent MFMAs with no memory traffic at 2 and 4 wavefronts per SIMD,
 * the best of three launches each — in TFLOP/s, the practical ceiling a
 * scan's roofline fraction can be read against beside the published peak. */
enum { SVDB_MFMA_FP64 = 0, SVDB_MFMA_BF16 = 1, SVDB_MFMA_FP8 = 2 };
int svdb_mfma_sustained(int kind, double* tflops);

/* ---- serving loop -------------------------------------------------------
 * A continuous-batching server over one store.  Searches and mutations are
 * submitted from any thread and answered in submission order by one device
 * thread: each tick takes every queued search (any sizes, k and metrics, up
 * to max_batch queries and while the tick's estimated device memory stays
 * within budget_bytes) into one batch; mutations are applied between ticks
 * in order (a search sees exactly the mutations submitted before it).
 * Requests arriving during a tick join the next one, which is staged on
 * the host while the current tick runs and takes searches until the device
 * thread starts it.  A smaller max_batch gives shorter ticks (lower tail
 * latency) at some cost in saturated throughput.  Results equal those of svdb_search
 * on the same store state.  Stop the server only after every submitted
 * request has been waited for. */
typedef struct svdb_server svdb_server;
typedef struct svdb_server_config {
    size_t max_batch;    /* queries per tick (default 8192) */
    size_t budget_bytes; /* device memory for one tick's requests (default 64 GiB) */
    int precision;       /* SVDB_PREC_* (default AUTO) */
} svdb_server_config;
typedef struct svdb_request_info {
    int64_t submit_us, start_us, done_us; /* CLOCK_MONOTONIC microseconds */
    uint64_t tick;                        /* tick that served the request */
    size_t tick_queries;                  /* queries in that tick */
} svdb_request_info;
void svdb_server_default_config(svdb_server_config* cfg);
svdb_server* svdb_server_start(VectorDatabase* db, const svdb_server_config* cfg);
void svdb_server_stop(svdb_server* s);
/* Each returns a ticket (>= 0) or a negative SVDB_E* code.  A search larger
 * than max_batch queries or than the memory budget is refused (SVDB_ENOMEM). */
int64_t svdb_server_submit_search(svdb_server* s, const double* queries, size_t nq, size_t dim,
                                  size_t k, int metric);
int64_t svdb_server_submit_insert(svdb_server* s, const double* row, size_t dim);
int64_t svdb_server_submit_update(svdb_server* s, size_t index, const double* row, size_t dim);
int64_t svdb_server_submit_delete(svdb_server* s, size_t index);
/* Blocks until the request is done.  Search: ids/scores get nq*k entries
 * (row stride k).  Insert: ids[0] gets the new index.  Returns the request's
 * status. */
int svdb_server_wait(svdb_server* s, int64_t ticket, size_t* ids, double* scores,
                     svdb_request_info* info);

/* Device memory for callers of svdb_search_device, on the engine's device
 * and HIP runtime (a second HIP runtime in the process, e.g. one bundled by
 * a framework, cannot share it).  kind: 0 host→device, 1 device→host,
 * 2 device→device (a plain 16-byte-per-lane copy kernel, which reads and
 * writes `bytes` exactly once).  svdb_device_sync waits for all work on the
 * device. */
void* svdb_device_alloc(size_t bytes);
void svdb_device_free(void* ptr);
int svdb_device_copy(void* dst, const void* src, size_t bytes, int kind);
int svdb_device_sync(void);

const char* svdb_last_error(void);

/* Test hook: the n-th host allocation point reached from now on (in any
 * thread) fails as an out-of-memory would (std::bad_alloc inside the
 * library, reported as SVDB_ENOMEM or the reference's sentinel); n = 0
 * disarms.  Returns the number of points still to pass before the armed
 * failure (0 once it has fired or when disarmed). */
int64_t svdb_debug_fail_host_alloc(int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* SVDB_H */
