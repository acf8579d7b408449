/*
 * svdb_oracle.h — CPU oracle for the similarity / nearest-vector path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (libsvdb.so) never links or calls it.
 *
 * Contents (each function cites the reference code it restates):
 *   - orc_kd_*      restatement of the reference kd-tree (kdtree.c:15-62,131-178)
 *   - orc_ref_*     restatement of the reference float-accumulating metrics
 *                   (vector_database.c:301-352)
 *   - orc_topk*     the canonical FP64 exhaustive top-k (the "FP64 oracle"):
 *                   sequential fused-multiply-add chains, ties by lowest id,
 *                   NaN ranked last
 *   - orc_pair_fp64 the canonical FP64 pair metric: 64 strided partial sums
 *                   then an xor butterfly (the order a 64-lane wavefront
 *                   reduction produces)
 */
#ifndef SVDB_ORACLE_H
#define SVDB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Metric ids; identical to SVDB_METRIC_* in include/svdb.h. */
enum { ORC_L2 = 0, ORC_COS = 1, ORC_DOT = 2 };

/* ---- reference kd-tree restatement (kdtree.c) ------------------------- */
typedef struct orc_kd orc_kd;
orc_kd* orc_kd_create(size_t dimension);
void    orc_kd_insert(orc_kd* t, const double* point, size_t index);
size_t  orc_kd_nearest(const orc_kd* t, const double* point);
size_t  orc_kd_size(const orc_kd* t);
size_t  orc_kd_depth(const orc_kd* t);
void    orc_kd_free(orc_kd* t);
/* Build a tree over rows [0, n) of X (row stride ldx) in row order, the way
 * vector_db_load rebuilds it (vector_database.c:277-279). */
orc_kd* orc_kd_build(const double* X, size_t n, size_t ldx, size_t dimension);
/* Answer nq queries with nthreads threads (each an independent stream). */
void    orc_kd_nearest_batch(const orc_kd* t, const double* Q, size_t nq, size_t ldq,
                             int nthreads, uint64_t* out);

/* ---- reference float metrics (vector_database.c:301-352) ------------- */
float orc_ref_cos(const double* a, const double* b, size_t n);
float orc_ref_l2(const double* a, const double* b, size_t n);
float orc_ref_dot(const double* a, const double* b, size_t n);
/* metric per call; out[p] for pair (A + ia[p]*lda, A + ib[p]*lda). */
void  orc_ref_pairs(const double* A, size_t lda, size_t n, const uint64_t* ia,
                    const uint64_t* ib, size_t np, int metric, float* out);

/* ---- canonical FP64 keys --------------------------------------------- */
double orc_l2_key(const double* q, const double* x, size_t D);   /* Σ fma(t,t,·), t=x_i-q_i */
double orc_dot_key(const double* q, const double* x, size_t D);  /* Σ fma(q_i,x_i,·)        */
double orc_sumsq(const double* x, size_t D);                     /* Σ fma(x_i,x_i,·)        */
double orc_cos_score(double dot, double nq, double nx);          /* dot/(sqrt(nq)*sqrt(nx)) */

/* ---- canonical FP64 exhaustive top-k --------------------------------- *
 * X: N rows, row stride ldx, first D columns used.  alive: optional byte
 * per row (0 = deleted; deleted rows are skipped and the returned ids are
 * logical ids = number of alive rows before the row).  Q: B queries,
 * stride ldq.  Ranking: L2 ascending on the squared distance, COS/DOT
 * descending on the score; equal keys by lowest id; NaN after numbers.
 * out_ids[b*k + j] / out_scores[b*k + j]: j-th best; missing → id
 * UINT64_MAX, score NaN.  Scores: L2 → sqrt(key), COS → cosine, DOT → dot.
 * Returns 0, or -1 on bad arguments. */
int orc_topk(const double* X, size_t N, size_t ldx, const uint8_t* alive,
             const double* Q, size_t B, size_t ldq, size_t D, int metric, size_t k,
             int nthreads, uint64_t* out_ids, double* out_scores);
/* Same contract, plain scalar loop (no SIMD, one thread): cross-check of
 * orc_topk on small inputs. */
int orc_topk_scalar(const double* X, size_t N, size_t ldx, const uint8_t* alive,
                    const double* Q, size_t B, size_t ldq, size_t D, int metric,
                    size_t k, uint64_t* out_ids, double* out_scores);

/* ---- canonical FP64 pair metric -------------------------------------- */
double orc_pair_fp64(const double* a, const double* b, size_t D, int metric);
void   orc_pairs_fp64(const double* A, size_t lda, size_t D, const uint64_t* ia,
                      const uint64_t* ib, size_t np, int metric, double* out);

/* ---- the reference's own CPU path, timed (refkd_bench.c) -------------- *
 * orc_reftree_build lays out, in the reference's KDTreeNode / KDTree
 * structures, exactly the tree kdtree_insert grows from rows 0..n-1 in
 * order (kdtree.c:47-62), with nthreads threads; orc_reftree_handle gives
 * the KDTree* to pass to the reference's kdtree_nearest.  orc_time_kd /
 * orc_time_pairs call a reference function through its address on
 * nthreads independent streams and return wall seconds. */
typedef struct orc_reftree orc_reftree;
orc_reftree* orc_reftree_build(const double* X, size_t n, size_t ldx, size_t D, int nthreads);
void* orc_reftree_handle(orc_reftree* t);
size_t orc_reftree_at(const void* tree, const unsigned char* path, size_t len);
void orc_reftree_shape(const void* tree, size_t* count, size_t* height);
void orc_reftree_free(orc_reftree* t);
double orc_time_kd(void* fn, void* tree, const double* Q, size_t nq, size_t ldq, int nthreads,
                   uint64_t* ids, double* ns);
double orc_time_pairs(void* fn, const double* X, size_t ldx, size_t d, const uint64_t* ia,
                      const uint64_t* ib, size_t np, int nthreads, float* out);

/* ---- helpers used by tests -------------------------------------------- */
/* Map a score to its ascending rank key exactly as the ranking uses it. */
double orc_rank_key(double score_or_l2key, int metric);

#ifdef __cplusplus
}
#endif
#endif
