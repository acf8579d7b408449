/*
 * svdb_nearest.h — the handler-independent part of POST /nearest.
 *
 * The reference's nearest_handler_callback (compare_handler.c:263-441)
 * parses one JSON array of db_vector_size numbers, calls
 * kdtree_nearest(db->kdtree, q) and answers {"index","vector","uuid"} of the
 * row vector_db_read returns.  Its README documents ?number= (README.md:287)
 * but the handler never reads it.  svdb_nearest_request does what the
 * handler does between "the body is complete" and "queue the response",
 * widened to:
 *   - ?number=k   (1..SVDB_MAX_K; default 1): the k nearest rows;
 *   - ?metric=m   euclidean (default), cosine or dot (also the /compare
 *                 names euclidean_distance, cosine_similarity, dot_product);
 *   - a batch body: a JSON array of query arrays.
 * Euclidean compares the first db->kdtree->dimension coordinates, as
 * kdtree_nearest does (kdtree.c:134-137); cosine and dot compare all of
 * them.  Equal scores go to the lowest index.
 *
 * Response (application/json):
 *   one query, k = 1:  {"index":i,"vector":[...],"uuid":"..."} — the
 *                      reference's object (compare_handler.c:409-416), with
 *                      the FP64 score added under the metric's /compare
 *                      name ("euclidean_distance", "cosine_similarity",
 *                      "dot_product") only when ?metric= was given;
 *   one query, k > 1:  [{"index":i,"<metric name>":s,"vector":[...],"uuid":"..."}, ...]
 *                      best first (fewer entries when the store holds fewer rows);
 *   a batch:           [answer for query 0, answer for query 1, ...], each as above;
 *   empty store:       {"error": "No nearest neighbor found"} with status 200,
 *                      as the reference answers (compare_handler.c:418-419).
 * Errors keep the reference's order, texts and statuses: an empty body →
 * 400 "Empty data"; unparsable JSON → 400 "Invalid JSON"; an array whose
 * length differs from expected_dim → 400 "Vector size mismatch"; a
 * non-number element → 400 "Invalid vector data"; a host allocation failure
 * → 500 "Memory allocation failed".  New: a bad ?number= → 400 "Invalid
 * number"; an unknown ?metric= → 400 "Unknown metric"; a device failure →
 * 500 with svdb_last_error().  Numbers are written as cJSON 1.7 prints
 * them (%1.15g, or %1.17g when that does not read back exactly).
 */
#ifndef SVDB_NEAREST_H
#define SVDB_NEAREST_H

#include <stddef.h>

#include "svdb.h"

#ifdef __cplusplus
extern "C" {
#endif

/* body: the request body (len bytes, need not be NUL-terminated).
 * number, metric: the query-string values, or NULL when absent.
 * srv: a serving loop over db (svdb_server_start), or NULL to call
 * svdb_search directly.  On return *response is a malloc'ed NUL-terminated
 * JSON text (the caller frees it; NULL only if even the error text could
 * not be allocated) of *response_len bytes.  Returns the HTTP status. */
int svdb_nearest_request(VectorDatabase* db, svdb_server* srv, size_t expected_dim,
                         const char* body, size_t len, const char* number, const char* metric,
                         int precision, char** response, size_t* response_len);

#ifdef __cplusplus
}
#endif

#endif /* SVDB_NEAREST_H */
