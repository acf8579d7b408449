/*
 * handlers_replay.c — a C99 caller of libsvdb.so that makes the reference
 * handlers' calls, in their order, through the include/ headers exactly as a
 * handler translation unit would (compiled with gcc -std=c99 -Wall
 * -Werror against include/, linked with -lsvdb).
 *
 * Sequences replayed (reference file:line):
 *   main.c:349-351          vector_db_load, else vector_db_init(0, kd_dim)
 *   post_handler.c:299-333  lazy kdtree_create when db->kdtree is NULL, then
 *                           vector_db_insert (ownership of vec.data passes)
 *   compare_handler.c:403-416  kdtree_nearest(db->kdtree, q) then
 *                           vector_db_read(db, index): index, vector, uuid
 *   compare_handler.c:101-170  db->size bounds check, vector_db_read x2,
 *                           dimension checks, cosine_similarity /
 *                           euclidean_distance / dot_product
 *   put_handler.c:247       vector_db_update
 *   delete_handler.c:84-96  db->size bounds check, vector_db_delete
 *   main.c:398-402          vector_db_save, vector_db_free
 * and the batched /nearest layer (svdb_nearest.h) a handler would call.
 *
 * Usage: handlers_replay IN OUT SAVEFILE
 *   IN  (native-endian): u64 N, d, kd_dim, nq, np; f64 rows[N][d];
 *       f64 queries[nq][d]; u64 pairs[np][2]; u64 nupd; per update u64 index,
 *       f64 row[d]; u64 ndel; u64 del_index[ndel]
 *   OUT: u64 kd_ids[nq]; u32 metric_bits[np][3]; u64 ids_after[nq];
 *        u64 size_after; u64 loaded_size; u64 json_len; char json[json_len]
 * Exit status 0 when every read-back matched, 2 on usage / IO errors, 3 when
 * the store could not be created (no GPU), 4 on a mismatch.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kdtree.h"
#include "svdb.h"
#include "svdb_nearest.h"
#include "vector_database.h"

static int rd(FILE* f, void* p, size_t sz, size_t n) { return fread(p, sz, n, f) == n; }

static double* read_rows(FILE* f, uint64_t n, uint64_t d) {
    double* p = (double*)malloc((size_t)(n * d > 0 ? n * d : 1) * sizeof(double));
    if (p && !rd(f, p, sizeof(double), (size_t)(n * d))) {
        free(p);
        return NULL;
    }
    return p;
}

/* post_handler.c:299-333: a malloc'ed row handed to vector_db_insert */
static size_t post_vector(VectorDatabase* db, const double* row, size_t d, size_t i) {
    Vector vec;
    memset(&vec, 0, sizeof vec);
    snprintf(vec.uuid, UUID_SIZE, "row-%08zu", i);
    vec.dimension = d;
    vec.data = (double*)malloc(d * sizeof(double));
    if (!vec.data) return (size_t)-1;
    memcpy(vec.data, row, d * sizeof(double));
    if (db->kdtree == NULL) {
        if (db->size > 0 && db->vectors && db->vectors[0].data)
            db->kdtree = kdtree_create(db->vectors[0].dimension);
        else
            db->kdtree = kdtree_create(d);
        if (db->kdtree == NULL) {
            free(vec.data);
            return (size_t)-1;
        }
    }
    return vector_db_insert(db, vec);
}

int main(int argc, char** argv) {
    FILE *in, *out;
    uint64_t hdr[5], N, d, kd, nq, np, nupd, ndel, i, j;
    double *rows, *queries, *model;
    uint64_t* pairs;
    VectorDatabase* db;
    int bad = 0;
    uint64_t* kd_ids;
    uint64_t* after;
    uint32_t* mbits;
    char* json = NULL;
    size_t json_len = 0;
    uint64_t loaded = 0, size_after = 0, model_n;

    if (argc != 4) {
        fprintf(stderr, "usage: %s IN OUT SAVEFILE\n", argv[0]);
        return 2;
    }
    in = fopen(argv[1], "rb");
    if (!in || !rd(in, hdr, sizeof(uint64_t), 5)) return 2;
    N = hdr[0]; d = hdr[1]; kd = hdr[2]; nq = hdr[3]; np = hdr[4];
    rows = read_rows(in, N, d);
    queries = read_rows(in, nq, d);
    pairs = (uint64_t*)malloc((size_t)(2 * np + 1) * sizeof(uint64_t));
    kd_ids = (uint64_t*)malloc((size_t)(nq + 1) * sizeof(uint64_t));
    after = (uint64_t*)malloc((size_t)(nq + 1) * sizeof(uint64_t));
    mbits = (uint32_t*)malloc((size_t)(3 * np + 1) * sizeof(uint32_t));
    model = (double*)malloc((size_t)(N * d + 1) * sizeof(double));
    if (!rows || !queries || !pairs || !kd_ids || !after || !mbits || !model ||
        !rd(in, pairs, sizeof(uint64_t), (size_t)(2 * np)))
        return 2;
    memcpy(model, rows, (size_t)(N * d) * sizeof(double));
    model_n = N;

    /* main.c:349-351 (no database file yet) */
    db = vector_db_load(argv[3], (size_t)kd);
    if (db == NULL) db = vector_db_init(0, (size_t)kd);
    if (!db) return 3;
    /* POST /vector for every row */
    for (i = 0; i < N; i++)
        if (post_vector(db, rows + i * d, (size_t)d, (size_t)i) != (size_t)i) {
            fprintf(stderr, "insert %llu failed\n", (unsigned long long)i);
            return 4;
        }
    /* POST /nearest: kdtree_nearest then vector_db_read */
    for (i = 0; i < nq; i++) {
        size_t idx = kdtree_nearest(db->kdtree, queries + i * d);
        Vector* v;
        kd_ids[i] = idx;
        if (idx == (size_t)-1) continue;
        v = vector_db_read(db, idx);
        if (!v || v->dimension != d || memcmp(v->data, model + idx * d, d * sizeof(double)))
            bad = 1;
    }
    /* GET /compare/<metric>?index1=&index2= */
    for (i = 0; i < np; i++) {
        size_t a = (size_t)pairs[2 * i], b = (size_t)pairs[2 * i + 1];
        Vector *v1, *v2;
        float r[3];
        if (a >= db->size || b >= db->size) return 4;
        v1 = vector_db_read(db, a);
        v2 = vector_db_read(db, b);
        if (!v1 || !v2 || v1->dimension != v2->dimension || v1->dimension != d) return 4;
        r[0] = cosine_similarity(*v1, *v2);
        r[1] = euclidean_distance(*v1, *v2);
        r[2] = dot_product(*v1, *v2);
        memcpy(mbits + 3 * i, r, sizeof r);
    }
    /* PUT /vector?index= */
    if (!rd(in, &nupd, sizeof nupd, 1)) return 2;
    for (j = 0; j < nupd; j++) {
        uint64_t idx;
        Vector vec;
        memset(&vec, 0, sizeof vec);
        vec.dimension = d;
        vec.data = (double*)malloc(d * sizeof(double));
        if (!vec.data || !rd(in, &idx, sizeof idx, 1) || !rd(in, vec.data, sizeof(double), d))
            return 2;
        memcpy(model + idx * d, vec.data, d * sizeof(double));
        snprintf(vec.uuid, UUID_SIZE, "upd-%08llu", (unsigned long long)idx);
        vector_db_update(db, (size_t)idx, vec);
    }
    /* DELETE /vector?index= */
    if (!rd(in, &ndel, sizeof ndel, 1)) return 2;
    for (j = 0; j < ndel; j++) {
        uint64_t idx;
        if (!rd(in, &idx, sizeof idx, 1)) return 2;
        if (idx >= db->size) continue; /* delete_handler.c:84 answers 400 */
        vector_db_delete(db, (size_t)idx);
        memmove(model + idx * d, model + (idx + 1) * d,
                (size_t)((model_n - idx - 1) * d) * sizeof(double));
        model_n--;
    }
    fclose(in);
    size_after = db->size;
    if (size_after != model_n) bad = 1;
    for (i = 0; i < db->size; i++) {
        Vector* v = vector_db_read(db, i);
        if (!v || memcmp(v->data, model + i * d, d * sizeof(double))) bad = 1;
    }
    /* /nearest again: the store as the mutations left it */
    for (i = 0; i < nq; i++) {
        size_t idx = kdtree_nearest(db->kdtree, queries + i * d);
        after[i] = idx;
        if (idx != (size_t)-1 &&
            memcmp(vector_db_read(db, idx)->data, model + idx * d, d * sizeof(double)))
            bad = 1;
    }
    /* the batched /nearest layer: the first two queries, top-3, dot */
    {
        char body[65536];
        size_t o = 0, q, c;
        o += (size_t)snprintf(body + o, sizeof body - o, "[");
        for (q = 0; q < 2 && q < nq; q++) {
            o += (size_t)snprintf(body + o, sizeof body - o, q ? ",[" : "[");
            for (c = 0; c < d && o < sizeof body - 64; c++)
                o += (size_t)snprintf(body + o, sizeof body - o, c ? ",%.17g" : "%.17g",
                                      queries[q * d + c]);
            o += (size_t)snprintf(body + o, sizeof body - o, "]");
        }
        o += (size_t)snprintf(body + o, sizeof body - o, "]");
        if (svdb_nearest_request(db, NULL, (size_t)d, body, o, "3", "dot", SVDB_PREC_AUTO, &json,
                                 &json_len) != 200)
            bad = 1;
    }
    /* main.c:398-402 */
    vector_db_save(db, argv[3]);
    vector_db_free(db);
    db = vector_db_load(argv[3], (size_t)kd);
    if (db) {
        loaded = db->size;
        for (i = 0; i < db->size; i++)
            if (memcmp(db->vectors[i].data, model + i * d, d * sizeof(double))) bad = 1;
        vector_db_free(db);
    }
    out = fopen(argv[2], "wb");
    if (!out) return 2;
    fwrite(kd_ids, sizeof(uint64_t), (size_t)nq, out);
    fwrite(mbits, sizeof(uint32_t), (size_t)(3 * np), out);
    fwrite(after, sizeof(uint64_t), (size_t)nq, out);
    fwrite(&size_after, sizeof size_after, 1, out);
    fwrite(&loaded, sizeof loaded, 1, out);
    {
        uint64_t jl = json_len;
        fwrite(&jl, sizeof jl, 1, out);
        if (json) fwrite(json, 1, json_len, out);
    }
    fclose(out);
    free(json);
    free(rows);
    free(queries);
    free(pairs);
    free(kd_ids);
    free(after);
    free(mbits);
    free(model);
    return bad ? 4 : 0;
}
