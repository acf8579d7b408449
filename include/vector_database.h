/*
 * vector_database.h — store and metric interface of simple-vector-db,
 * served on MI355X.
 *
 * Same declarations as the reference's include/vector_database.h
 * (vector_database.h:14-135); libsvdb.so exports them with C linkage.
 *
 * Layout (LP64): Vector = uuid[37] @0, dimension @40, data @48 (56 bytes);
 * VectorDatabase = vectors @0, size @8, capacity @16, kdtree @24, mutex @32,
 * as in the reference, then the engine handle.  `vectors` stays a host
 * mirror in logical (positional) order, so callers that read
 * db->vectors[i].data or .uuid keep working; every row also lives in HBM as
 * an FP64 master row with its norm and a low-precision shadow, and all
 * similarity work runs there.
 */
#ifndef VECTOR_DATABASE_H
#define VECTOR_DATABASE_H

#include <pthread.h>
#include <stddef.h>

#include "kdtree.h"

#ifdef __cplusplus
extern "C" {
#endif

#define UUID_SIZE 37 /* 36 characters + NUL */

typedef struct Vector {
    char uuid[UUID_SIZE];
    size_t dimension;
    double* data;
} Vector;

typedef struct VectorDatabase {
    Vector* vectors;        /* host mirror, logical order */
    size_t size;            /* number of rows */
    size_t capacity;        /* host mirror capacity */
    KDTree* kdtree;         /* find-nearest view of this store */
    pthread_mutex_t mutex;  /* guards the host mirror, as in the reference */
    void* svdb_engine;      /* device store + search engine */
    void* svdb_uuid;        /* uuid -> position index of the host mirror */
} VectorDatabase;

/* Replaces vector_db_init (vector_database.c:17-54; vector_database.h:39).
 * `dimension` is the kd-tree dimension (the number of leading coordinates
 * find-nearest compares); capacity 0 becomes 10.  NULL on failure. */
VectorDatabase* vector_db_init(size_t initial_capacity, size_t dimension);

/* Replaces vector_db_free (vector_database.c:61-72; vector_database.h:46). */
void vector_db_free(VectorDatabase* db);

/* Replaces vector_db_insert (vector_database.c:81-119; vector_database.h:55).
 * Takes ownership of vec.data (malloc'ed).  The first row fixes the store's
 * vector dimension; a row of another dimension is refused.  Returns the new
 * row's index or (size_t)-1. */
size_t vector_db_insert(VectorDatabase* db, Vector vec);

/* Replaces vector_db_read (vector_database.c:128-136; vector_database.h:64).
 * Borrowed pointer into the host mirror, or NULL when out of range. */
Vector* vector_db_read(VectorDatabase* db, size_t index);

/* Replaces vector_db_read_by_uuid (vector_database.c:145-158;
 * vector_database.h:73): first row whose uuid matches, or NULL. */
Vector* vector_db_read_by_uuid(VectorDatabase* db, const char* uuid);

/* Replaces vector_db_update (vector_database.c:169-177; vector_database.h:83).
 * Frees the old row, takes ownership of vec.data and rewrites the device
 * row, norm and shadow in place.  Out-of-range index or a different
 * dimension: nothing changes and vec.data is freed. */
void vector_db_update(VectorDatabase* db, size_t index, Vector vec);

/* Replaces vector_db_delete (vector_database.c:185-195; vector_database.h:91).
 * Later rows move down one index, as in the reference. */
void vector_db_delete(VectorDatabase* db, size_t index);

/* Replaces vector_db_save (vector_database.c:203-228; vector_database.h:99).
 * Same native-endian file format: size_t count, then per row uuid[37],
 * size_t dimension, double data[dimension].  The count written equals the
 * number of records written. */
void vector_db_save(VectorDatabase* db, const char* filename);

/* Replaces vector_db_load (vector_database.c:237-292; vector_database.h:108).
 * Reads that format and uploads all rows to the device in one bulk pass. */
VectorDatabase* vector_db_load(const char* filename, size_t dimension);

/* Replace cosine_similarity / euclidean_distance / dot_product
 * (vector_database.c:301-352; vector_database.h:117,126,135).  Computed on
 * the GPU with the reference's arithmetic (FP64 products, float
 * accumulation, double sqrt/division, float result), so results are
 * bit-identical to the reference.  -1.0f on mismatched dimensions.  For FP64
 * scores use svdb_pair_metrics (svdb.h). */
float cosine_similarity(Vector vec1, Vector vec2);
float euclidean_distance(Vector vec1, Vector vec2);
float dot_product(Vector vec1, Vector vec2);

#ifdef __cplusplus
}
#endif

#endif /* VECTOR_DATABASE_H */
