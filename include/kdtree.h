/*
 * kdtree.h — find-nearest interface of simple-vector-db, served on MI355X.
 *
 * Same declarations as the reference's include/kdtree.h (kdtree.h:10-57);
 * libsvdb.so exports them with C linkage.  The tree is no longer a linked
 * structure: find-nearest is an exhaustive batched scan on the GPU over the
 * first `dimension` coordinates of every stored row, so the answer is the
 * exact nearest row under squared Euclidean distance with ties broken by the
 * lowest index (the reference keeps the first node its depth-first search
 * visits, kdtree.c:139).
 *
 * Layout: `root` @0 and `dimension` @8 keep the reference offsets; the engine
 * handle is appended after them.  `root` stays NULL: nothing outside the
 * reference's kdtree.c dereferences it.
 */
#ifndef KDTREE_H
#define KDTREE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kept for source compatibility (kdtree.h:10-15); no nodes are allocated. */
typedef struct KDTreeNode {
    double* point;
    size_t index;
    struct KDTreeNode* left;
    struct KDTreeNode* right;
} KDTreeNode;

typedef struct KDTree {
    KDTreeNode* root;  /* always NULL (no linked nodes) */
    size_t dimension;  /* number of leading coordinates compared */
    void* svdb_index;  /* engine view: the owning database's device store, or a
                          private device point set for a stand-alone tree */
} KDTree;

/* Replaces kdtree_create (kdtree.c:70-78; kdtree.h:32).  A stand-alone tree
 * owns a device point set of `dimension` doubles per point.  NULL if
 * dimension is 0 or no GPU is usable. */
KDTree* kdtree_create(size_t dimension);

/* Replaces kdtree_insert (kdtree.c:87-91; kdtree.h:41).  Copies the first
 * `dimension` doubles of `point` (not owned) under label `index`.  A label
 * that is already present is overwritten in place (the reference adds a
 * second node).  On a tree owned by a VectorDatabase this is a no-op: the
 * database store already holds the row. */
void kdtree_insert(KDTree* tree, const double* point, size_t index);

/* Replaces kdtree_free (kdtree.c:112-118; kdtree.h:48). */
void kdtree_free(KDTree* tree);

/* Replaces kdtree_nearest (kdtree.c:171-178; kdtree.h:57).  Label of the
 * point nearest to `point` over the first `dimension` coordinates; lowest
 * label on ties; (size_t)-1 for a NULL or empty tree. */
size_t kdtree_nearest(KDTree* tree, const double* point);

#ifdef __cplusplus
}
#endif

#endif /* KDTREE_H */
