/*
 * refkd_bench.c — timing harness for the reference's own CPU path.
 *
 * TEST INFRASTRUCTURE ONLY: used by bench.py's cpu_baseline leg and by
 * tests/, never by the product library.
 *
 * The reference answers find-nearest with kdtree_nearest over a tree that
 * vector_db_insert grows one kdtree_insert at a time (kdtree.c:47-91), each
 * insert printing a line per level — at ten million rows that is minutes of
 * stdout before the first query.  orc_reftree_build lays out the same tree
 * directly, in the reference's node structure (kdtree.h: KDTreeNode
 * {point, index, left, right}, KDTree {root, dimension}):
 *
 *   the tree kdtree_insert builds from rows 0..n-1 in order is: row 0 at the
 *   root; below it, on the left the tree of the later rows with
 *   row[cd] < root[cd] (cd = depth % D), in their order, on the right the
 *   tree of the others, in their order (kdtree.c:55-58)
 *
 * so a stable partition per level, recursing into both sides (in parallel
 * near the top), gives node for node the tree of the sequential inserts.
 * Each node's point is its own copy of the row's first D coordinates, as
 * kdtree_create_node makes (kdtree.c:20-28); the copies sit in insertion
 * order in one block, where the reference's mallocs would put them one
 * after the other.  The timed search is the reference's compiled
 * kdtree_nearest, called through a function pointer (oracle/_ref).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "svdb_oracle.h"

typedef struct RefNode {
    double* point;
    size_t index;
    struct RefNode* left;
    struct RefNode* right;
} RefNode;

typedef struct RefTree {
    RefNode* root;
    size_t dimension;
} RefTree;

struct orc_reftree {
    RefTree tree;  /* first member: &t->tree is what kdtree_nearest takes */
    RefNode* nodes;
    double* pts;
    size_t n;
};

typedef struct {
    const double* X;
    size_t ldx, D;
    RefNode* nodes;
    uint32_t* idx;   /* rows of this subtree, insertion order */
    uint32_t* tmp;   /* scratch of the same length */
    size_t cnt, depth;
    int spawn;       /* levels below which this call may still spawn threads */
    RefNode** link;  /* where the subtree's root goes */
} build_job;

static void* build_rec(void* arg);

static void build_run(build_job* j) {
    while (j->cnt) {
        const uint32_t r = j->idx[0];
        RefNode* root = &j->nodes[r];
        const size_t cd = j->depth % j->D;
        const double s = j->X[(size_t)r * j->ldx + cd];
        size_t nl = 0, nr = 0, i;
        *j->link = root;
        /* stable partition of the later rows: left (< s) to the front of
         * tmp, right (>= s) to its back, then the right run reversed back
         * into order */
        for (i = 1; i < j->cnt; i++) {
            const uint32_t v = j->idx[i];
            if (j->X[(size_t)v * j->ldx + cd] < s)
                j->tmp[nl++] = v;
            else
                j->tmp[j->cnt - 1 - nr++] = v;
        }
        for (i = 0; i < nl; i++) j->idx[i] = j->tmp[i];
        for (i = 0; i < nr; i++) j->idx[nl + i] = j->tmp[j->cnt - 1 - i];
        {
            build_job L = *j, R = *j;
            L.idx = j->idx;
            L.tmp = j->tmp;
            L.cnt = nl;
            L.depth = j->depth + 1;
            L.link = &root->left;
            R.idx = j->idx + nl;
            R.tmp = j->tmp + nl;
            R.cnt = nr;
            R.depth = j->depth + 1;
            R.link = &root->right;
            if (j->spawn > 0 && nl > 65536 && nr > 65536) {
                pthread_t th;
                L.spawn = R.spawn = j->spawn - 1;
                if (pthread_create(&th, NULL, build_rec, &L) == 0) {
                    build_run(&R);
                    pthread_join(th, NULL);
                    return;
                }
            }
            L.spawn = R.spawn = 0;
            build_run(&L);
            *j = R;  /* the right side iteratively */
        }
    }
}

static void* build_rec(void* arg) {
    build_run((build_job*)arg);
    return NULL;
}

typedef struct {
    const double* X;
    size_t ldx, D, lo, hi;
    RefNode* nodes;
    double* pts;
} copy_job;

static void* copy_rows(void* arg) {
    copy_job* c = (copy_job*)arg;
    size_t i;
    for (i = c->lo; i < c->hi; i++) {
        memcpy(c->pts + i * c->D, c->X + i * c->ldx, c->D * sizeof(double));
        c->nodes[i].point = c->pts + i * c->D;
        c->nodes[i].index = i;
        c->nodes[i].left = c->nodes[i].right = NULL;
    }
    return NULL;
}

orc_reftree* orc_reftree_build(const double* X, size_t n, size_t ldx, size_t D, int nthreads) {
    orc_reftree* t;
    uint32_t *idx, *tmp;
    size_t i;
    int w, spawn = 0;
    pthread_t th[256];
    copy_job cj[256];
    build_job root;
    if (!X || n == 0 || D == 0 || D > ldx || n > UINT32_MAX) return NULL;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    t = (orc_reftree*)calloc(1, sizeof *t);
    if (!t) return NULL;
    t->nodes = (RefNode*)malloc(n * sizeof(RefNode));
    t->pts = (double*)malloc(n * D * sizeof(double));
    idx = (uint32_t*)malloc(n * sizeof(uint32_t));
    tmp = (uint32_t*)malloc(n * sizeof(uint32_t));
    if (!t->nodes || !t->pts || !idx || !tmp) {
        free(idx);
        free(tmp);
        orc_reftree_free(t);
        return NULL;
    }
    for (w = 0; w < nthreads; w++) {
        cj[w] = (copy_job){X, ldx, D, n * (size_t)w / nthreads, n * (size_t)(w + 1) / nthreads,
                           t->nodes, t->pts};
        if (pthread_create(&th[w], NULL, copy_rows, &cj[w]) != 0) {
            copy_rows(&cj[w]);
            th[w] = 0;
        }
    }
    for (w = 0; w < nthreads; w++)
        if (th[w]) pthread_join(th[w], NULL);
    for (i = 0; i < n; i++) idx[i] = (uint32_t)i;
    while ((1 << spawn) < nthreads) spawn++;
    root = (build_job){t->pts, D, D, t->nodes, idx, tmp, n, 0, spawn + 1, &t->tree.root};
    build_run(&root);
    free(idx);
    free(tmp);
    t->tree.dimension = D;
    t->n = n;
    return t;
}

void* orc_reftree_handle(orc_reftree* t) { return t ? &t->tree : NULL; }

/* Row index at the node reached from the root by `path` (0 = left,
 * 1 = right, one byte per level), or (size_t)-1 when that node is absent:
 * lets tests compare this layout with a tree grown by kdtree_insert. */
size_t orc_reftree_at(const void* tree, const unsigned char* path, size_t len) {
    const RefNode* nd = ((const RefTree*)tree)->root;
    size_t i;
    for (i = 0; i < len && nd; i++) nd = path[i] ? nd->right : nd->left;
    return nd ? nd->index : (size_t)-1;
}

/* Nodes and height of a tree in the reference's layout. */
static void ref_walk(const RefNode* nd, size_t depth, size_t* count, size_t* height) {
    while (nd) {
        (*count)++;
        if (depth + 1 > *height) *height = depth + 1;
        ref_walk(nd->left, depth + 1, count, height);
        nd = nd->right;
        depth++;
    }
}

void orc_reftree_shape(const void* tree, size_t* count, size_t* height) {
    *count = *height = 0;
    ref_walk(((const RefTree*)tree)->root, 0, count, height);
}

void orc_reftree_free(orc_reftree* t) {
    if (!t) return;
    free(t->nodes);
    free(t->pts);
    free(t);
}

/* ------------------------------------------------------------- timing */

typedef size_t (*kd_nearest_fn)(void* tree, const double* point);

typedef struct {
    kd_nearest_fn fn;
    void* tree;
    const double* Q;
    size_t nq, ldq, w, stride;
    uint64_t* ids;
    double* ns;
} kd_job;

static double now_ns(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e9 + (double)ts.tv_nsec;
}

static void* kd_time_worker(void* arg) {
    kd_job* j = (kd_job*)arg;
    size_t q;
    for (q = j->w; q < j->nq; q += j->stride) {
        const double t0 = now_ns();
        j->ids[q] = j->fn(j->tree, j->Q + q * j->ldq);
        j->ns[q] = now_ns() - t0;
    }
    return NULL;
}

/* Calls fn(tree, Q + q*ldq) for every query on nthreads threads (thread w
 * takes queries w, w + nthreads, ...: independent query streams, as the
 * reference's thread-per-connection server runs them).  ids[q] gets the
 * answer, ns[q] the call's own time.  Returns the wall time in seconds. */
double orc_time_kd(void* fn, void* tree, const double* Q, size_t nq, size_t ldq, int nthreads,
                   uint64_t* ids, double* ns) {
    pthread_t th[256];
    kd_job jobs[256];
    int w;
    double t0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((size_t)nthreads > nq) nthreads = nq ? (int)nq : 1;
    t0 = now_ns();
    for (w = 0; w < nthreads; w++) {
        jobs[w] = (kd_job){(kd_nearest_fn)fn, tree, Q, nq, ldq, (size_t)w, (size_t)nthreads, ids,
                           ns};
        if (pthread_create(&th[w], NULL, kd_time_worker, &jobs[w]) != 0) {
            kd_time_worker(&jobs[w]);
            th[w] = 0;
        }
    }
    for (w = 0; w < nthreads; w++)
        if (th[w]) pthread_join(th[w], NULL);
    return (now_ns() - t0) * 1e-9;
}

/* The reference's Vector (vector_database.h:14-18), passed by value to its
 * metric functions. */
typedef struct {
    char uuid[37];
    size_t dimension;
    double* data;
} RefVector;

typedef float (*ref_metric_fn)(RefVector a, RefVector b);

typedef struct {
    ref_metric_fn fn;
    const double* X;
    size_t ldx, d;
    const uint64_t *ia, *ib;
    size_t lo, hi;
    float* out;
} pair_job;

static void* pair_worker(void* arg) {
    pair_job* j = (pair_job*)arg;
    RefVector a, b;
    size_t p;
    memset(&a, 0, sizeof a);
    memset(&b, 0, sizeof b);
    a.dimension = b.dimension = j->d;
    for (p = j->lo; p < j->hi; p++) {
        a.data = (double*)(j->X + j->ia[p] * j->ldx);
        b.data = (double*)(j->X + j->ib[p] * j->ldx);
        j->out[p] = j->fn(a, b);
    }
    return NULL;
}

/* out[p] = fn(row ia[p], row ib[p]) for np pairs on nthreads threads, fn
 * being one of the reference's cosine_similarity / euclidean_distance /
 * dot_product.  Returns the wall time in seconds. */
double orc_time_pairs(void* fn, const double* X, size_t ldx, size_t d, const uint64_t* ia,
                      const uint64_t* ib, size_t np, int nthreads, float* out) {
    pthread_t th[256];
    pair_job jobs[256];
    int w;
    double t0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    t0 = now_ns();
    for (w = 0; w < nthreads; w++) {
        jobs[w] = (pair_job){(ref_metric_fn)fn, X, ldx, d, ia, ib, np * (size_t)w / nthreads,
                             np * (size_t)(w + 1) / nthreads, out};
        if (pthread_create(&th[w], NULL, pair_worker, &jobs[w]) != 0) {
            pair_worker(&jobs[w]);
            th[w] = 0;
        }
    }
    for (w = 0; w < nthreads; w++)
        if (th[w]) pthread_join(th[w], NULL);
    return (now_ns() - t0) * 1e-9;
}
