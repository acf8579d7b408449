/*
 * svdb_oracle.c — CPU oracle for the similarity / nearest-vector path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.
 *
 * Build: gcc -O2 -mavx2 -mfma -ffp-contract=off (oracle/Makefile).  With
 * contraction off every a*b+c below is a rounded multiply then a rounded
 * add, exactly as the reference's gcc build (no -march, so no FMA) computes
 * it; the explicit fma() calls are the canonical FP64 oracle's fused steps
 * and are correctly rounded on every x86-64 host (hardware FMA or glibc's
 * exact software fma).
 */
#include "svdb_oracle.h"

#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== *
 *  Reference kd-tree restatement                                            *
 *  kdtree.c:15-35  node = private copy of point[0..D) + index               *
 *  kdtree.c:47-62  insert: axis cd = depth % D; point[cd] < node[cd] → left *
 *                  else right                                               *
 *  kdtree.c:131-162 nearest: d = Σ_{i<D} (node_i - q_i)*(node_i - q_i)      *
 *                  (sequential, unfused); improve iff d < best (strict);    *
 *                  descend the q[cd] < node[cd] side first; visit the other *
 *                  side iff (q[cd]-node[cd])^2 < best (strict)              *
 *  kdtree.c:171-178 empty / NULL tree → (size_t)-1                         *
 * ======================================================================== */
typedef struct {
    size_t index;
    int64_t left, right;
} kdn;

struct orc_kd {
    size_t D;
    kdn* nodes;
    double* pts;
    size_t n, cap;
    int64_t root;
    size_t maxdepth;
};

orc_kd* orc_kd_create(size_t dimension) {
    if (dimension == 0) return NULL;
    orc_kd* t = (orc_kd*)calloc(1, sizeof(orc_kd));
    if (!t) return NULL;
    t->D = dimension;
    t->root = -1;
    return t;
}

static int kd_grow(orc_kd* t) {
    size_t nc = t->cap ? t->cap * 2 : 1024;
    kdn* nn = (kdn*)realloc(t->nodes, nc * sizeof(kdn));
    if (!nn) return -1;
    t->nodes = nn;
    double* np = (double*)realloc(t->pts, nc * t->D * sizeof(double));
    if (!np) return -1;
    t->pts = np;
    t->cap = nc;
    return 0;
}

void orc_kd_insert(orc_kd* t, const double* point, size_t index) {
    if (!t) return;
    if (t->n == t->cap && kd_grow(t) != 0) return;
    int64_t me = (int64_t)t->n++;
    memcpy(t->pts + (size_t)me * t->D, point, t->D * sizeof(double));
    t->nodes[me].index = index;
    t->nodes[me].left = t->nodes[me].right = -1;
    if (t->root < 0) {
        t->root = me;
        return;
    }
    int64_t cur = t->root;
    size_t depth = 0;
    for (;;) {
        size_t cd = depth % t->D;
        const double* p = t->pts + (size_t)cur * t->D;
        int64_t* link = (point[cd] < p[cd]) ? &t->nodes[cur].left : &t->nodes[cur].right;
        depth++;
        if (*link < 0) {
            *link = me;
            if (depth > t->maxdepth) t->maxdepth = depth;
            return;
        }
        cur = *link;
    }
}

size_t orc_kd_size(const orc_kd* t) { return t ? t->n : 0; }
size_t orc_kd_depth(const orc_kd* t) { return t ? t->maxdepth : 0; }

void orc_kd_free(orc_kd* t) {
    if (!t) return;
    free(t->nodes);
    free(t->pts);
    free(t);
}

orc_kd* orc_kd_build(const double* X, size_t n, size_t ldx, size_t dimension) {
    orc_kd* t = orc_kd_create(dimension);
    if (!t) return NULL;
    for (size_t i = 0; i < n; i++) orc_kd_insert(t, X + i * ldx, i);
    return t;
}

/* Explicit-stack form of the reference's recursion; the visiting order is
 * the same: a frame either visits a node or, after the near subtree is
 * done, tests whether the far subtree can still hold a closer point. */
typedef struct {
    int64_t node;
    size_t depth;
    int check_other;
} kdframe;

size_t orc_kd_nearest(const orc_kd* t, const double* q) {
    if (!t || t->root < 0) return (size_t)-1;
    const size_t D = t->D;
    size_t cap = 2 * (t->maxdepth + 2);
    kdframe stackbuf[512];
    kdframe* st = cap <= 512 ? stackbuf : (kdframe*)malloc(cap * sizeof(kdframe));
    if (!st) return (size_t)-1;
    size_t sp = 0;
    double best_d = INFINITY;
    int64_t best = -1;
    st[sp++] = (kdframe){t->root, 0, 0};
    while (sp) {
        kdframe f = st[--sp];
        const double* p = t->pts + (size_t)f.node * D;
        size_t cd = f.depth % D;
        if (f.check_other) {
            double diff = q[cd] - p[cd];
            if (diff * diff < best_d) {
                int64_t other = (q[cd] < p[cd]) ? t->nodes[f.node].right : t->nodes[f.node].left;
                if (other >= 0) st[sp++] = (kdframe){other, f.depth + 1, 0};
            }
            continue;
        }
        double d = 0;
        for (size_t i = 0; i < D; i++) d += (p[i] - q[i]) * (p[i] - q[i]);
        if (d < best_d) {
            best_d = d;
            best = f.node;
        }
        int64_t next = (q[cd] < p[cd]) ? t->nodes[f.node].left : t->nodes[f.node].right;
        st[sp++] = (kdframe){f.node, f.depth, 1};
        if (next >= 0) st[sp++] = (kdframe){next, f.depth + 1, 0};
    }
    if (st != stackbuf) free(st);
    return best >= 0 ? t->nodes[best].index : (size_t)-1;
}

typedef struct {
    const orc_kd* t;
    const double* Q;
    size_t ldq, lo, hi;
    uint64_t* out;
} kdjob;

static void* kd_worker(void* arg) {
    kdjob* j = (kdjob*)arg;
    for (size_t i = j->lo; i < j->hi; i++) j->out[i] = orc_kd_nearest(j->t, j->Q + i * j->ldq);
    return NULL;
}

void orc_kd_nearest_batch(const orc_kd* t, const double* Q, size_t nq, size_t ldq,
                          int nthreads, uint64_t* out) {
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > nq) nthreads = nq ? (int)nq : 1;
    pthread_t th[256];
    kdjob jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (int w = 0; w < nthreads; w++) {
        jobs[w] = (kdjob){t, Q, ldq, nq * w / nthreads, nq * (w + 1) / nthreads, out};
        if (nthreads == 1) kd_worker(&jobs[w]);
        else pthread_create(&th[w], NULL, kd_worker, &jobs[w]);
    }
    if (nthreads > 1)
        for (int w = 0; w < nthreads; w++) pthread_join(th[w], NULL);
}

/* ======================================================================== *
 *  Reference float metrics (vector_database.c:301-352)                       *
 *  Accumulators are float; each double product is added in double and the  *
 *  sum rounded back to float (C's float += double).  Euclidean rounds the   *
 *  double difference to float and squares/accumulates in float.  sqrt and   *
 *  the cosine division run in double; the result is returned as float.     *
 * ======================================================================== */
float orc_ref_cos(const double* a, const double* b, size_t n) {
    float dp = 0.0f, na = 0.0f, nb = 0.0f;
    for (size_t i = 0; i < n; i++) {
        dp = (float)((double)dp + a[i] * b[i]);
        na = (float)((double)na + a[i] * a[i]);
        nb = (float)((double)nb + b[i] * b[i]);
    }
    return (float)((double)dp / (sqrt((double)na) * sqrt((double)nb)));
}

float orc_ref_l2(const double* a, const double* b, size_t n) {
    float sum = 0.0f;
    for (size_t i = 0; i < n; i++) {
        float diff = (float)(a[i] - b[i]);
        float sq = diff * diff;
        sum = sum + sq;
    }
    return (float)sqrt((double)sum);
}

float orc_ref_dot(const double* a, const double* b, size_t n) {
    float r = 0.0f;
    for (size_t i = 0; i < n; i++) r = (float)((double)r + a[i] * b[i]);
    return r;
}

void orc_ref_pairs(const double* A, size_t lda, size_t n, const uint64_t* ia,
                   const uint64_t* ib, size_t np, int metric, float* out) {
    for (size_t p = 0; p < np; p++) {
        const double* a = A + ia[p] * lda;
        const double* b = A + ib[p] * lda;
        out[p] = metric == ORC_COS ? orc_ref_cos(a, b, n)
               : metric == ORC_L2  ? orc_ref_l2(a, b, n)
                                   : orc_ref_dot(a, b, n);
    }
}

/* ======================================================================== *
 *  Canonical FP64 arithmetic                                                *
 *  L2 key  : s = 0; for i in [0,D): t = x_i - q_i; s = fma(t, t, s)         *
 *  dot key : s = 0; for i in [0,D): s = fma(q_i, x_i, s)                    *
 *  sumsq   : s = 0; for i in [0,D): s = fma(x_i, x_i, s)                    *
 *  cosine  : dot / (sqrt(nq) * sqrt(nx))  (operand order of               *
 *            vector_database.c:312; * is commutative in IEEE)              *
 * ======================================================================== */
double orc_l2_key(const double* q, const double* x, size_t D) {
    double s = 0.0;
    for (size_t i = 0; i < D; i++) {
        double t = x[i] - q[i];
        s = fma(t, t, s);
    }
    return s;
}

double orc_dot_key(const double* q, const double* x, size_t D) {
    double s = 0.0;
    for (size_t i = 0; i < D; i++) s = fma(q[i], x[i], s);
    return s;
}

double orc_sumsq(const double* x, size_t D) {
    double s = 0.0;
    for (size_t i = 0; i < D; i++) s = fma(x[i], x[i], s);
    return s;
}

double orc_cos_score(double dot, double nq, double nx) { return dot / (sqrt(nq) * sqrt(nx)); }

double orc_rank_key(double v, int metric) { return metric == ORC_L2 ? v : -v; }

/* Ranking order: ascending rank key, NaN after every number, then lowest id. */
static inline int rk_less(double ka, uint64_t ia, double kb, uint64_t ib) {
    int na = isnan(ka), nb = isnan(kb);
    if (na || nb) {
        if (na && nb) return ia < ib;
        return nb;
    }
    if (ka < kb) return 1;
    if (ka > kb) return 0;
    return ia < ib;
}

typedef struct {
    double* key;
    uint64_t* id;
    size_t cnt;
} topk_t;

static inline void topk_push(topk_t* t, size_t k, double key, uint64_t id) {
    size_t pos;
    if (t->cnt == k) {
        if (!rk_less(key, id, t->key[k - 1], t->id[k - 1])) return;
        pos = k - 1;
    } else {
        pos = t->cnt++;
    }
    while (pos > 0 && rk_less(key, id, t->key[pos - 1], t->id[pos - 1])) {
        t->key[pos] = t->key[pos - 1];
        t->id[pos] = t->id[pos - 1];
        pos--;
    }
    t->key[pos] = key;
    t->id[pos] = id;
}

static void topk_emit(const topk_t* t, size_t k, int metric, uint64_t* ids, double* sc) {
    for (size_t j = 0; j < k; j++) {
        if (j < t->cnt) {
            ids[j] = t->id[j];
            /* every NaN score is emitted as the canonical quiet NaN and
             * every zero score as +0: the ranking folds -0 onto +0, so
             * neither a NaN's payload nor a zero's sign carries meaning */
            sc[j] = isnan(t->key[j])   ? NAN
                    : metric == ORC_L2 ? sqrt(t->key[j])
                    : t->key[j] == 0.0 ? 0.0
                                       : -t->key[j];
        } else {
            ids[j] = UINT64_MAX;
            sc[j] = NAN;
        }
    }
}

static uint64_t* logical_ids(const uint8_t* alive, size_t N) {
    uint64_t* lg = (uint64_t*)malloc((N ? N : 1) * sizeof(uint64_t));
    if (!lg) return NULL;
    uint64_t c = 0;
    for (size_t r = 0; r < N; r++) {
        lg[r] = c;
        if (!alive || alive[r]) c++;
    }
    return lg;
}

static inline double rank_key_of(int metric, double key_or_dot, double nq, double nx) {
    if (metric == ORC_L2) return key_or_dot;
    if (metric == ORC_DOT) return -key_or_dot;
    return -orc_cos_score(key_or_dot, nq, nx);
}

int orc_topk_scalar(const double* X, size_t N, size_t ldx, const uint8_t* alive,
                    const double* Q, size_t B, size_t ldq, size_t D, int metric,
                    size_t k, uint64_t* out_ids, double* out_scores) {
    if (k == 0 || (metric != ORC_L2 && metric != ORC_COS && metric != ORC_DOT)) return -1;
    uint64_t* lg = logical_ids(alive, N);
    double* kk = (double*)malloc(k * sizeof(double));
    uint64_t* ii = (uint64_t*)malloc(k * sizeof(uint64_t));
    if (!lg || !kk || !ii) {
        free(lg); free(kk); free(ii);
        return -1;
    }
    for (size_t b = 0; b < B; b++) {
        const double* q = Q + b * ldq;
        double nq = orc_sumsq(q, D);
        topk_t t = {kk, ii, 0};
        for (size_t r = 0; r < N; r++) {
            if (alive && !alive[r]) continue;
            const double* x = X + r * ldx;
            double v = metric == ORC_L2 ? orc_l2_key(q, x, D) : orc_dot_key(q, x, D);
            double nx = metric == ORC_COS ? orc_sumsq(x, D) : 0.0;
            topk_push(&t, k, rank_key_of(metric, v, nq, nx), lg[r]);
        }
        topk_emit(&t, k, metric, out_ids + b * k, out_scores + b * k);
    }
    free(lg); free(kk); free(ii);
    return 0;
}

/* ---- vectorised, threaded form ------------------------------------------
 * Rows are processed in blocks of RB, transposed so that four consecutive
 * rows share one 256-bit register; each lane runs the same sequential fma
 * chain as orc_l2_key / orc_dot_key, so results are bit-identical to the
 * scalar form.  Threads split the queries. */
#define RB 32

typedef struct {
    const double* X;
    size_t N, ldx;
    const uint8_t* alive;
    const uint64_t* lg;
    const double* Q;
    size_t ldq, D, k, b0, b1;
    int metric;
    uint64_t* out_ids;
    double* out_scores;
    int err;
} tkjob;

__attribute__((target("avx2,fma"))) static void block_keys(const double* buf, const double* q,
                                                            size_t D, int metric, double* keys) {
    __m256d acc[RB / 4];
    for (int v = 0; v < RB / 4; v++) acc[v] = _mm256_setzero_pd();
    if (metric == ORC_L2) {
        for (size_t i = 0; i < D; i++) {
            __m256d qb = _mm256_broadcast_sd(q + i);
            const double* xr = buf + i * RB;
            for (int v = 0; v < RB / 4; v++) {
                __m256d t = _mm256_sub_pd(_mm256_loadu_pd(xr + 4 * v), qb);
                acc[v] = _mm256_fmadd_pd(t, t, acc[v]);
            }
        }
    } else {
        for (size_t i = 0; i < D; i++) {
            __m256d qb = _mm256_broadcast_sd(q + i);
            const double* xr = buf + i * RB;
            for (int v = 0; v < RB / 4; v++)
                acc[v] = _mm256_fmadd_pd(qb, _mm256_loadu_pd(xr + 4 * v), acc[v]);
        }
    }
    for (int v = 0; v < RB / 4; v++) _mm256_storeu_pd(keys + 4 * v, acc[v]);
}

__attribute__((target("avx2,fma"))) static void block_sumsq(const double* buf, size_t D,
                                                             double* out) {
    __m256d acc[RB / 4];
    for (int v = 0; v < RB / 4; v++) acc[v] = _mm256_setzero_pd();
    for (size_t i = 0; i < D; i++) {
        const double* xr = buf + i * RB;
        for (int v = 0; v < RB / 4; v++) {
            __m256d x = _mm256_loadu_pd(xr + 4 * v);
            acc[v] = _mm256_fmadd_pd(x, x, acc[v]);
        }
    }
    for (int v = 0; v < RB / 4; v++) _mm256_storeu_pd(out + 4 * v, acc[v]);
}

static void* topk_worker(void* arg) {
    tkjob* j = (tkjob*)arg;
    const size_t D = j->D, k = j->k, nb = j->b1 - j->b0;
    double* buf = (double*)aligned_alloc(64, ((D * RB * sizeof(double)) + 63) / 64 * 64);
    double* kk = (double*)malloc(nb * k * sizeof(double));
    uint64_t* ii = (uint64_t*)malloc(nb * k * sizeof(uint64_t));
    topk_t* tk = (topk_t*)malloc(nb * sizeof(topk_t));
    double* nq = (double*)malloc(nb * sizeof(double));
    if (!buf || !kk || !ii || !tk || !nq) {
        j->err = -1;
        free(buf); free(kk); free(ii); free(tk); free(nq);
        return NULL;
    }
    for (size_t b = 0; b < nb; b++) {
        tk[b] = (topk_t){kk + b * k, ii + b * k, 0};
        nq[b] = orc_sumsq(j->Q + (j->b0 + b) * j->ldq, D);
    }
    double keys[RB], nx[RB];
    for (size_t r0 = 0; r0 < j->N; r0 += RB) {
        size_t rn = j->N - r0 < RB ? j->N - r0 : RB;
        for (size_t i = 0; i < D; i++)
            for (size_t c = 0; c < RB; c++)
                buf[i * RB + c] = c < rn ? j->X[(r0 + c) * j->ldx + i] : 0.0;
        if (j->metric == ORC_COS) block_sumsq(buf, D, nx);
        for (size_t b = 0; b < nb; b++) {
            block_keys(buf, j->Q + (j->b0 + b) * j->ldq, D, j->metric, keys);
            for (size_t c = 0; c < rn; c++) {
                size_t r = r0 + c;
                if (j->alive && !j->alive[r]) continue;
                topk_push(&tk[b], k, rank_key_of(j->metric, keys[c], nq[b], nx[c]), j->lg[r]);
            }
        }
    }
    for (size_t b = 0; b < nb; b++)
        topk_emit(&tk[b], k, j->metric, j->out_ids + (j->b0 + b) * k,
                  j->out_scores + (j->b0 + b) * k);
    free(buf); free(kk); free(ii); free(tk); free(nq);
    return NULL;
}

int orc_topk(const double* X, size_t N, size_t ldx, const uint8_t* alive,
             const double* Q, size_t B, size_t ldq, size_t D, int metric, size_t k,
             int nthreads, uint64_t* out_ids, double* out_scores) {
    if (k == 0 || (metric != ORC_L2 && metric != ORC_COS && metric != ORC_DOT)) return -1;
    if (B == 0) return 0;
    uint64_t* lg = logical_ids(alive, N);
    if (!lg) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((size_t)nthreads > B) nthreads = (int)B;
    pthread_t th[256];
    tkjob jobs[256];
    int rc = 0;
    for (int w = 0; w < nthreads; w++) {
        jobs[w] = (tkjob){X, N, ldx, alive, lg, Q, ldq, D, k,
                          B * w / nthreads, B * (w + 1) / nthreads, metric,
                          out_ids, out_scores, 0};
        if (nthreads == 1) topk_worker(&jobs[w]);
        else pthread_create(&th[w], NULL, topk_worker, &jobs[w]);
    }
    for (int w = 0; w < nthreads; w++) {
        if (nthreads > 1) pthread_join(th[w], NULL);
        if (jobs[w].err) rc = -1;
    }
    free(lg);
    return rc;
}

/* ======================================================================== *
 *  Canonical FP64 pair metric: lane l (0..63) accumulates the elements      *
 *  i ≡ l (mod 64) in increasing i with fma, then six xor-butterfly steps    *
 *  (offsets 32,16,8,4,2,1) add partner partials.  IEEE addition is          *
 *  commutative, so every lane ends with the same value; lane 0's is the     *
 *  result.  L2 → sqrt(Σ t^2) with t = a_i - b_i; DOT → Σ a_i b_i;           *
 *  COS → dot / (sqrt(Σ a_i^2) * sqrt(Σ b_i^2)).                             *
 * ======================================================================== */
static double lanesum(double* part) {
    for (int off = 32; off >= 1; off >>= 1) {
        double nxt[64];
        for (int l = 0; l < 64; l++) nxt[l] = part[l] + part[l ^ off];
        memcpy(part, nxt, sizeof(nxt));
    }
    return part[0];
}

double orc_pair_fp64(const double* a, const double* b, size_t D, int metric) {
    double pd[64] = {0}, pa[64] = {0}, pb[64] = {0};
    for (size_t i = 0; i < D; i++) {
        int l = (int)(i & 63);
        if (metric == ORC_L2) {
            double t = a[i] - b[i];
            pd[l] = fma(t, t, pd[l]);
        } else {
            pd[l] = fma(a[i], b[i], pd[l]);
            if (metric == ORC_COS) {
                pa[l] = fma(a[i], a[i], pa[l]);
                pb[l] = fma(b[i], b[i], pb[l]);
            }
        }
    }
    double d = lanesum(pd);
    if (metric == ORC_L2) return sqrt(d);
    if (metric == ORC_DOT) return d;
    double na = lanesum(pa), nb = lanesum(pb);
    return d / (sqrt(na) * sqrt(nb));
}

void orc_pairs_fp64(const double* A, size_t lda, size_t D, const uint64_t* ia,
                    const uint64_t* ib, size_t np, int metric, double* out) {
    for (size_t p = 0; p < np; p++)
        out[p] = orc_pair_fp64(A + ia[p] * lda, A + ib[p] * lda, D, metric);
}
