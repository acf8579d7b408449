/*
 * svdb_nearest.c — POST /nearest between "the body is complete" and "queue
 * the response": parse, validate, search, lay out the JSON answer.
 *
 * Follows nearest_handler_callback (compare_handler.c:263-441) for the
 * validation order, error texts and the k = 1 response object, and adds
 * ?number= (README.md:287), ?metric= and batch bodies; the contract is in
 * svdb_nearest.h.  Plain C99: the reference's handlers are C, and this is
 * the part of a handler they would call.
 */
#include "svdb_nearest.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ JSON */

/* A parsed JSON value.  Arrays keep their elements and objects their member
 * values (cJSON_GetArraySize / cJSON_GetArrayItem count and index both the
 * same way, compare_handler.c:319,372); strings and literals keep only their
 * type. */
enum { J_NUMBER, J_ARRAY, J_OBJECT, J_OTHER };

typedef struct JNode {
    int type;
    double num;
    size_t n, cap;
    struct JNode* kids;
} JNode;

typedef struct {
    const char* p;
    const char* end;
    int oom;
} JCur;

#define J_MAX_DEPTH 512

static void j_free(JNode* n) {
    size_t i;
    for (i = 0; i < n->n; i++) j_free(&n->kids[i]);
    free(n->kids);
    n->kids = NULL;
    n->n = n->cap = 0;
}

static void j_ws(JCur* c) {
    while (c->p < c->end && (*c->p == ' ' || *c->p == '\t' || *c->p == '\n' || *c->p == '\r'))
        c->p++;
}

static int j_lit(JCur* c, const char* w) {
    size_t l = strlen(w);
    if ((size_t)(c->end - c->p) < l || memcmp(c->p, w, l) != 0) return 0;
    c->p += l;
    return 1;
}

static int j_string(JCur* c) {
    if (c->p >= c->end || *c->p != '"') return 0;
    c->p++;
    while (c->p < c->end && *c->p != '"') {
        if ((unsigned char)*c->p < 0x20) return 0;
        if (*c->p == '\\') {
            c->p++;
            if (c->p >= c->end) return 0;
            if (*c->p == 'u') {
                int i;
                for (i = 0; i < 4; i++) {
                    c->p++;
                    if (c->p >= c->end || !strchr("0123456789abcdefABCDEF", *c->p)) return 0;
                }
            } else if (!strchr("\"\\/bfnrt", *c->p)) {
                return 0;
            }
        }
        c->p++;
    }
    if (c->p >= c->end) return 0;
    c->p++;
    return 1;
}

/* A number as cJSON reads one: it starts with '-' or a digit; the run of
 * characters from "0123456789+-eE." that follows goes to strtod, and the
 * parse continues after what strtod consumed (so "01" and "1." are read as
 * numbers, and "1.2.3" leaves ".3" behind to fail the enclosing value). */
static int j_number(JCur* c, double* out) {
    char buf[128];
    char* h = buf;
    char* after;
    size_t n = 0;
    if (c->p >= c->end || !(*c->p == '-' || (*c->p >= '0' && *c->p <= '9'))) return 0;
    while (c->p + n < c->end && strchr("0123456789+-eE.", c->p[n]) && c->p[n]) n++;
    if (n >= sizeof buf) {
        h = (char*)malloc(n + 1);
        if (!h) {
            c->oom = 1;
            return 0;
        }
    }
    memcpy(h, c->p, n);
    h[n] = '\0';
    *out = strtod(h, &after);
    n = (size_t)(after - h);
    if (h != buf) free(h);
    if (n == 0) return 0;
    c->p += n;
    return 1;
}

static int j_value(JCur* c, JNode* out, int depth);

static int j_push(JCur* c, JNode* parent, JNode* kid) {
    if (parent->n == parent->cap) {
        size_t nc = parent->cap ? 2 * parent->cap : 8;
        JNode* k = (JNode*)realloc(parent->kids, nc * sizeof(JNode));
        if (!k) {
            c->oom = 1;
            return 0;
        }
        parent->kids = k;
        parent->cap = nc;
    }
    parent->kids[parent->n++] = *kid;
    return 1;
}

static int j_container(JCur* c, JNode* out, int depth, int object) {
    const char close = object ? '}' : ']';
    out->type = object ? J_OBJECT : J_ARRAY;
    c->p++;
    j_ws(c);
    if (c->p < c->end && *c->p == close) {
        c->p++;
        return 1;
    }
    for (;;) {
        JNode kid;
        memset(&kid, 0, sizeof kid);
        j_ws(c);
        if (object) {
            if (!j_string(c)) return 0;
            j_ws(c);
            if (c->p >= c->end || *c->p != ':') return 0;
            c->p++;
        }
        if (!j_value(c, &kid, depth + 1)) {
            j_free(&kid);
            return 0;
        }
        if (!j_push(c, out, &kid)) {
            j_free(&kid);
            return 0;
        }
        j_ws(c);
        if (c->p < c->end && *c->p == ',') {
            c->p++;
            continue;
        }
        if (c->p < c->end && *c->p == close) {
            c->p++;
            return 1;
        }
        return 0;
    }
}

static int j_value(JCur* c, JNode* out, int depth) {
    memset(out, 0, sizeof *out);
    if (depth > J_MAX_DEPTH) return 0;
    j_ws(c);
    if (c->p >= c->end) return 0;
    switch (*c->p) {
        case '[': return j_container(c, out, depth, 0);
        case '{': return j_container(c, out, depth, 1);
        case '"': out->type = J_OTHER; return j_string(c);
        case 't': out->type = J_OTHER; return j_lit(c, "true");
        case 'f': out->type = J_OTHER; return j_lit(c, "false");
        case 'n': out->type = J_OTHER; return j_lit(c, "null");
        default: out->type = J_NUMBER; return j_number(c, &out->num);
    }
}

/* ------------------------------------------------------------- response */

typedef struct {
    char* s;
    size_t n, cap;
    int oom;
} Buf;

static void b_put(Buf* b, const char* s, size_t n) {
    if (b->oom) return;
    if (b->n + n + 1 > b->cap) {
        size_t nc = b->cap ? b->cap : 256;
        char* t;
        while (nc < b->n + n + 1) nc *= 2;
        t = (char*)realloc(b->s, nc);
        if (!t) {
            b->oom = 1;
            return;
        }
        b->s = t;
        b->cap = nc;
    }
    memcpy(b->s + b->n, s, n);
    b->n += n;
    b->s[b->n] = '\0';
}

static void b_str(Buf* b, const char* s) { b_put(b, s, strlen(s)); }

/* A number as cJSON 1.7 prints it: integral values in int range with %d,
 * NaN and infinities as null, otherwise %1.15g unless that does not read
 * back within cJSON's comparison, then %1.17g. */
static void b_num(Buf* b, double d) {
    char t[64];
    int valueint;
    if (d != d || d - d != 0.0) {
        b_str(b, "null");
        return;
    }
    valueint = d >= (double)INT32_MAX ? INT32_MAX : d <= (double)INT32_MIN ? INT32_MIN : (int)d;
    if (d == (double)valueint) {
        snprintf(t, sizeof t, "%d", valueint);
    } else {
        double back = 0.0;
        snprintf(t, sizeof t, "%1.15g", d);
        if (sscanf(t, "%lg", &back) != 1 ||
            !(fabs(back - d) <= (fabs(back) > fabs(d) ? fabs(back) : fabs(d)) * DBL_EPSILON))
            snprintf(t, sizeof t, "%1.17g", d);
    }
    b_str(b, t);
}

/* A string as cJSON escapes it. */
static void b_jstr(Buf* b, const char* s, size_t n) {
    size_t i;
    b_put(b, "\"", 1);
    for (i = 0; i < n && s[i]; i++) {
        const unsigned char ch = (unsigned char)s[i];
        char t[8];
        switch (ch) {
            case '"': b_put(b, "\\\"", 2); break;
            case '\\': b_put(b, "\\\\", 2); break;
            case '\b': b_put(b, "\\b", 2); break;
            case '\f': b_put(b, "\\f", 2); break;
            case '\n': b_put(b, "\\n", 2); break;
            case '\r': b_put(b, "\\r", 2); break;
            case '\t': b_put(b, "\\t", 2); break;
            default:
                if (ch < 32) {
                    snprintf(t, sizeof t, "\\u%04x", ch);
                    b_str(b, t);
                } else {
                    b_put(b, (const char*)&ch, 1);
                }
        }
    }
    b_put(b, "\"", 1);
}

static int finish(Buf* b, int status, char** response, size_t* response_len) {
    if (b->oom) {
        static const char oom[] = "{\"error\": \"Memory allocation failed\"}";
        free(b->s);
        b->s = (char*)malloc(sizeof oom);
        if (b->s) memcpy(b->s, oom, sizeof oom);
        b->n = b->s ? sizeof oom - 1 : 0;
        status = 500;
    }
    *response = b->s;
    if (response_len) *response_len = b->n;
    return status;
}

static int error_response(int status, const char* text, char** response, size_t* response_len) {
    Buf b = {NULL, 0, 0, 0};
    b_str(&b, "{\"error\": ");
    b_jstr(&b, text, strlen(text));
    b_str(&b, "}");
    return finish(&b, status, response, response_len);
}

/* ------------------------------------------------------------ parameters */

static int parse_number_param(const char* s, size_t* k) {
    char* end;
    unsigned long v;
    if (!s) {
        *k = 1;
        return 1;
    }
    if (*s < '0' || *s > '9') return 0;
    v = strtoul(s, &end, 10);
    if (*end != '\0' || v < 1 || v > SVDB_MAX_K) return 0;
    *k = (size_t)v;
    return 1;
}

static int parse_metric_param(const char* s, int* metric, const char** name) {
    static const char* names[3] = {"euclidean_distance", "cosine_similarity", "dot_product"};
    if (!s || !strcmp(s, "euclidean") || !strcmp(s, "euclidean_distance") || !strcmp(s, "l2"))
        *metric = SVDB_METRIC_L2;
    else if (!strcmp(s, "cosine") || !strcmp(s, "cosine_similarity") || !strcmp(s, "cos"))
        *metric = SVDB_METRIC_COS;
    else if (!strcmp(s, "dot") || !strcmp(s, "dot_product"))
        *metric = SVDB_METRIC_DOT;
    else
        return 0;
    *name = names[*metric];
    return 1;
}

/* ----------------------------------------------------------------- entry */

/* One result row: {"index":i[,"<metric>":s],"vector":[...],"uuid":"..."}. */
static void put_row(Buf* b, VectorDatabase* db, size_t id, const char* score_key, double score) {
    Vector* v = vector_db_read(db, id);
    size_t i;
    if (!v) {
        b_str(b, "{\"error\":\"Nearest neighbor not found\"}");
        return;
    }
    b_str(b, "{\"index\":");
    b_num(b, (double)id);
    if (score_key) {
        b_str(b, ",\"");
        b_str(b, score_key);
        b_str(b, "\":");
        b_num(b, score);
    }
    b_str(b, ",\"vector\":[");
    for (i = 0; i < v->dimension; i++) {
        if (i) b_put(b, ",", 1);
        b_num(b, v->data ? v->data[i] : 0.0);
    }
    b_str(b, "],\"uuid\":");
    b_jstr(b, v->uuid, UUID_SIZE);
    b_str(b, "}");
}

int svdb_nearest_request(VectorDatabase* db, svdb_server* srv, size_t expected_dim,
                         const char* body, size_t len, const char* number, const char* metric,
                         int precision, char** response, size_t* response_len) {
    JCur cur;
    JNode top;
    const JNode* rows;   /* the query arrays: &top (one query) or top.kids (a batch) */
    size_t nq, q, i, k = 1, D, j;
    int batch, m = SVDB_METRIC_L2, rc, status;
    const char* mname = NULL;
    double *Q = NULL, *sc = NULL;
    size_t *ids = NULL, *ks = NULL;
    int* ms = NULL;
    Buf b = {NULL, 0, 0, 0};

    *response = NULL;
    if (response_len) *response_len = 0;
    /* compare_handler.c:287-297 */
    if (!body || len == 0) return error_response(400, "Empty data", response, response_len);
    /* compare_handler.c:300-313 */
    cur.p = body;
    cur.end = body + len;
    cur.oom = 0;
    if (!j_value(&cur, &top, 0)) {
        j_free(&top);
        return cur.oom ? error_response(500, "Memory allocation failed", response, response_len)
                       : error_response(400, "Invalid JSON", response, response_len);
    }
    batch = top.type == J_ARRAY && top.n > 0 && top.kids[0].type == J_ARRAY;
    rows = batch ? top.kids : &top;
    nq = batch ? top.n : 1;
    /* compare_handler.c:316-340 (size) and 369-386 (elements), per query */
    for (q = 0; q < nq; q++) {
        const JNode* r = &rows[q];
        const size_t n = (r->type == J_ARRAY || r->type == J_OBJECT) ? r->n : 0;
        if (batch && r->type != J_ARRAY) {
            j_free(&top);
            return error_response(400, "Invalid vector data", response, response_len);
        }
        if (n != expected_dim) {
            j_free(&top);
            return error_response(400, "Vector size mismatch", response, response_len);
        }
        for (i = 0; i < n; i++)
            if (r->kids[i].type != J_NUMBER) {
                j_free(&top);
                return error_response(400, "Invalid vector data", response, response_len);
            }
    }
    if (!parse_number_param(number, &k)) {
        j_free(&top);
        return error_response(400, "Invalid number", response, response_len);
    }
    if (!parse_metric_param(metric, &m, &mname)) {
        j_free(&top);
        return error_response(400, "Unknown metric", response, response_len);
    }
    if (!db) {
        j_free(&top);
        return error_response(500, "Database not initialised", response, response_len);
    }
    /* Euclidean: the kd-tree's leading coordinates (kdtree.c:134-137) */
    D = expected_dim;
    if (m == SVDB_METRIC_L2 && db->kdtree && db->kdtree->dimension > 0 &&
        db->kdtree->dimension < D)
        D = db->kdtree->dimension;
    Q = (double*)malloc((nq * D > 0 ? nq * D : 1) * sizeof(double));
    ids = (size_t*)malloc(nq * k * sizeof(size_t));
    sc = (double*)malloc(nq * k * sizeof(double));
    ks = (size_t*)malloc(nq * sizeof(size_t));
    ms = (int*)malloc(nq * sizeof(int));
    if (!Q || !ids || !sc || !ks || !ms) {
        status = error_response(500, "Memory allocation failed", response, response_len);
        goto out;
    }
    for (q = 0; q < nq; q++) {
        for (i = 0; i < D; i++) Q[q * D + i] = rows[q].kids[i].num;
        ks[q] = k;
        ms[q] = m;
    }
    j_free(&top);
    /* the serving loop compares whole rows; kd-tree projections go direct */
    if (srv && D == expected_dim) {
        int64_t t = svdb_server_submit_search(srv, Q, nq, D, k, m);
        rc = t < 0 ? (int)t : svdb_server_wait(srv, t, ids, sc, NULL);
    } else {
        rc = svdb_search(db, Q, nq, D, ks, ms, precision, k, ids, sc);
    }
    if (rc == SVDB_EEMPTY) {
        for (j = 0; j < nq * k; j++) ids[j] = (size_t)-1;
        rc = SVDB_OK;
    }
    if (rc != SVDB_OK) {
        status = error_response(500, svdb_last_error(), response, response_len);
        goto out;
    }
    if (batch) b_put(&b, "[", 1);
    for (q = 0; q < nq; q++) {
        const size_t* qi = ids + q * k;
        if (q) b_put(&b, ",", 1);
        if (k == 1) {
            /* compare_handler.c:402-420 */
            if (qi[0] == (size_t)-1)
                b_str(&b, "{\"error\":\"No nearest neighbor found\"}");
            else
                put_row(&b, db, qi[0], metric ? mname : NULL, sc[q * k]);
        } else {
            int first = 1;
            b_put(&b, "[", 1);
            for (j = 0; j < k; j++) {
                if (qi[j] == (size_t)-1) continue;
                if (!first) b_put(&b, ",", 1);
                first = 0;
                put_row(&b, db, qi[j], mname, sc[q * k + j]);
            }
            b_put(&b, "]", 1);
        }
    }
    if (batch) b_put(&b, "]", 1);
    status = finish(&b, 200, response, response_len);
    b.s = NULL;
out:
    j_free(&top);
    free(b.s);
    free(Q);
    free(ids);
    free(sc);
    free(ks);
    free(ms);
    return status;
}
