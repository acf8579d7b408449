"""The FP64 scan on the matrix cores (widths that are multiples of 32),
through the C ABI, bit for bit against the CPU oracle.

The matrix-core scan computes dot products with v_mfma_f64_16x16x4f64
chained in column order (the canonical fma chain), cosine from them, and
Euclidean keys through a bound on (xx + qq) − 2·dot that sends every pair
that could still enter a list to the canonical fma chain.  These cases aim
at what that design could get wrong: non-finite and overflowing values
(the bound turns non-finite), subnormal products, exact ties and zero
distances, large offsets where (xx + qq) − 2·dot cancels almost every bit,
deleted rows, stores shorter than one 64-row stage, partial query groups,
and k on both sides of 64 (one or two list registers per lane).
"""
import sys

import numpy as np
import pytest

import oracle as O
import svdb

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    a = np.array(a, dtype=np.float64)
    a[np.isnan(a)] = np.nan
    return a.view(np.uint64)


def check(db, X, Q, k, metric, alive=None, what=""):
    ids, sc = db.search(Q, k, metric, "fp64")
    rid, rsc = O.topk(X, Q, k, metric, alive=alive)
    assert np.array_equal(ids, rid), f"{what} {metric} k={k}: ids differ at {np.argwhere(ids != rid)[:5]}"
    assert np.array_equal(bits(sc), bits(rsc)), f"{what} {metric} k={k}: scores differ"


def edge_store(D, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((5000, D))
    X[100:140] = X[0:40]                      # exact duplicates: ties broken by id
    X[200:220] *= 1e155                       # squares overflow: the bound is infinite
    X[300:320] *= 1e-160                      # products are subnormal
    X[400, 3] = np.nan
    X[401, 0] = np.inf
    X[402, D - 1] = -np.inf
    X[403] = 0.0
    X[404] = -0.0
    X[405:410] = 1e-300                       # tiny, all equal
    X[500:510] = X[500] + rng.standard_normal((10, D)) * 1e-12   # near-ties
    Q = rng.standard_normal((70, D))          # 70 queries: a partial group of 64
    Q[1] = X[0]                               # distance 0 to rows 0 and 100
    Q[2] = X[500]
    Q[3] = 0.0
    Q[4, 5] = np.nan
    Q[5, 1] = np.inf
    Q[6] *= 1e155
    Q[7] *= 1e-160
    return X, Q


@pytest.mark.parametrize("D", [32, 96])
def test_matrix_core_edge_values(D):
    X, Q = edge_store(D, 5 + D)
    say(f"matrix-core scan, edge values, 5000x{D}, 70 queries")
    db = svdb.Database(kd_dimension=D)
    db.bulk_insert(X)
    for metric in ("l2", "cos", "dot"):
        for k in (1, 7, 32, 33, 100):
            check(db, X, Q, k, metric, what=f"D={D}")
    db.close()


def test_matrix_core_large_offset_near_ties():
    # rows c + ε·z with ‖c‖ ≈ 1e6 and ε = 1e-3: (xx + qq) − 2·dot loses
    # about 18 of its digits, so nearly every pair near the query needs its
    # canonical key
    rng = np.random.default_rng(21)
    D = 64
    c = rng.standard_normal(D) * 1.25e5
    X = c + rng.standard_normal((20000, D)) * 1e-3
    Q = c + rng.standard_normal((80, D)) * 1e-3
    say("matrix-core scan, large offset near-ties, 20000x64, 80 queries")
    db = svdb.Database(kd_dimension=D)
    db.bulk_insert(X)
    for metric in ("l2", "cos", "dot"):
        for k in (1, 10, 32):
            check(db, X, Q, k, metric, what="offset")
    db.close()


@pytest.mark.parametrize("n", [1, 15, 63, 64, 65, 200])
def test_matrix_core_short_stores_and_deletes(n):
    rng = np.random.default_rng(100 + n)
    D = 32
    X = rng.standard_normal((n, D))
    Q = rng.standard_normal((70, D))
    say(f"matrix-core scan, {n} rows")
    db = svdb.Database(kd_dimension=D)
    db.bulk_insert(X)
    for metric in ("l2", "cos", "dot"):
        for k in (1, 10):
            check(db, X, Q, k, metric, what=f"n={n}")
    if n > 3:
        # delete a few rows: ids are the logical (compacted) positions
        gone = sorted({0, n // 2, n - 1})
        for i in reversed(gone):
            db.delete(i)
        keep = np.ones(n, dtype=bool)
        keep[gone] = False
        Xk = X[keep]
        for metric in ("l2", "cos", "dot"):
            check(db, Xk, Q, 5, metric, what=f"n={n} after deletes")
    db.close()


def test_matrix_core_cosine_underflowing_quotients():
    # rows whose dot product with the query is subnormal or tiny while
    # ‖q‖·‖x‖ is ~1e11: cos = dot / (‖q‖‖x‖) underflows to subnormals or
    # zero, so the k-th best key sits in the subnormal range, where a
    # pre-test that skips the division must still let every candidate in
    rng = np.random.default_rng(77)
    D = 32
    n = 6000
    X = rng.standard_normal((n, D)) * 1e10
    tiny = np.array([1e-300, -1e-300, 1e-305, -1e-305, 1e-310, -1e-310, 5e-324, -5e-324,
                     0.0, -0.0, 2.5e-320, 1e-200, 1e-10])
    X[:, 0] = tiny[rng.integers(0, len(tiny), n)]
    Q = rng.standard_normal((70, D))
    Q[0] = 0.0
    Q[0, 0] = 1.0
    Q[1] = 0.0
    Q[1, 0] = -1.0
    Q[2] = 0.0
    Q[2, 0] = 3.0
    Q[3] = 0.0
    Q[3, 0] = 1e-300
    say("matrix-core scan, cosine with underflowing quotients, 6000x32, 70 queries")
    db = svdb.Database(kd_dimension=D)
    db.bulk_insert(X)
    for k in (1, 10, 40, 64):
        check(db, X, Q, k, "cos", what="underflow")
        check(db, X, Q, k, "dot", what="underflow")
    db.close()
