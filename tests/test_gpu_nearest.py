"""POST /nearest through svdb_nearest_request on the GPU: the answers the
reference's handler gives (compare_handler.c:263-441: kdtree_nearest over
the kd-tree's leading coordinates, then vector_db_read), and the top-k,
metric and batch forms it adds, checked against the FP64 oracle — directly
and through the serving loop."""
import json

import numpy as np
import pytest

import svdb
import oracle as O

pytestmark = pytest.mark.gpu

N, D, KD = 20_000, 64, 3


@pytest.fixture(scope="module")
def store():
    rng = np.random.default_rng(21)
    X = rng.standard_normal((N, D))
    db = svdb.Database(kd_dimension=KD)
    db.bulk_insert(X, uuids=[f"id-{i}" for i in range(N)])
    yield db, X, rng
    db.close()


def body(Q):
    Q = np.atleast_2d(Q)
    rows = ["[" + ",".join(repr(float(v)) for v in q) + "]" for q in Q]
    return rows[0] if len(rows) == 1 else "[" + ",".join(rows) + "]"


def close(a, b):
    return abs(a - b) <= 2.3e-16 * abs(b)


def test_single_query_is_the_reference_object(store):
    db, X, rng = store
    q = rng.standard_normal(D)
    st, text = svdb.nearest_request(db, body(q), D)
    assert st == 200
    ans = json.loads(text)
    assert list(ans) == ["index", "vector", "uuid"]  # compare_handler.c:413-415
    rid, _ = O.topk(X, q, 1, "l2", D=KD)  # the kd-tree's projection
    assert ans["index"] == int(rid[0, 0]) == db.kd_nearest(q)
    assert ans["uuid"] == f"id-{ans['index']}"
    np.testing.assert_allclose(ans["vector"], X[ans["index"]], rtol=2.3e-16, atol=0)


@pytest.mark.parametrize("metric,name", [("euclidean", "euclidean_distance"),
                                         ("cosine", "cosine_similarity"), ("dot", "dot_product")])
@pytest.mark.parametrize("k", [1, 7])
def test_number_and_metric(store, metric, name, k):
    db, X, rng = store
    q = rng.standard_normal(D)
    st, text = svdb.nearest_request(db, body(q), D, number=str(k), metric=metric)
    assert st == 200
    ans = json.loads(text)
    m = {"euclidean": "l2", "cosine": "cos", "dot": "dot"}[metric]
    rid, rsc = O.topk(X, q, k, m, D=KD if m == "l2" else None)
    rows = [ans] if k == 1 else ans
    assert [r["index"] for r in rows] == [int(v) for v in rid[0]]
    for j, r in enumerate(rows):
        assert close(r[name], rsc[0, j])


def test_batch_body_and_server(store):
    db, X, rng = store
    Q = rng.standard_normal((9, D))
    rid, rsc = O.topk(X, Q, 5, "cos")
    st, text = svdb.nearest_request(db, body(Q), D, number="5", metric="cosine_similarity")
    assert st == 200
    direct = json.loads(text)
    srv = svdb.Server(db, max_batch=1024)
    try:
        st2, text2 = svdb.nearest_request(db, body(Q), D, number="5", metric="cosine",
                                          server=srv)
    finally:
        srv.stop()
    assert st2 == 200 and text2 == text
    assert len(direct) == 9
    for qi in range(9):
        assert [r["index"] for r in direct[qi]] == [int(v) for v in rid[qi]]
        assert all(close(r["cosine_similarity"], rsc[qi, j]) for j, r in enumerate(direct[qi]))


def test_more_neighbours_than_rows_and_empty_store():
    rng = np.random.default_rng(22)
    X = rng.standard_normal((3, 4))
    db = svdb.Database(kd_dimension=4)
    st, text = svdb.nearest_request(db, "[1,2,3,4]", 4)
    assert (st, json.loads(text)) == (200, {"error": "No nearest neighbor found"})
    db.bulk_insert(X)
    st, text = svdb.nearest_request(db, "[1,2,3,4]", 4, number="10", metric="dot")
    ans = json.loads(text)
    rid, _ = O.topk(X, np.array([1.0, 2, 3, 4]), 3, "dot")
    assert st == 200 and [r["index"] for r in ans] == [int(v) for v in rid[0]]
    db.close()
