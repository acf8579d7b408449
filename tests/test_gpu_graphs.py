"""Graph-captured exact searches, through the C ABI, against the CPU oracle.

With SVDB_GRAPHS=1, a host-query exact search whose shape repeats (a
handler calling kdtree_nearest once per request, or any FP64 search of the
same size, k and metric) runs from the second time on as one launch of a HIP
graph captured from its launch sequence.  Every replay here is checked bit
for bit against the oracle on the store as it is at that moment, with eager
work, host synchronisations, store mutations and scratch growth between the
replays — the things that could leave a replay running stale launches.
"""
import sys

import numpy as np
import pytest

import oracle as O
import svdb

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def graphs_on(monkeypatch):
    # graph replay is opt-in; the engine reads the switch when a store is made
    monkeypatch.setenv("SVDB_GRAPHS", "1")
    yield


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    a = np.array(a, dtype=np.float64)
    a[np.isnan(a)] = np.nan
    return a.view(np.uint64)


def test_kd_nearest_replays_follow_the_store():
    rng = np.random.default_rng(41)
    X = rng.standard_normal((20000, 64))
    say("graphs: kdtree_nearest at kd dimension 3 over 20000x64, 40 calls with mutations")
    db = svdb.Database(kd_dimension=3)
    db.bulk_insert(X)
    rows = [r for r in X]
    s0 = db.stats()
    for i in range(40):
        q = rng.standard_normal(64)
        got = db.kd_nearest(q)
        want, _ = O.topk(np.array(rows), q[None, :], 1, "l2", D=3)
        assert got == int(want[0, 0]), f"call {i}"
        if i % 7 == 3:
            # eager work of another shape and a host sync between replays
            Q2 = rng.standard_normal((5, 64))
            ids, sc = db.search(Q2, 4, "dot", "fp64")
            rid, rsc = O.topk(np.array(rows), Q2, 4, "dot")
            assert np.array_equal(ids, rid) and np.array_equal(bits(sc), bits(rsc))
        if i % 10 == 5:
            # a mutation: the next call must see it (a row planted on the query)
            q_next = rng.standard_normal(64)
            db.insert(q_next)
            rows.append(q_next)
            assert db.kd_nearest(q_next) == len(rows) - 1
        if i == 22:
            db.delete(0)
            rows.pop(0)
        if i == 30:
            db.update(5, rows[7])
            rows[5] = rows[7].copy()
    s1 = db.stats()
    say(f"  graph captures {s1['graph_captures'] - s0['graph_captures']}, "
        f"launches {s1['graph_launches'] - s0['graph_launches']}")
    assert s1["graph_launches"] - s0["graph_launches"] >= 20
    db.close()


@pytest.mark.parametrize("metric", ["l2", "cos", "dot"])
def test_repeated_fp64_batches_replay_exactly(metric):
    rng = np.random.default_rng(43)
    X = rng.standard_normal((30000, 32))
    say(f"graphs: repeated 3-query FP64 batches, {metric}")
    db = svdb.Database(kd_dimension=32)
    db.bulk_insert(X)
    s0 = db.stats()
    for i in range(10):
        Q = rng.standard_normal((3, 32))
        ids, sc = db.search(Q, 5, metric, "fp64")
        rid, rsc = O.topk(X, Q, 5, metric)
        assert np.array_equal(ids, rid), f"replay {i}"
        assert np.array_equal(bits(sc), bits(rsc)), f"replay {i}"
        if i == 4:
            # scratch growth between replays: a large batch re-allocates the
            # search buffers, which must retire the graphs that used them
            Qb = rng.standard_normal((3000, 32))
            ids, sc = db.search(Qb, 20, metric, "fp64")
            rid, rsc = O.topk(X, Qb[:50], 20, metric)
            assert np.array_equal(ids[:50], rid) and np.array_equal(bits(sc[:50]), bits(rsc))
    s1 = db.stats()
    assert s1["graph_launches"] - s0["graph_launches"] >= 5
    db.close()
