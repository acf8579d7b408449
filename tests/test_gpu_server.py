"""Serving loop (svdb_server_*): a request answered inside a mixed,
continuously batched tick gets exactly the ids and FP64 scores it gets when
run alone on the same store state, and mutations act as barriers in
submission order.

The requests follow the serving workload's shape: batch sizes log-uniform in
1..512, k in {1, 10, 100}, metric uniform over the three, about 5% of the
requests mutations (insert / update / delete).  The reference serves one
request per connection thread (server.c, compare_handler.c:263-441); the
"alone" replay below is that one-at-a-time order.
"""
import numpy as np
import pytest

import svdb
import workloads
import oracle as O

pytestmark = pytest.mark.gpu

N, D = 60_000, 128


def _requests(rng, n_req, mut_frac=0.05):
    reqs = []
    for _ in range(n_req):
        if rng.random() < mut_frac:
            reqs.append(("mut", ["insert", "update", "delete"][rng.integers(3)]))
        else:
            nq = int(np.exp(rng.uniform(0, np.log(512))))
            reqs.append(("search", max(1, nq), int(rng.choice([1, 10, 100])),
                         ["l2", "cos", "dot"][rng.integers(3)]))
    return reqs


def _store(rows):
    db = svdb.Database(kd_dimension=3)
    db.bulk_insert(rows)
    return db


def _bits(a):
    a = np.asarray(a, dtype=np.float64).copy()
    a[np.isnan(a)] = np.nan
    return a.view(np.uint64)


def test_mixed_ticks_equal_alone():
    rng = np.random.default_rng(7)
    rows = workloads.mixture(N, D, seed=41, n_centers=64)
    pool = workloads.mixture(4096, D, seed=42, n_centers=64)
    reqs = _requests(rng, 120)
    # materialise query blocks and mutation payloads once
    plan = []
    size = N
    for r in reqs:
        if r[0] == "search":
            _, nq, k, m = r
            o = int(rng.integers(0, pool.shape[0] - nq + 1))
            plan.append(("search", pool[o:o + nq].copy(), k, m))
        else:
            kind = r[1]
            if kind == "insert":
                plan.append(("insert", rng.standard_normal(D)))
                size += 1
            elif kind == "update":
                plan.append(("update", int(rng.integers(size)), rng.standard_normal(D)))
            else:
                plan.append(("delete", int(rng.integers(size))))
                size -= 1

    # mixed: everything submitted at once, then waited for
    db = _store(rows)
    srv = svdb.Server(db, max_batch=8192)
    tickets = []
    for p in plan:
        if p[0] == "search":
            tickets.append(srv.submit_search(p[1], p[2], p[3]))
        elif p[0] == "insert":
            tickets.append(srv.submit_insert(p[1]))
        elif p[0] == "update":
            tickets.append(srv.submit_update(p[1], p[2]))
        else:
            tickets.append(srv.submit_delete(p[1]))
    mixed = [srv.wait(t) for t in tickets]
    srv.stop()
    final_mixed = len(db)
    db.close()

    # alone: the same requests one at a time through svdb_search; and every
    # served search against the FP64 oracle on the rows as the mutations
    # before it left them (a list model replayed in submission order)
    db = _store(rows)
    model = rows.copy()
    batched = 0
    for p, (rc, ids, sc, info) in zip(plan, mixed):
        assert rc == 0, p[0]
        if p[0] == "search":
            Q, k, m = p[1], p[2], p[3]
            ids_a, sc_a = db.search(Q, k, m, precision="auto")
            assert np.array_equal(ids, ids_a), (Q.shape, k, m)
            assert np.array_equal(_bits(sc), _bits(sc_a)), (Q.shape, k, m)
            ids_o, sc_o = O.topk(model, Q, k, m)
            assert np.array_equal(ids, ids_o), ("oracle", Q.shape, k, m)
            assert np.array_equal(_bits(sc), _bits(sc_o)), ("oracle", Q.shape, k, m)
            batched += info["tick_queries"] > Q.shape[0]
        elif p[0] == "insert":
            assert db.insert(p[1]) == ids == model.shape[0]
            model = np.vstack([model, p[1]])
        elif p[0] == "update":
            db.update(p[1], p[2])
            model[p[1]] = p[2]
        else:
            db.delete(p[1])
            model = np.delete(model, p[1], axis=0)
    assert len(db) == final_mixed == model.shape[0]
    assert batched > 0  # some requests really shared a tick
    # the final store agrees with the FP64 oracle
    Q = pool[:64]
    ids_a, sc_a = db.search(Q, 10, "l2", precision="auto")
    X = np.stack([db.read(i)[1] for i in range(len(db))])
    ids_o, sc_o = O.topk(X, Q, 10, "l2")
    assert np.array_equal(ids_a, ids_o)
    assert np.array_equal(_bits(sc_a), _bits(sc_o))
    db.close()


def test_ticks_are_ordered_and_timed():
    rows = workloads.gaussian(20_000, D, seed=5)
    db = _store(rows)
    srv = svdb.Server(db)
    q = rows[:3] + 1e-3
    t1 = srv.submit_search(q, 1, "l2")
    t2 = srv.submit_delete(0)
    t3 = srv.submit_search(q, 1, "l2")
    r1, r2, r3 = srv.wait(t1), srv.wait(t2), srv.wait(t3)
    assert r1[0] == r2[0] == r3[0] == 0
    assert list(r1[1][:, 0]) == [0, 1, 2]
    # after deleting row 0 every logical index shifts down by one
    assert list(r3[1][1:, 0]) == [0, 1]
    for r in (r1, r2, r3):
        i = r[3]
        assert i["submit_us"] <= i["start_us"] <= i["done_us"]
    assert r1[3]["tick"] < r2[3]["tick"] < r3[3]["tick"]
    srv.stop()
    db.close()


def test_admission_limits():
    rows = workloads.gaussian(4096, D, seed=6)
    db = _store(rows)
    srv = svdb.Server(db, max_batch=64, budget_bytes=1 << 20)
    with pytest.raises(svdb.SvdbError):
        srv.submit_search(rows[:65], 10, "l2")  # more queries than max_batch
    with pytest.raises(svdb.SvdbError):
        srv.submit_search(rows[:8], 10, "l2")  # more memory than the budget
    with pytest.raises(svdb.SvdbError):
        srv.submit_search(rows[:1, :64], 10, "l2")  # wrong dimension
    # within both limits: served, and ticks never exceed max_batch queries
    ts = [srv.submit_search(rows[i:i + 1], 5, "cos") for i in range(40)]
    out = [srv.wait(t) for t in ts]
    for i, (rc, ids, sc, info) in enumerate(out):
        assert rc == 0 and ids[0, 0] == i
        assert info["tick_queries"] <= 64
    srv.stop()
    db.close()


def test_concurrent_submitters():
    """Four client threads submit at once (searches and mutations); every
    result equals the same request run alone, replayed in ticket order (the
    order the server admitted them)."""
    import threading

    rows = workloads.mixture(40_000, D, seed=43, n_centers=32)
    pool = workloads.mixture(2048, D, seed=44, n_centers=32)
    db = _store(rows)
    srv = svdb.Server(db, max_batch=4096)
    lock = threading.Lock()
    done = []  # (ticket, plan entry, result)
    size = [rows.shape[0]]

    def client(seed):
        rng = np.random.default_rng(seed)
        mine = []
        for _ in range(30):
            if rng.random() < 0.08:
                with lock:  # keep indices valid for whatever order the server sees
                    kind = int(rng.integers(2))
                    if kind == 0:
                        p = ("insert", rng.standard_normal(D))
                        t = srv.submit_insert(p[1])
                        size[0] += 1
                    else:
                        p = ("update", int(rng.integers(size[0])), rng.standard_normal(D))
                        t = srv.submit_update(p[1], p[2])
            else:
                nq = max(1, int(np.exp(rng.uniform(0, np.log(128)))))
                o = int(rng.integers(0, pool.shape[0] - nq + 1))
                p = ("search", pool[o:o + nq].copy(), int(rng.choice([1, 10, 100])),
                     ["l2", "cos", "dot"][rng.integers(3)])
                t = srv.submit_search(p[1], p[2], p[3])
            mine.append((t, p))
        for t, p in mine:
            r = srv.wait(t)
            with lock:
                done.append((t, p, r))

    th = [threading.Thread(target=client, args=(100 + i,)) for i in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.stop()
    db.close()
    assert len(done) == 120
    done.sort(key=lambda e: e[0])
    db = _store(rows)
    shared = 0
    for t, p, (rc, ids, sc, info) in done:
        assert rc == 0, p[0]
        if p[0] == "search":
            ids_a, sc_a = db.search(p[1], p[2], p[3], precision="auto")
            assert np.array_equal(ids, ids_a), (t, p[1].shape, p[2], p[3])
            assert np.array_equal(_bits(sc), _bits(sc_a)), (t, p[2], p[3])
            shared += info["tick_queries"] > p[1].shape[0]
        elif p[0] == "insert":
            assert db.insert(p[1]) == ids
        else:
            db.update(p[1], p[2])
    assert shared > 0
    db.close()


def test_serving_mix_on_the_fp64_scan_store():
    """The serving workload over the FP64-scan workload's store shape:
    1,000,000 x 128 rows i.i.d. N(0,1) (seed 1), Poisson-free burst of 60
    requests with log-uniform 1..512 queries, k in {1, 10, 100}, metrics
    uniform, 5 % mutations; the BF16 path (the server's default).  Each
    request's first queries (up to 4) are checked against the FP64 oracle
    on the rows as they stood when it was admitted, ids and score bits."""
    rng = np.random.default_rng(1)
    rows = rng.standard_normal((1_000_000, D))
    qrng = np.random.default_rng(2)
    reqs = _requests(np.random.default_rng(12), 60)
    plan, size = [], rows.shape[0]
    for r in reqs:
        if r[0] == "search":
            _, nq, k, m = r
            plan.append(("search", qrng.standard_normal((nq, D)), k, m))
        elif r[1] == "insert":
            plan.append(("insert", qrng.standard_normal(D)))
            size += 1
        elif r[1] == "update":
            plan.append(("update", int(qrng.integers(size)), qrng.standard_normal(D)))
        else:
            plan.append(("delete", int(qrng.integers(size))))
            size -= 1
    print("serving mix on 1M x 128 N(0,1):", sum(p[0] == "search" for p in plan), "searches,",
          sum(p[0] != "search" for p in plan), "mutations", flush=True)
    db = _store(rows)
    srv = svdb.Server(db)
    tickets = []
    for p in plan:
        if p[0] == "search":
            tickets.append(srv.submit_search(p[1], p[2], p[3]))
        elif p[0] == "insert":
            tickets.append(srv.submit_insert(p[1]))
        elif p[0] == "update":
            tickets.append(srv.submit_update(p[1], p[2]))
        else:
            tickets.append(srv.submit_delete(p[1]))
    out = [srv.wait(t) for t in tickets]
    srv.stop()
    stats = db.stats()
    db.close()
    model = rows
    checked = 0
    for p, (rc, ids, sc, info) in zip(plan, out):
        assert rc == 0, p[0]
        if p[0] == "search":
            Q, k, m = p[1], p[2], p[3]
            n = min(4, Q.shape[0])
            ids_o, sc_o = O.topk(model, Q[:n], k, m)
            assert np.array_equal(ids[:n], ids_o), (Q.shape, k, m)
            assert np.array_equal(_bits(sc[:n]), _bits(sc_o)), (Q.shape, k, m)
            checked += n
        elif p[0] == "insert":
            assert ids == model.shape[0]
            model = np.vstack([model, p[1]])
        elif p[0] == "update":
            model = model.copy() if model is rows else model
            model[p[1]] = p[2]
        else:
            model = np.delete(model, p[1], axis=0)
    assert checked >= 100
    assert stats["certified"] > 0
