"""Failure paths and store upkeep, through the C ABI on the GPU.

* Host allocation failures (forced with svdb_debug_fail_host_alloc) come back
  as SVDB_ENOMEM or the reference's sentinels, and leave the store exactly as
  it was: every later search still equals the FP64 oracle.
* The serving loop answers a search only once the host mirror
  (db->vectors, read by vector_db_read as compare_handler.c:403-416 does
  after kdtree_nearest) shows every mutation submitted before it.
* Deleted or rewritten degenerate rows leave the shadow's degenerate list,
  and tombstone compaction rebuilds the shadows, so certification recovers.
* vector_db_load sizes the store by what the file can hold, not by a
  damaged header count.
"""
import os

import numpy as np
import pytest

import svdb
import oracle as O

pytestmark = pytest.mark.gpu

ENOMEM = -2


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def _exact(db, X, Q, k, metric, precision="fp64"):
    ids, sc = db.search(Q, k, metric, precision)
    rid, rsc = O.topk(X, Q, k, metric)
    assert np.array_equal(ids, rid), f"{metric}/{precision}: ids differ from the FP64 oracle"
    assert np.array_equal(_bits(sc), _bits(rsc)), f"{metric}/{precision}: scores differ"


@pytest.fixture(autouse=True)
def _disarm():
    yield
    svdb.fail_host_alloc(0)


def test_search_allocation_failure_then_recovers():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((5000, 64))
    Q = rng.standard_normal((16, 64))
    db = svdb.Database(kd_dimension=64)
    db.bulk_insert(X)
    svdb.fail_host_alloc(1)
    with pytest.raises(svdb.SvdbError) as e:
        db.search(Q, 5, "l2", "fp64")
    assert e.value.code == ENOMEM
    assert svdb.fail_host_alloc(-1) == 0  # fired
    # kdtree_nearest's one-launch path (small store, kd dimension <= 256)
    # allocates no host memory: it answers, and the hook stays armed
    svdb.fail_host_alloc(1)
    rid, _ = O.topk(X, Q[:1], 1, "l2", D=64)
    assert db.kd_nearest(Q[0]) == int(rid[0, 0])
    assert svdb.fail_host_alloc(-1) == 1  # not fired
    svdb.fail_host_alloc(0)
    # on the batched scan (kd dimension above that path's 256 columns) the
    # reference entry point reports its sentinel
    wide = svdb.Database(kd_dimension=300)
    wide.bulk_insert(rng.standard_normal((500, 300)))
    svdb.fail_host_alloc(1)
    assert wide.kd_nearest(rng.standard_normal(300)) == svdb.NOID
    svdb.fail_host_alloc(0)
    wide.close()
    for m in ("l2", "cos", "dot"):
        _exact(db, X, Q, 5, m, "fp64")
        _exact(db, X, Q, 5, m, "bf16")
    svdb.fail_host_alloc(1)
    with pytest.raises(svdb.SvdbError) as e:
        db.pair_metrics([0, 1], [2, 3], "dot")
    assert e.value.code == ENOMEM
    db.close()


def test_growth_allocation_failure_leaves_store_intact():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((1024, 32))
    db = svdb.Database(kd_dimension=32, capacity=1024)
    db.bulk_insert(X)  # the device store is now full
    more = rng.standard_normal((3000, 32))
    svdb.fail_host_alloc(1)  # the store's growth (host bitmap) fails
    with pytest.raises(svdb.SvdbError) as e:
        db.bulk_insert(more)
    assert e.value.code == ENOMEM
    assert len(db) == 1024 and db.stats()["rows"] == 1024
    svdb.fail_host_alloc(1)  # the reference insert reports (size_t)-1
    assert db.insert(more[0]) == svdb.NOID
    assert len(db) == 1024
    svdb.fail_host_alloc(0)
    # growth works afterwards, and searches see exactly the rows stored
    db.bulk_insert(more)
    assert db.insert(more[0]) == 4024
    Xall = np.vstack([X, more, more[:1]])
    Q = rng.standard_normal((8, 32))
    for m in ("l2", "cos", "dot"):
        _exact(db, Xall, Q, 10, m, "fp64")
        _exact(db, Xall, Q, 10, m, "bf16")
    db.close()


def test_server_allocation_failures_fail_only_their_request():
    rng = np.random.default_rng(5)
    X = rng.standard_normal((4000, 32))
    model = [r for r in X]
    db = svdb.Database(kd_dimension=32)
    db.bulk_insert(X)
    srv = svdb.Server(db, max_batch=1024)
    Q = rng.standard_normal((4, 32))
    # a search whose staging fails
    svdb.fail_host_alloc(1)
    t = srv.submit_search(Q, 5, "l2")
    rc, _, _, _ = srv.wait(t)
    assert rc == ENOMEM
    # a delete whose tick cannot be allocated, then one whose mirror node
    # cannot: neither touches the store
    for n in (1, 2):
        svdb.fail_host_alloc(n)
        t = srv.submit_delete(0)
        rc, _, _, _ = srv.wait(t)
        assert rc == ENOMEM
    svdb.fail_host_alloc(0)
    assert len(db) == 4000
    # the loop still serves, and mutations after the failures apply once
    t = srv.submit_delete(0)
    assert srv.wait(t)[0] == 0
    del model[0]
    t = srv.submit_search(Q, 5, "dot")
    rc, ids, sc, _ = srv.wait(t)
    assert rc == 0
    rid, rsc = O.topk(np.array(model), Q, 5, "dot")
    assert np.array_equal(ids, rid) and np.array_equal(_bits(sc), _bits(rsc))
    srv.stop()
    assert len(db) == 3999
    db.close()


def test_served_ids_resolve_through_the_mirror():
    # Deletes at low indices shift every later row of db->vectors (an O(N)
    # move); a search submitted after them must not complete before the
    # mirror has moved, or vector_db_read(id) returns another row.
    rng = np.random.default_rng(6)
    N, d = 300_000, 64
    X = rng.standard_normal((N, d))
    db = svdb.Database(kd_dimension=d)
    db.bulk_insert(X)
    model = X.copy()
    srv = svdb.Server(db, max_batch=1024)
    for rnd in range(6):
        tickets = []
        for _ in range(3):
            i = int(rng.integers(0, 16))
            tickets.append(srv.submit_delete(i))
            model = np.delete(model, i, axis=0)
        row = rng.standard_normal(d)
        tickets.append(srv.submit_insert(row))
        model = np.vstack([model, row])
        Q = rng.standard_normal((4, d))
        metric = ["l2", "cos", "dot"][rnd % 3]
        ts = srv.submit_search(Q, 3, metric)
        rc, ids, sc, _ = srv.wait(ts)  # waited first: the mutations may still be pending
        assert rc == 0
        # every returned id names, in the mirror as it is now, the row scored
        for qi in range(Q.shape[0]):
            for j in range(3):
                _, v = db.read(int(ids[qi, j]))
                if metric == "l2":
                    want = np.sqrt(O.lib().orc_l2_key(O._f64(Q[qi]).ctypes.data,
                                                       O._f64(v).ctypes.data, d))
                elif metric == "dot":
                    want = O.lib().orc_dot_key(O._f64(Q[qi]).ctypes.data,
                                               O._f64(v).ctypes.data, d)
                else:
                    nq = O.lib().orc_sumsq(O._f64(Q[qi]).ctypes.data, d)
                    nx = O.lib().orc_sumsq(O._f64(v).ctypes.data, d)
                    dt = O.lib().orc_dot_key(O._f64(Q[qi]).ctypes.data, O._f64(v).ctypes.data, d)
                    want = O.lib().orc_cos_score(dt, nq, nx)
                assert _bits(want) == _bits(sc[qi, j]), (rnd, qi, j)
        rid, rsc = O.topk(model, Q, 3, metric)
        assert np.array_equal(ids, rid)
        for t in tickets:
            assert srv.wait(t)[0] == 0
    srv.stop()
    db.close()


def test_deleted_degenerate_rows_leave_the_list():
    # More degenerate rows than the list holds (64) switch certification
    # off while they live; once they are deleted (no compaction: far fewer
    # than a quarter of the slots) the list is rebuilt and queries certify.
    rng = np.random.default_rng(8)
    N, d = 20_000, 128
    X = rng.standard_normal((N, d))
    zpos = np.sort(rng.choice(N, 65, replace=False))
    X[zpos] = 0.0
    db = svdb.Database(kd_dimension=d)
    db.bulk_insert(X)
    Q = rng.standard_normal((64, d))
    _exact(db, X, Q, 10, "l2", "bf16")
    s0 = db.stats()
    assert s0["fallbacks"] >= 64  # 65 degenerate rows: nothing certifies
    for i in zpos[::-1]:
        db.delete(int(i))
    live = np.delete(X, zpos, axis=0)
    assert db.stats()["compactions"] == 0
    s1 = db.stats()
    for m in ("l2", "cos", "dot"):
        _exact(db, live, Q, 10, m, "bf16")
    s2 = db.stats()
    assert s2["certified"] - s1["certified"] >= 3 * 64 * 0.9
    db.close()


def test_compaction_keeps_results_exact():
    rng = np.random.default_rng(9)
    N, d = 4000, 96
    X = rng.standard_normal((N, d))
    X[17] = 0.0  # a degenerate row that survives compaction
    db = svdb.Database(kd_dimension=d)
    db.bulk_insert(X)
    db.enable_shadow("fp8")
    model = X.copy()
    orig = np.arange(N)
    Q = rng.standard_normal((32, d))
    # delete a third of the rows from all over the store
    for _ in range(N // 3):
        i = int(rng.integers(0, model.shape[0]))
        if orig[i] == 17:
            continue
        db.delete(i)
        model = np.delete(model, i, axis=0)
        orig = np.delete(orig, i)
    st = db.stats()
    assert st["compactions"] >= 1
    assert st["phys_rows"] < N
    assert len(db) == model.shape[0] == st["rows"]
    for m in ("l2", "cos", "dot"):
        for p in ("fp64", "bf16", "fp8"):
            _exact(db, model, Q, 10, m, p)
    # inserts and updates after compaction
    new = rng.standard_normal((50, d))
    db.bulk_insert(new)
    model = np.vstack([model, new])
    db.update(5, new[0])
    model[5] = new[0]
    for i in range(model.shape[0]):
        _, v = db.read(i)
        assert np.array_equal(v, model[i])
    for m in ("l2", "cos", "dot"):
        _exact(db, model, Q, 10, m, "bf16")
        _exact(db, model, Q, 10, m, "fp8")
    db.close()


def test_load_ignores_an_impossible_header_count(tmp_path):
    rng = np.random.default_rng(10)
    X = rng.standard_normal((3, 8))
    good = tmp_path / "good.bin"
    db = svdb.Database(kd_dimension=8)
    for r in X:
        db.insert(r, b"u")
    db.save(str(good))
    db.close()
    body = good.read_bytes()[8:]
    bad = tmp_path / "bad.bin"
    bad.write_bytes(np.uint64(1 << 62).tobytes() + body)
    # three records where the header claims 2^62: the load reads what is
    # there, finds the file short and fails cleanly (no 2^62-row allocation)
    assert not svdb.lib().vector_db_load(os.fsencode(str(bad)), 8)
    db = svdb.Database.load(str(good), kd_dimension=8)
    assert len(db) == 3
    db.close()
