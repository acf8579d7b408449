"""The allocation-free entry point (svdb_search_device_noalloc), through the
C ABI: its answers equal svdb_search_device's FP64 scan, bit for bit, on two
calls in a row (the second smaller, so the scratch array holds the first
call's leftovers), and sizes past its scratch are refused with an error."""
import numpy as np
import pytest

import svdb
import oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(db, dq, nq, ldq, k, metric, dids, dsc):
    return svdb.lib().svdb_search_device_noalloc(db.p, dq, nq, ldq, k, svdb.METRICS[metric],
                                                 dids, dsc)


@pytest.mark.parametrize("metric", ["l2", "cos", "dot"])
def test_noalloc_equals_the_allocating_entry_point(metric):
    rng = np.random.default_rng(71)
    N, d = 200_000, 96
    X = rng.standard_normal((N, d))
    db = svdb.Database(kd_dimension=3)
    db.bulk_insert(X)
    db.delete(5)  # tombstones: logical ids differ from slots
    X = np.delete(X, 5, axis=0)
    for nq, k in ((700, 32), (37, 10)):  # the second call smaller
        Q = rng.standard_normal((nq, d))
        dQ = svdb.DeviceArray.from_host(Q)
        ids_a = svdb.DeviceArray((nq, k), np.uint64)
        sc_a = svdb.DeviceArray((nq, k), np.float64)
        ids_b = svdb.DeviceArray((nq, k), np.uint64)
        sc_b = svdb.DeviceArray((nq, k), np.float64)
        db.search_device(dQ.ptr, nq, d, k, metric, "fp64", ids_a.ptr, sc_a.ptr)
        rc = _noalloc(db, dQ.ptr, nq, d, k, metric, ids_b.ptr, sc_b.ptr)
        assert rc == 0, svdb.lib().svdb_last_error()
        ia, sa, ib, sb = ids_a.to_host(), sc_a.to_host(), ids_b.to_host(), sc_b.to_host()
        assert np.array_equal(ia, ib), (nq, k)
        assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64)), (nq, k)
        rid, rsc = O.topk(X, Q[:8], k, metric)
        assert np.array_equal(ib[:8], rid) and np.array_equal(sb[:8].view(np.uint64),
                                                             rsc.view(np.uint64))
        for a in (dQ, ids_a, sc_a, ids_b, sc_b):
            a.free()
    db.close()


def test_noalloc_refuses_what_its_scratch_cannot_hold():
    rng = np.random.default_rng(72)
    X = rng.standard_normal((50_000, 32))
    db = svdb.Database(kd_dimension=3)
    db.bulk_insert(X)
    nq, k = 8192, 128  # 8192 x 32 lists x 128 entries x 12 B: far past 64 MiB
    dQ = svdb.DeviceArray((nq, 32), np.float64)
    ids = svdb.DeviceArray((nq, k), np.uint64)
    sc = svdb.DeviceArray((nq, k), np.float64)
    rc = _noalloc(db, dQ.ptr, nq, 32, k, "l2", ids.ptr, sc.ptr)
    assert rc == -1
    assert b"scratch" in svdb.lib().svdb_last_error()
    for a in (dQ, ids, sc):
        a.free()
    db.close()
