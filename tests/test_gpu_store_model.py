"""The store behind the C ABI against a plain Python list model, driven by
hypothesis: random sequences of insert, bulk insert, update (new row and
uuid), delete (positions shift down, as vector_database.c:185-195 does),
read, read-by-uuid (the first row with that uuid, vector_database.c:145-158)
and exhaustive top-k search, both the FP64 scan and the certified BF16 path.
After every step the store must equal the model row for row, and every
search must return the oracle's ids and FP64 scores bit for bit.
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, settings
from hypothesis import strategies as st
from hypothesis.stateful import (RuleBasedStateMachine, invariant, precondition, rule,
                                 run_state_machine_as_test)

import oracle as O
import svdb

pytestmark = pytest.mark.gpu

D = 16
# operations run, over all examples (checked after the run)
RAN = {"insert": 0, "bulk": 0, "update": 0, "delete": 0, "search": 0}
UUIDS = [b"", b"a", b"b", b"c", b"dup-uuid-0123456789"]


def bits(a):
    a = np.array(a, dtype=np.float64)
    a[np.isnan(a)] = np.nan
    return a.view(np.uint64)


class StoreModel(RuleBasedStateMachine):
    def __init__(self):
        super().__init__()
        self.db = svdb.Database(kd_dimension=3)
        self.rows = []   # model: rows in position order
        self.uuids = []
        self.rng = np.random.default_rng(0)
        self.counter = 0

    def fresh(self, n=1):
        r = self.rng.standard_normal((n, D))
        # small integers now and then, so exact ties between rows occur
        if self.rng.random() < 0.3:
            r = np.round(r)
        return r

    @rule(u=st.sampled_from(UUIDS))
    def insert(self, u):
        RAN["insert"] += 1
        r = self.fresh()[0]
        assert self.db.insert(r, u) == len(self.rows)
        self.rows.append(r)
        self.uuids.append(u)

    @rule(n=st.integers(1, 40), u=st.sampled_from(UUIDS))
    def bulk(self, n, u):
        RAN["bulk"] += 1
        R = self.fresh(n)
        self.db.bulk_insert(R, [u] * n)
        self.rows.extend(R)
        self.uuids.extend([u] * n)

    @precondition(lambda self: self.rows)
    @rule(data=st.data(), u=st.sampled_from(UUIDS))
    def update(self, data, u):
        i = data.draw(st.integers(0, len(self.rows) - 1))
        RAN["update"] += 1
        r = self.fresh()[0]
        self.db.update(i, r, u)
        self.rows[i] = r
        self.uuids[i] = u

    @precondition(lambda self: self.rows)
    @rule(data=st.data())
    def delete(self, data):
        i = data.draw(st.integers(0, len(self.rows) - 1))
        RAN["delete"] += 1
        self.db.delete(i)
        del self.rows[i]
        del self.uuids[i]

    @precondition(lambda self: self.rows)
    @rule(k=st.sampled_from([1, 3, 10]), metric=st.sampled_from(["l2", "cos", "dot"]),
          prec=st.sampled_from(["fp64", "bf16"]), nq=st.integers(1, 5))
    def search(self, k, metric, prec, nq):
        RAN["search"] += 1
        M = np.array(self.rows)
        Q = self.rng.standard_normal((nq, D))
        if self.rng.random() < 0.3:  # queries equal to stored rows
            Q[0] = M[self.rng.integers(len(M))]
        ids, sc = self.db.search(Q, k, metric, prec)
        rid, rsc = O.topk(M, Q, k, metric)
        assert np.array_equal(ids, rid), (metric, prec, ids, rid)
        assert np.array_equal(bits(sc), bits(rsc)), (metric, prec)

    @invariant()
    def same_rows(self):
        assert len(self.db) == len(self.rows)
        for i in {0, len(self.rows) // 2, len(self.rows) - 1} if self.rows else ():
            u, r = self.db.read(i)
            assert u == self.uuids[i] and np.array_equal(r, self.rows[i])
        assert self.db.read(len(self.rows)) is None

    @invariant()
    def uuid_lookup(self):
        for u in UUIDS:
            got = self.db.read_by_uuid(u)
            if u in self.uuids:
                i = self.uuids.index(u)
                assert got is not None and got[0] == u and np.array_equal(got[1], self.rows[i])
            else:
                assert got is None

    def teardown(self):
        self.db.close()


def test_store_matches_list_model():
    for key in RAN:
        RAN[key] = 0
    run_state_machine_as_test(StoreModel, settings=settings(
        max_examples=25, stateful_step_count=30, deadline=None, derandomize=True,
        suppress_health_check=list(HealthCheck)))
    # the state machine really ran its operations
    assert RAN["search"] >= 50 and RAN["delete"] >= 20 and RAN["update"] >= 20, RAN
    assert RAN["insert"] >= 20 and RAN["bulk"] >= 20, RAN
