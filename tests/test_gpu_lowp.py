"""GPU parity of the BF16 MFMA first pass + FP64 rerank, through the C ABI.

Bar: returned ids and FP64 scores bit-identical to the CPU FP64 oracle (the
certification guarantees it; uncertified queries are rescanned in FP64).
Because a broken first pass would still give correct answers through that
fallback, the tests also require most queries to be certified and the
first-pass score error to be at BF16 level (not zero, not large).
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


def assert_same(ids, sc, rid, rsc, what=""):
    bad = np.argwhere(ids != rid)
    assert bad.size == 0, f"{what}: ids differ at {bad[:5].tolist()}"
    assert np.array_equal(bits(sc), bits(rsc)), f"{what}: scores differ"


def cluster_mixture(rng, n, d, centers, sigma=0.05, lognorm=0.5):
    """Rows of the cluster-mixture workload: x = r·u/‖u‖, u = c_j + σz."""
    j = rng.integers(0, centers.shape[0], n)
    u = centers[j] + sigma * rng.standard_normal((n, d))
    return rng.lognormal(0.0, lognorm, (n, 1)) * u / np.linalg.norm(u, axis=1, keepdims=True)


@pytest.fixture(scope="module")
def mix():
    rng = np.random.default_rng(42)
    d = 128
    C = rng.standard_normal((32, d))
    C /= np.linalg.norm(C, axis=1, keepdims=True)
    X = cluster_mixture(rng, 300_000, d, C)
    Q = cluster_mixture(rng, 700, d, C)
    say("mixture store: 300000x128, 32 clusters (~9400 rows each)")
    db = svdb.Database(kd_dimension=d)
    db.bulk_insert(X)
    yield X, Q, db
    db.close()


@pytest.mark.parametrize("metric", ["l2", "cos", "dot"])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_bf16_rerank_matches_oracle(mix, metric, k):
    X, Q, db = mix
    say(f"bf16 {metric} k={k}: 700 queries")
    s0 = db.stats()
    ids, sc = db.search(Q, k, metric, "bf16")
    s1 = db.stats()
    rid, rsc = O.topk(X, Q, k, metric)
    assert_same(ids, sc, rid, rsc, f"{metric} k={k}")
    cert = s1["certified"] - s0["certified"]
    fb = s1["fallbacks"] - s0["fallbacks"]
    say(f"  certified {cert} fallback {fb} max first-pass rel err {s1['last_max_rel_err']:.2e}")
    assert cert + fb == Q.shape[0]
    assert cert >= (0.9 if k < 100 else 0.5) * Q.shape[0]
    assert 1e-7 < s1["last_max_rel_err"] < 5e-2


def test_bf16_gaussian(gauss_lowp=None):
    rng = np.random.default_rng(7)
    X = rng.standard_normal((250_000, 128))
    Q = rng.standard_normal((600, 128))
    db = svdb.Database(kd_dimension=128)
    db.bulk_insert(X)
    for m in ("l2", "cos", "dot"):
        say(f"bf16 gaussian {m}")
        ids, sc = db.search(Q, 10, m, "bf16")
        rid, rsc = O.topk(X, Q, 10, m)
        assert_same(ids, sc, rid, rsc, m)
    st = db.stats()
    assert st["certified"] >= 0.9 * st["queries"]
    db.close()


def _check_mixed(X, Q, ks, ms, ids, sc):
    for m in ("l2", "cos", "dot"):
        for k in sorted(set(int(v) for v in ks)):
            sel = np.flatnonzero((np.asarray(ms) == m) & (np.asarray(ks) == k))
            if sel.size == 0:
                continue
            rid, rsc = O.topk(X, Q[sel], k, m)
            assert_same(ids[sel, :k], sc[sel, :k], rid, rsc, f"mixed {m} k={k}")


def test_auto_mixed_batch(mix):
    """Several metrics in one call: one mixed-metric first pass, every query
    checked against the oracle."""
    X, Q, db = mix
    rng = np.random.default_rng(3)
    ks = rng.choice([1, 10, 100], size=Q.shape[0])
    ms = rng.choice(["l2", "cos", "dot"], size=Q.shape[0])
    say("auto precision, mixed k/metric")
    s0 = db.stats()
    ids, sc = db.search(Q, ks, list(ms), "auto")
    s1 = db.stats()
    _check_mixed(X, Q, ks, ms, ids, sc)
    assert s1["certified"] - s0["certified"] >= 0.9 * Q.shape[0]


def test_mixed_unbalanced_ranges(mix):
    """Metric ranges of 1, 130 and 65 queries (padding between wavefront-
    aligned ranges) with a degenerate query inside one range."""
    X, Q, db = mix
    Qm = np.vstack([Q[:196], np.zeros((1, Q.shape[1]))])
    ms = ["dot"] + ["l2"] * 130 + ["cos"] * 65 + ["l2"]
    ks = np.array([7] * 131 + [3] * 65 + [7])
    say("mixed ranges 1/131/65 with a zero query")
    ids, sc = db.search(Qm, ks, ms, "auto")
    _check_mixed(X, Qm, ks, ms, ids, sc)


@pytest.mark.parametrize("metric", ["l2", "dot"])
def test_single_metric_k_classes(mix, metric):
    """One metric, k on both sides of 32 and the largest k allowed in one
    call (one range at the largest k, each query's own k returned)."""
    X, Q, db = mix
    rng = np.random.default_rng(8)
    ks = rng.choice([1, 5, 33, 50], size=600)
    ks[17] = 128
    say(f"{metric}: k classes <=32 and >32 in one call")
    ids, sc = db.search(Q[:600], ks, metric, "auto")
    _check_mixed(X, Q[:600], ks, [metric] * 600, ids, sc)


def test_bf16_after_mutations_and_degenerate_rows():
    rng = np.random.default_rng(12)
    d = 96  # padded to 128 in the shadow
    rows = [r for r in rng.standard_normal((20_000, d))]
    db = svdb.Database(kd_dimension=d)
    db.bulk_insert(np.array(rows))
    say("bf16 after 400 mutations incl. zero / huge / NaN rows")
    specials = [np.zeros(d), np.full(d, 1e30), np.full(d, np.nan), np.full(d, 1e-30)]
    for step in range(400):
        op = rng.integers(3)
        r = specials[step % 4] if step % 50 == 0 else rng.standard_normal(d)
        if op == 0:
            db.insert(r)
            rows.append(r)
        elif op == 1:
            i = int(rng.integers(len(rows)))
            db.update(i, r)
            rows[i] = r
        else:
            i = int(rng.integers(len(rows)))
            db.delete(i)
            del rows[i]
    M = np.array(rows)
    Q = rng.standard_normal((300, d))
    Q[0] = 0.0  # degenerate query: answered by the FP64 scan
    for m in ("l2", "cos", "dot"):
        ids, sc = db.search(Q, 10, m, "bf16")
        rid, rsc = O.topk(M, Q, 10, m)
        assert_same(ids, sc, rid, rsc, m)
    db.close()


def test_kd_nearest_full_dimension_paths(golden, c1_data):
    # kd dimension == vector size.  A small store (rows x width up to 2^26)
    # answers one query with the FP64 scan; a larger one through the BF16
    # first pass; both must equal the reference's own answers / the oracle.
    X, Q, _ = c1_data
    db = svdb.Database(kd_dimension=128)
    db.bulk_insert(X)
    got = np.array([db.kd_nearest(q) for q in Q], dtype=np.uint64)
    assert np.array_equal(got, golden["c1_kd128_ids"])
    assert db.stats()["certified"] == 0
    db.close()
    rng = np.random.default_rng(17)
    Xl = rng.standard_normal((540_000, 128))  # 69.1M elements, above 2^26
    Ql = rng.standard_normal((40, 128))
    db = svdb.Database(kd_dimension=128)
    db.bulk_insert(Xl)
    got = np.array([db.kd_nearest(q) for q in Ql], dtype=np.uint64)
    rid, _ = O.topk(Xl, Ql, 1, "l2")
    assert np.array_equal(got, rid[:, 0])
    assert db.stats()["certified"] >= 36
    db.close()


def test_device_entry_point(mix):
    X, Q, db = mix
    dq = svdb.DeviceArray.from_host(Q)
    dids = svdb.DeviceArray((Q.shape[0], 10), np.uint64)
    dsc = svdb.DeviceArray((Q.shape[0], 10), np.float64)
    db.search_device(dq.ptr, Q.shape[0], 128, 10, "dot", "bf16", dids.ptr, dsc.ptr)
    svdb.device_sync()
    rid, rsc = O.topk(X, Q, 10, "dot")
    assert_same(dids.to_host(), dsc.to_host(), rid, rsc, "device")


def test_mixed_ranges_retries_with_exact_ties():
    """Widening retries inside a mixed-metric call.  Rows are small integers
    (exact in BF16), each copied ~500 times, more than a 256-wide shortlist
    holds, so a query near a base row cannot be certified before a retry
    covers all its copies.  Each retry starts from the earlier stage's list
    and re-scores only the copies keyed exactly at that stage's largest key
    (all of them here), which must not enter the list twice; each metric
    range keeps its own lists and k.  Every answer equals the oracle's."""
    rng = np.random.default_rng(81)
    d, nb = 128, 400
    base = rng.integers(-8, 9, size=(nb, d)).astype(np.float64)
    X = base[rng.integers(0, nb, size=200_000)]
    db = svdb.Database(kd_dimension=3)
    db.bulk_insert(X)
    nq = 192
    Q = np.vstack([base[rng.integers(0, nb, size=nq - 32)] +
                   0.25 * rng.standard_normal((nq - 32, d)),
                   base[rng.integers(0, nb, size=32)]])
    ks = rng.choice([10, 100], size=nq)
    ms = rng.choice(["l2", "cos", "dot"], size=nq)
    say("bf16 mixed ranges over ~500 copies per row, k in {10, 100}")
    s0 = db.stats()
    ids, sc = db.search(Q, ks, list(ms), "bf16")
    s1 = db.stats()
    _check_mixed(X, Q, ks, ms, ids, sc)
    wid = s1["widened"] - s0["widened"]
    say(f"  widened {wid}, fp64 {s1['fallbacks'] - s0['fallbacks']}")
    assert wid > 0
    db.close()
