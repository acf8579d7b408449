"""GPU parity of the FP8 (OCP e4m3) first pass + FP64 rerank, through the C ABI.

Bar: returned ids and FP64 scores bit-identical to the CPU FP64 oracle on
wide rows (the wide-vector workload's shapes, d = 768 and 1536, at sizes the
oracle checks in seconds) and on the other supported padded widths.  The
certification keeps the ids exact whatever the first pass does, so the tests
also check that the first pass is really FP8: its score error sits at e4m3
level (well above BF16's, well below 1), and that most top-10 queries are
certified without the FP64 rescan.
"""
import sys

import numpy as np
import pytest

import oracle as O
import svdb
import workloads

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


def store(X, fp8=True):
    db = svdb.Database(kd_dimension=3)
    db.bulk_insert(X)
    if fp8:
        db.enable_shadow("fp8")
    return db


@pytest.fixture(scope="module")
def wide768():
    # cluster mixture with LogNormal(0, 1) norms (dot ranking is the hard case)
    X = workloads.mixture(60_000, 768, seed=51, n_centers=64, lognorm=1.0)
    Q = workloads.mixture(96, 768, seed=52, n_centers=64, lognorm=1.0)
    say("fp8 store: 60000x768 mixture, LogNormal(0,1) norms")
    db = store(X)
    yield X, Q, db
    db.close()


@pytest.mark.parametrize("metric", ["l2", "cos", "dot"])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_fp8_exact_768(wide768, metric, k):
    X, Q, db = wide768
    say(f"fp8 768 {metric} k={k}")
    s0 = db.stats()
    ids, sc = db.search(Q, k, metric, precision="fp8")
    s1 = db.stats()
    rid, rsc = O.topk(X, Q, k, metric)
    assert_same(ids, sc, rid, rsc, f"fp8 768 {metric} k={k}")
    cert = s1["certified"] - s0["certified"]
    fb = s1["fallbacks"] - s0["fallbacks"]
    wid = s1["widened"] - s0["widened"]
    sec = s1["second_pass"] - s0["second_pass"]
    say(f"  certified {cert} ({wid} by a wider retry), {sec} handed to the BF16 pass, "
        f"fp64 rescans {fb}, first-pass err {s1['last_max_rel_err']:.3g}")
    assert cert + fb == Q.shape[0]
    if k <= 10:
        assert cert >= 0.5 * Q.shape[0]
    if cert:
        # e4m3 carries 3 mantissa bits: errors of a few percent of a score
        assert 1e-4 < s1["last_max_rel_err"] < 0.5


def test_fp8_1536_and_auto():
    X = workloads.mixture(20_000, 1536, seed=53, n_centers=32, lognorm=1.0)
    Q = workloads.mixture(48, 1536, seed=54, n_centers=32, lognorm=1.0)
    db = store(X)
    for metric in ("l2", "cos", "dot"):
        say(f"fp8 1536 {metric}")
        ids, sc = db.search(Q, 10, metric, precision="fp8")
        rid, rsc = O.topk(X, Q, 10, metric)
        assert_same(ids, sc, rid, rsc, f"fp8 1536 {metric}")
    # the BF16 first pass at padded width 1536 (4 wavefronts per block), and
    # "auto" (BF16 when the store has that shadow)
    for metric in ("l2", "cos", "dot"):
        for prec in ("bf16", "auto"):
            say(f"{prec} 1536 {metric}")
            s0 = db.stats()
            ids, sc = db.search(Q, 10, metric, precision=prec)
            s1 = db.stats()
            rid, rsc = O.topk(X, Q, 10, metric)
            assert_same(ids, sc, rid, rsc, f"{prec} 1536 {metric}")
            assert s1["certified"] - s0["certified"] >= 0.9 * Q.shape[0]
    db.close()


@pytest.mark.parametrize("d", [100, 256, 300, 1000])
def test_fp8_other_widths(d):
    # padded widths 256, 256, 512, 1024 (K steps 2, 2, 4, 8)
    rng = np.random.default_rng(d)
    X = rng.standard_normal((12_000, d))
    Q = rng.standard_normal((40, d))
    db = store(X)
    for metric in ("l2", "dot"):
        say(f"fp8 d={d} {metric}")
        ids, sc = db.search(Q, 5, metric, precision="fp8")
        rid, rsc = O.topk(X, Q, 5, metric)
        assert_same(ids, sc, rid, rsc, f"fp8 d={d} {metric}")
    db.close()


def test_fp8_mutations_and_degenerate_rows():
    rng = np.random.default_rng(9)
    d = 768
    X = workloads.mixture(8_000, d, seed=55, n_centers=16, lognorm=1.0)
    db = store(X)
    rows = [r for r in X]
    # updates, deletes, inserts, and rows the first pass cannot scale
    for i in (5, 77, 4000):
        v = rng.standard_normal(d)
        db.update(i, v)
        rows[i] = v
    for i in (3, 3, 1000):
        db.delete(i)
        del rows[i]
    special = [np.zeros(d), np.full(d, 1e30), np.full(d, 1e-30), rng.standard_normal(d)]
    special[3][7] = np.nan
    for v in special + [rng.standard_normal(d) for _ in range(5)]:
        db.insert(v)
        rows.append(v)
    Xm = np.stack(rows)
    assert len(db) == Xm.shape[0]
    Q = np.vstack([Xm[[10, 200, 7000]] + 1e-3, rng.standard_normal((5, d))])
    for metric in ("l2", "cos", "dot"):
        say(f"fp8 mutations {metric}")
        ids, sc = db.search(Q, 10, metric, precision="fp8")
        rid, rsc = O.topk(Xm, Q, 10, metric)
        assert_same(ids, sc, rid, rsc, f"fp8 mutations {metric}")
    db.close()


def test_fp8_enable_errors():
    db = svdb.Database(kd_dimension=3)
    with pytest.raises(svdb.SvdbError):
        db.enable_shadow("fp8")  # no rows yet
    db.bulk_insert(np.random.default_rng(1).standard_normal((100, 2000)))
    with pytest.raises(svdb.SvdbError):
        db.enable_shadow("fp8")  # wider than 1536
    with pytest.raises(svdb.SvdbError):
        db.search(np.zeros((1, 2000)), 1, "l2", precision="fp8")
    db.close()


@pytest.mark.parametrize("metric", ["l2", "cos", "dot"])
def test_bf16_wide_768(wide768, metric):
    """The BF16 first pass at padded width 768 (k_scan_wide, BF16 operands)."""
    X, Q, db = wide768
    for k in (10, 100):
        say(f"bf16 768 {metric} k={k}")
        s0 = db.stats()
        ids, sc = db.search(Q, k, metric, precision="bf16")
        s1 = db.stats()
        rid, rsc = O.topk(X, Q, k, metric)
        assert_same(ids, sc, rid, rsc, f"bf16 768 {metric} k={k}")
        cert = s1["certified"] - s0["certified"]
        say(f"  certified {cert} of {Q.shape[0]}, first-pass err {s1['last_max_rel_err']:.3g}")
        assert cert >= 0.9 * Q.shape[0]
        assert 1e-7 < s1["last_max_rel_err"] < 5e-2


def test_bf16_wide_512_mutations():
    rng = np.random.default_rng(12)
    d = 400  # padded to 512
    X = workloads.mixture(30_000, d, seed=57, n_centers=16, lognorm=0.5)
    db = store(X, fp8=False)
    rows = [r for r in X]
    for i in (2, 900):
        v = rng.standard_normal(d)
        db.update(i, v)
        rows[i] = v
    db.delete(17)
    del rows[17]
    for v in (np.zeros(d), rng.standard_normal(d)):
        db.insert(v)
        rows.append(v)
    Xm = np.stack(rows)
    Q = np.vstack([Xm[[1, 500]] + 1e-3, workloads.mixture(30, d, seed=58, n_centers=16)])
    for metric in ("l2", "cos", "dot"):
        say(f"bf16 512 mutations {metric}")
        ids, sc = db.search(Q, 10, metric, precision="auto")
        rid, rsc = O.topk(Xm, Q, 10, metric)
        assert_same(ids, sc, rid, rsc, f"bf16 512 {metric}")
    assert db.stats()["certified"] > 0
    db.close()


@pytest.mark.parametrize("n,d", [(200_000, 768), (100_000, 1536)])
def test_fp8_wide_workload_shape(n, d):
    """The wide-vector workload's shape at a size the oracle can check: a
    cluster mixture (1,024 centres, σ 0.05, LogNormal(0, 1) norms), one batch
    of 4,096 queries, top-100, a 128-wide shortlist (the widening retries
    and the BF16 second pass behind it).  All 4,096 answers equal the FP64
    GPU scan's; 256 of them are also checked against the CPU FP64 oracle."""
    X = workloads.mixture(n, d, seed=61, n_centers=1024, lognorm=1.0)
    Q = workloads.mixture(4096, d, seed=62, n_centers=1024, lognorm=1.0)
    db = store(X)
    db.set_shortlist(128)
    sample = np.random.default_rng(63).choice(4096, 256, replace=False)
    for metric in ("dot", "l2", "cos"):
        say(f"fp8 {n}x{d} {metric} k=100 L=128, 4096 queries")
        s0 = db.stats()
        ids, sc = db.search(Q, 100, metric, precision="fp8")
        s1 = db.stats()
        fids, fsc = db.search(Q, 100, metric, precision="fp64")
        assert_same(ids, sc, fids, fsc, f"fp8 vs fp64 scan {d} {metric}")
        rid, rsc = O.topk(X, Q[sample], 100, metric)
        assert_same(ids[sample], sc[sample], rid, rsc, f"fp8 vs oracle {d} {metric}")
        cert = s1["certified"] - s0["certified"]
        wid = s1["widened"] - s0["widened"]
        sec = s1["second_pass"] - s0["second_pass"]
        fb = s1["fallbacks"] - s0["fallbacks"]
        say(f"  certified {cert} (stages counted as they certify: {wid} only after a wider "
            f"retry; {sec} handed to the BF16 pass, counted there again), FP64 rescans {fb}")
    db.close()


def test_fp8_retries_with_exact_ties():
    """Widening retries prune shortlisted rows by lower bounds (the FP8 key's,
    then the BF16 row's) against an earlier list's k-th value.  Rows here are
    small integers, exact in e4m3 (so the keys' rounding slack is all that
    separates a bound from the value), copied ~100 times each: the top 100
    of a query near a base row is its copies, whose shared value is the
    k-th, ties broken by id, and a query whose copies outnumber the
    128-wide shortlist goes on to the wider retries.  Searched before and
    after deletes and updates; every answer equals the oracle's."""
    rng = np.random.default_rng(71)
    d = 768
    nb = 2000
    base = rng.integers(-8, 9, size=(nb, d)).astype(np.float64)
    X = base[rng.integers(0, nb, size=200_000)]
    db = store(X)
    db.set_shortlist(128)
    rows = [r for r in X]
    Q = np.vstack([base[rng.integers(0, nb, size=48)] + 0.25 * rng.standard_normal((48, d)),
                   base[rng.integers(0, nb, size=16)]])
    for phase in ("initial", "mutated"):
        if phase == "mutated":
            for i in (11, 500, 120_000, 199_999):
                db.update(i, base[(i * 7) % nb])
                rows[i] = base[(i * 7) % nb].copy()
            for i in sorted(rng.choice(len(rows), 40, replace=False).tolist(), reverse=True):
                db.delete(i)
                del rows[i]
        Xm = np.stack(rows)
        for metric in ("l2", "cos", "dot"):
            say(f"fp8 ties {phase} {metric} k=100 L=128")
            s0 = db.stats()
            ids, sc = db.search(Q, 100, metric, precision="fp8")
            s1 = db.stats()
            rid, rsc = O.topk(Xm, Q, 100, metric)
            assert_same(ids, sc, rid, rsc, f"fp8 ties {phase} {metric}")
            say(f"  widened {s1['widened'] - s0['widened']}, second pass "
                f"{s1['second_pass'] - s0['second_pass']}, fp64 {s1['fallbacks'] - s0['fallbacks']}")
    db.close()
