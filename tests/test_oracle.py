"""Pins of the CPU oracle (oracle/svdb_oracle.c) — CPU only.

Each check is independent of the oracle's own code: the reference's compiled
outputs (tests/golden, written by gen_golden.py from kdtree.c and
vector_database.c), exact rational arithmetic (fractions.Fraction, one
correctly rounded step at a time), numpy brute force, and closed forms
(duplicates, D=1 order statistics, permutation invariance).
"""
import hashlib
import os
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

U64MAX = np.uint64(2**64 - 1)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def xfma(a, b, c):
    """Correctly rounded a*b+c, evaluated exactly then rounded once."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


# --------------------------------------------------------------------------
# reference kd-tree restatement vs the reference's own answers
# --------------------------------------------------------------------------
def test_c1_inputs_match_fixture(golden, c1_data):
    X, Q, P = c1_data
    assert sha(X) == str(golden["c1_db_sha256"])
    assert sha(Q) == str(golden["c1_q_sha256"])
    assert np.array_equal(P.astype(np.uint64), golden["c1_pairs"])


@pytest.mark.parametrize("D,key", [(3, "c1_kd3_ids"), (128, "c1_kd128_ids")])
def test_kd_restatement_matches_reference(golden, c1_data, D, key):
    X, Q, _ = c1_data
    t = O.KDTree(X, D)
    assert np.array_equal(t.nearest_batch(Q, threads=4), golden[key])


@pytest.mark.parametrize("name,D", [("tie", 3), ("grid", 2), ("dfs", 2)])
def test_kd_restatement_tie_order(golden, name, D):
    t = O.KDTree(golden[name + "_db"], D)
    assert np.array_equal(t.nearest_batch(golden[name + "_q"]), golden[name + "_ref_ids"])


def test_kd_restatement_stale_tree(golden):
    # Reproduce vector_db_update/delete's effect on the tree (a second node
    # for index 1; delete leaves the tree alone) and get the reference's
    # stale answers.
    S = golden["stale_rows"]
    t = O.KDTree(None, 2)
    for i in range(4):
        t.insert(S[i], i)
    t.insert(np.array([100.0, 100.0]), 1)
    assert np.array_equal(t.nearest_batch(golden["stale_q"]), golden["stale_ref_ids"])


def test_kd_is_exact_nn_on_leading_columns(c1_data):
    # Closed form: with distinct distances the kd answer is the argmin of the
    # squared distance over the first D coordinates.
    X, Q, _ = c1_data
    for D in (1, 3, 8):
        t = O.KDTree(X, D)
        got = t.nearest_batch(Q[:64])
        d2 = ((Q[:64, None, :D] - X[None, :, :D]) ** 2).sum(-1)
        assert np.array_equal(got, d2.argmin(1).astype(np.uint64))


# --------------------------------------------------------------------------
# reference float metrics restatement: bit-exact vs the reference
# --------------------------------------------------------------------------
def test_float_metrics_bit_exact(golden, c1_data):
    X, _, P = c1_data
    ref = golden["c1_metric_bits"]
    for col, m in enumerate((O.COS, O.L2, O.DOT)):
        got = O.ref_pairs(X, P[:, 0], P[:, 1], m).view(np.uint32)
        assert np.array_equal(got, ref[:, col]), m


def test_float_metrics_special_inputs(golden):
    A, B = golden["spec_a"], golden["spec_b"]
    ref = golden["spec_metric_bits"]
    AB = np.concatenate([A, B])
    n = A.shape[0]
    for col, m in enumerate((O.COS, O.L2, O.DOT)):
        got = O.ref_pairs(AB, np.arange(n), np.arange(n) + n, m).view(np.uint32)
        assert np.array_equal(got, ref[:, col]), m
    # zero vector → cosine NaN; identical vectors → euclid exactly 0
    f = ref.view(np.float32)
    assert math.isnan(f[0, 0]) and f[2, 1] == 0.0


# --------------------------------------------------------------------------
# canonical FP64 keys: every step correctly rounded in sequence
# --------------------------------------------------------------------------
def exact_chain_l2(q, x):
    s = 0.0
    for a, b in zip(q, x):
        t = b - a  # one IEEE subtraction (Python floats are IEEE doubles)
        s = xfma(t, t, s)
    return s


def exact_chain_dot(q, x):
    s = 0.0
    for a, b in zip(q, x):
        s = xfma(a, b, s)
    return s


@pytest.mark.parametrize("seed", range(5))
def test_canonical_keys_are_sequential_fma_chains(seed):
    rng = np.random.default_rng(100 + seed)
    d = int(rng.integers(1, 200))
    scale = 10.0 ** rng.integers(-3, 4)
    q = rng.standard_normal(d) * scale
    x = rng.standard_normal(d) * scale
    assert O.l2_key(q, x) == exact_chain_l2(q.tolist(), x.tolist())
    assert O.dot_key(q, x) == exact_chain_dot(q.tolist(), x.tolist())
    assert O.sumsq(x) == exact_chain_dot(x.tolist(), x.tolist())
    # and the chain is within the sequential-summation bound of the exact sum
    ex = sum(Fraction(b - a) ** 2 for a, b in zip(q.tolist(), x.tolist()))
    g = d * 2.0 ** -53 / (1 - d * 2.0 ** -53)
    assert abs(Fraction(O.l2_key(q, x)) - ex) <= Fraction(g) * ex


def test_fma_chain_differs_from_unfused():
    # A plausible mistake (mul then add) must be visible: find inputs where
    # the fused chain and the unfused chain disagree, and check the oracle
    # follows the fused one.
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(200):
        q, x = rng.standard_normal(64), rng.standard_normal(64)
        unfused = 0.0
        for a, b in zip(q.tolist(), x.tolist()):
            unfused = unfused + a * b
        fused = exact_chain_dot(q.tolist(), x.tolist())
        if fused != unfused:
            hits += 1
            assert O.dot_key(q, x) == fused
    assert hits > 10


# --------------------------------------------------------------------------
# exhaustive top-k
# --------------------------------------------------------------------------
@pytest.mark.parametrize("metric", [O.L2, O.COS, O.DOT])
def test_topk_vector_form_matches_scalar_form(metric):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((1000, 37))
    Q = rng.standard_normal((17, 37))
    alive = (rng.random(1000) > 0.2).astype(np.uint8)
    for al in (None, alive):
        i1, s1 = O.topk(X, Q, 13, metric, alive=al, threads=3)
        i2, s2 = O.topk(X, Q, 13, metric, alive=al, scalar=True)
        assert np.array_equal(i1, i2)
        assert np.array_equal(s1.view(np.uint64), s2.view(np.uint64))


def test_topk_l2_top1_equals_reference_kd(golden, c1_data):
    # The reference's own find-nearest (kd dimension 128 and 3), rows with
    # distinct distances: the oracle's L2 top-1 must give the same ids.
    X, Q, _ = c1_data
    ids, _ = O.topk(X, Q, 1, O.L2, threads=4)
    assert np.array_equal(ids[:, 0], golden["c1_kd128_ids"])
    ids3, _ = O.topk(X, Q, 1, O.L2, D=3, threads=4)
    assert np.array_equal(ids3[:, 0], golden["c1_kd3_ids"])


@pytest.mark.parametrize("metric", [O.L2, O.COS, O.DOT])
def test_topk_matches_numpy_bruteforce(metric):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((5000, 64))
    Q = rng.standard_normal((20, 64))
    ids, sc = O.topk(X, Q, 10, metric, threads=4)
    if metric == O.L2:
        key = ((Q[:, None, :] - X[None]) ** 2).sum(-1)
        exp_sc = np.sqrt(np.sort(key, 1)[:, :10])
    elif metric == O.DOT:
        key = -(Q @ X.T)
        exp_sc = -np.sort(key, 1)[:, :10]
    else:
        key = -(Q @ X.T) / np.linalg.norm(Q, axis=1)[:, None] / np.linalg.norm(X, axis=1)[None]
        exp_sc = -np.sort(key, 1)[:, :10]
    exp_ids = np.argsort(key, 1, kind="stable")[:, :10]
    assert np.array_equal(ids, exp_ids.astype(np.uint64))
    np.testing.assert_allclose(sc, exp_sc, rtol=1e-12)


def test_topk_ties_go_to_lowest_id(golden):
    # duplicate rows → lowest index; the DFS case where the reference
    # answers 2 → the lowest-id rule answers 1 (documented divergence)
    T, TQ = golden["tie_db"], golden["tie_q"]
    ids, sc = O.topk(T, TQ, 2, O.L2)
    assert ids[0].tolist() == [0, 2] and ids[1].tolist() == [1, 4] and ids[2].tolist() == [3, 5]
    assert sc[0, 0] == 0.0 and sc[0, 1] == 0.0
    ids, _ = O.topk(golden["dfs_db"], golden["dfs_q"], 1, O.L2)
    assert ids[0, 0] == 1 and golden["dfs_ref_ids"][0] == 2
    ids, _ = O.topk(golden["grid_db"], golden["grid_q"], 2, O.L2)
    assert ids[0].tolist() == [1, 3]


def test_topk_nan_ranks_last_and_fill():
    X = np.array([[1.0, 0.0], [np.nan, 0.0], [0.0, 0.0], [0.5, 0.5]])
    q = np.array([[1.0, 0.1]])
    ids, sc = O.topk(X, q, 6, O.COS)
    # row 2 is the zero vector → cosine 0/0 = NaN; row 1 holds a NaN
    assert ids[0, :2].tolist() == [0, 3]
    assert ids[0, 2:4].tolist() == [1, 2] and np.isnan(sc[0, 2:4]).all()
    assert (ids[0, 4:] == U64MAX).all() and np.isnan(sc[0, 4:]).all()
    ids, _ = O.topk(X, q, 4, O.L2)
    assert ids[0, 3] == 1


def test_topk_alive_mask_gives_logical_ids():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((300, 16))
    Q = rng.standard_normal((5, 16))
    alive = np.ones(300, np.uint8)
    alive[[0, 10, 11, 200]] = 0
    ids, sc = O.topk(X, Q, 7, O.DOT, alive=alive)
    ids2, sc2 = O.topk(X[alive == 1], Q, 7, O.DOT)
    assert np.array_equal(ids, ids2) and np.array_equal(sc, sc2)


def test_topk_permutation_invariance():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((2000, 24))
    Q = rng.standard_normal((8, 24))
    perm = rng.permutation(2000)
    ids, sc = O.topk(X, Q, 5, O.L2)
    ids_p, sc_p = O.topk(X[perm], Q, 5, O.L2)
    assert np.array_equal(perm[ids_p.astype(np.int64)], ids.astype(np.int64))
    assert np.array_equal(sc, sc_p)


def test_topk_one_dimension_is_sorted_neighbour():
    rng = np.random.default_rng(6)
    x = np.sort(rng.uniform(-1, 1, 4000))
    xp = rng.permutation(x)
    for qv in rng.uniform(-1.2, 1.2, 40):
        ids, _ = O.topk(xp[:, None], np.array([[qv]]), 1, O.L2)
        j = np.searchsorted(x, qv)
        cands = [c for c in (j - 1, j) if 0 <= c < x.size]
        best = min(cands, key=lambda c: abs(x[c] - qv))
        assert xp[int(ids[0, 0])] == x[best]


def test_topk_cos_self_match():
    rng = np.random.default_rng(8)
    X = rng.standard_normal((500, 32)) * rng.lognormal(0, 1, (500, 1))
    ids, sc = O.topk(X, X[:50], 1, O.COS)
    assert np.array_equal(ids[:, 0], np.arange(50, dtype=np.uint64))
    assert np.all(np.abs(sc[:, 0] - 1.0) <= 4 * 2.0 ** -52)


# --------------------------------------------------------------------------
# canonical FP64 pair metric: strided lanes then xor butterfly
# --------------------------------------------------------------------------
def butterfly(part):
    part = list(part)
    off = 32
    while off >= 1:
        part = [part[l] + part[l ^ off] for l in range(64)]
        off //= 2
    return part[0]


def exact_pair(a, b, metric):
    def lanes(f):
        p = [0.0] * 64
        for i, (u, v) in enumerate(zip(a, b)):
            p[i % 64] = f(u, v, p[i % 64])
        return butterfly(p)
    if metric == O.L2:
        return math.sqrt(lanes(lambda u, v, s: xfma(u - v, u - v, s)))
    d = lanes(lambda u, v, s: xfma(u, v, s))
    if metric == O.DOT:
        return d
    na = lanes(lambda u, v, s: xfma(u, u, s))
    nb = lanes(lambda u, v, s: xfma(v, v, s))
    return d / (math.sqrt(na) * math.sqrt(nb))


@pytest.mark.parametrize("d", [1, 5, 64, 65, 128, 300])
@pytest.mark.parametrize("metric", [O.L2, O.COS, O.DOT])
def test_pair_fp64_order(d, metric):
    rng = np.random.default_rng(d)
    a, b = rng.standard_normal(d), rng.standard_normal(d)
    assert O.pair_fp64(a, b, metric) == exact_pair(a.tolist(), b.tolist(), metric)
    ex = {O.DOT: float(sum(Fraction(u) * Fraction(v) for u, v in zip(a.tolist(), b.tolist())))}
    if metric == O.DOT:
        assert abs(O.pair_fp64(a, b, metric) - ex[O.DOT]) <= 1e-12 * np.abs(a * b).sum()


def test_pair_fp64_small_d_is_pairwise_tree_of_products():
    # D <= 64: each lane holds one correctly rounded product; the butterfly
    # is then a fixed pairwise tree — computed here with plain float adds.
    rng = np.random.default_rng(9)
    a, b = rng.standard_normal(64), rng.standard_normal(64)
    p = (a * b).tolist()
    for off in (32, 16, 8, 4, 2, 1):
        p = [p[l] + p[l ^ off] for l in range(64)]
    assert O.pair_fp64(a, b, O.DOT) == p[0]


# ---- the harness that times the reference's own CPU path ----------------

def _ref_tree_by_inserts(L, X, D):
    """Grow the tree with the reference's own kdtree_insert (stdout silenced:
    it prints per node and per level)."""
    import ctypes as C
    import sys
    L.kdtree_create.restype = C.c_void_p
    L.kdtree_create.argtypes = [C.c_size_t]
    L.kdtree_insert.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    libc = C.CDLL(None)
    sys.stdout.flush()
    libc.fflush(None)
    saved, null = os.dup(1), os.open(os.devnull, os.O_WRONLY)
    os.dup2(null, 1)
    try:
        t = L.kdtree_create(D)
        rows = [np.ascontiguousarray(r) for r in X]
        for i, r in enumerate(rows):
            L.kdtree_insert(t, r.ctypes.data, i)
        libc.fflush(None)
    finally:
        os.dup2(saved, 1)
        os.close(null)
        os.close(saved)
    return t, rows


@pytest.mark.parametrize("D,threads", [(3, 1), (3, 8), (16, 4)])
def test_reftree_equals_the_tree_kdtree_insert_grows(D, threads):
    """orc_reftree_build lays out node for node the tree the reference's own
    kdtree_insert grows (kdtree.c:47-62), walked path by path."""
    L = O.ref_lib()
    if L is None:
        pytest.skip("reference build (oracle/_ref) absent")
    rng = np.random.default_rng(5)
    X = rng.standard_normal((3000, 16))
    X[100:140] = X[7]  # duplicates: equal coordinates go right (kdtree.c:55-58)
    t_ref, keep = _ref_tree_by_inserts(L, X, D)
    ours = O.RefTree(X, D, threads=threads)
    assert ours.shape()[0] == 3000
    # breadth-first over every path of the reference's tree
    frontier = [b""]
    seen = 0
    while frontier:
        nxt = []
        for p in frontier:
            a = O.lib().orc_reftree_at(t_ref, p, len(p))
            b = ours.at(p)
            assert a == b, (p, a, b)
            if a != (1 << 64) - 1:
                seen += 1
                nxt += [p + b"\x00", p + b"\x01"]
        frontier = nxt
    assert seen == 3000


def test_reference_nearest_through_the_harness(c1_data, golden):
    """The reference's compiled kdtree_nearest over the harness's tree gives
    the reference's own C1 answers (golden, from trees grown by inserts), in
    both builds and on several threads."""
    X, Q, _ = c1_data
    for build in ("as_shipped", "O2"):
        L = O.ref_lib(build)
        if L is None:
            pytest.skip("reference build (oracle/_ref) absent")
        t = O.RefTree(X, 3, threads=4)
        ids, ns, wall = t.time_nearest(O.fn_addr(L, "kdtree_nearest"), Q, threads=4)
        assert np.array_equal(ids, golden["c1_kd3_ids"]), build
        assert wall > 0 and (ns > 0).all()


def test_reference_metrics_through_the_harness(c1_data, golden):
    X, _, P = c1_data
    for build in ("as_shipped", "O2"):
        L = O.ref_lib(build)
        if L is None:
            pytest.skip("reference build (oracle/_ref) absent")
        for col, name in enumerate(("cosine_similarity", "euclidean_distance", "dot_product")):
            out, wall = O.time_ref_pairs(O.fn_addr(L, name), X, P[:, 0], P[:, 1], 3)
            assert np.array_equal(out.view(np.uint32), golden["c1_metric_bits"][:, col]), \
                (build, name)
