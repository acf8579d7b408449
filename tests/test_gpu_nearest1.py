"""kdtree_nearest's one-launch path (k_nearest1: small stores, kd dimension
up to 256), through the C ABI, against the FP64 oracle: ties broken by the
lowest id, NaN and infinite rows, deleted rows (logical ids), a store with
every row deleted, stores that need many blocks, and the widths either side
of the path's limit (above it kdtree_nearest takes the batched scan)."""
import sys

import numpy as np
import pytest

import oracle as O
import svdb

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def expect(rows, q, D):
    ids, _ = O.topk(np.array(rows), np.asarray(q)[None, :], 1, "l2", D=D)
    return int(ids[0, 0])


@pytest.mark.parametrize("D", [1, 3, 32, 128])
def test_nearest_matches_oracle_with_ties_and_specials(D):
    rng = np.random.default_rng(300 + D)
    n, d = 3000, 128
    X = rng.integers(-3, 4, (n, d)).astype(np.float64)  # many exact ties
    X[10, 0] = np.nan
    X[11, min(D, d) - 1] = np.inf
    X[12] = -0.0
    say(f"nearest1: {n}x{d}, kd dimension {D}")
    db = svdb.Database(kd_dimension=D)
    db.bulk_insert(X)
    rows = [r for r in X]
    for i in range(40):
        q = rng.integers(-3, 4, d).astype(np.float64)
        if i == 0:
            q[:] = X[500]  # distance 0, ties with its duplicates
        assert db.kd_nearest(q) == expect(rows, q, D), f"query {i}"
    # deletes shift the logical ids; rows planted on later queries
    for j in sorted(rng.choice(n, 200, replace=False))[::-1]:
        db.delete(int(j))
        rows.pop(int(j))
    for i in range(20):
        q = rng.integers(-3, 4, d).astype(np.float64)
        assert db.kd_nearest(q) == expect(rows, q, D), f"after deletes, query {i}"
    db.close()


def test_many_blocks_and_every_row_deleted():
    rng = np.random.default_rng(7)
    n, d = 300_000, 8   # 1,024 blocks, each thread over several rows
    X = rng.standard_normal((n, d))
    say(f"nearest1: {n}x{d}, kd dimension 8")
    db = svdb.Database(kd_dimension=d)
    db.bulk_insert(X)
    for i in range(10):
        q = rng.standard_normal(d)
        assert db.kd_nearest(q) == expect(X, q, d), f"query {i}"
    db.close()
    small = svdb.Database(kd_dimension=4)
    small.bulk_insert(rng.standard_normal((3, 4)))
    for _ in range(3):
        small.delete(0)
    assert small.kd_nearest(rng.standard_normal(4)) == 2**64 - 1  # (size_t)-1
    small.close()


@pytest.mark.parametrize("D", [256, 288])
def test_widths_either_side_of_the_limit(D):
    rng = np.random.default_rng(D)
    n = 2000
    X = rng.standard_normal((n, D))
    say(f"nearest1: {n}x{D}, kd dimension {D}")
    db = svdb.Database(kd_dimension=D)
    db.bulk_insert(X)
    for i in range(8):
        q = rng.standard_normal(D)
        assert db.kd_nearest(q) == expect(X, q, D), f"query {i}"
    db.close()
