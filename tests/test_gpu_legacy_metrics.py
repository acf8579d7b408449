"""The drop-in float metrics (cosine_similarity, euclidean_distance,
dot_product; vector_database.c:301-352) one call at a time, through the C
ABI, bit for bit against the oracle's restatement of the reference: lengths
either side of the one-launch form's limit (224 elements per vector, the
pair carried in the kernel arguments) and the copy path beyond it, special
values, and a call after many others (the polled result word)."""
import sys

import numpy as np
import pytest

import oracle as O
import svdb

pytestmark = pytest.mark.gpu


def same(a, b):
    a, b = np.float32(a), np.float32(b)
    return (np.isnan(a) and np.isnan(b)) or a.view(np.uint32) == b.view(np.uint32)


@pytest.mark.parametrize("n", [0, 1, 3, 128, 223, 224, 225, 300, 1000])
def test_float_metrics_one_call_at_a_time(n):
    print(f"legacy metrics, n = {n}", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(900 + n)
    A = rng.standard_normal((40, max(n, 1)))[:, :n] * rng.choice([1e-3, 1.0, 1e3], (40, 1))
    if n:
        A[3, 0] = np.nan
        A[4, n - 1] = np.inf
        A[5] = 0.0
    ia = np.arange(0, 40, 2, dtype=np.uint64)
    ib = np.arange(1, 40, 2, dtype=np.uint64)
    for metric in ("cos", "l2", "dot"):
        want = O.ref_pairs(A, ia, ib, metric) if n else None
        for p in range(len(ia)):
            got = svdb.legacy_metric(A[ia[p]], A[ib[p]], metric)
            if n:
                assert same(got, want[p]), f"{metric} pair {p}: {got!r} vs {want[p]!r}"
            elif metric == "cos":
                assert np.isnan(got)
            else:
                assert got == 0.0
