"""svdb_mfma_sustained: the matrix-core rate loops run and report rates
between a floor no working device falls below and the published peaks
(FP64 78.6 TF/s, BF16 2.5 PF/s, FP8 5.0 PF/s dense) with a margin for the
clock."""
import math
import sys

import pytest

import svdb

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,lo,hi", [("fp64", 20.0, 90.0), ("bf16", 500.0, 2900.0),
                                        ("fp8", 1000.0, 5800.0)])
def test_sustained_rate_in_range(kind, lo, hi):
    print(f"mfma_sustained {kind}", flush=True)
    sys.stdout.flush()
    r = svdb.mfma_sustained(kind)
    print(f"  {r:.1f} TF/s", flush=True)
    assert math.isfinite(r) and lo < r < hi
