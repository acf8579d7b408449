import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "simple-vector-db_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "c1_reference.npz")))


@pytest.fixture(scope="session")
def c1_data():
    import numpy as np
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (10000, 128))
    Q = rng.uniform(-1, 1, (256, 128))
    P = rng.integers(0, 10000, (256, 2))
    return X, Q, P
