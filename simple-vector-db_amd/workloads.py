"""Synthetic workloads for the nearest-vector path (host-side numpy, PCG64).

All data is drawn on the host with numpy.random.default_rng and copied to
the device, so every run (and the CPU oracle) sees identical doubles.

  gaussian(n, d, seed)         rows i.i.d. N(0, 1)
  mixture(n, d, seed, ...)     cluster mixture: centres uniform on the unit
                               sphere; a row is x = r·u/‖u‖ with u = c_j + σz,
                               j uniform over the centres, z ~ N(0, I) (σ per
                               coordinate) and r ~ LogNormal(0, s)
"""
import numpy as np


def gaussian(n, d, seed):
    return np.random.default_rng(seed).standard_normal((n, d))


def centers(n_centers, d, seed):
    c = np.random.default_rng(seed).standard_normal((n_centers, d))
    return c / np.linalg.norm(c, axis=1, keepdims=True)


def mixture(n, d, seed, n_centers=1024, sigma=0.05, lognorm=0.5, center_seed=3,
            chunk=1_000_000):
    """Rows of the cluster mixture, generated in chunks (bounded temporaries)."""
    C = centers(n_centers, d, center_seed)
    rng = np.random.default_rng(seed)
    out = np.empty((n, d))
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        j = rng.integers(0, n_centers, m)
        u = C[j] + sigma * rng.standard_normal((m, d))
        r = rng.lognormal(0.0, lognorm, (m, 1))
        out[s:s + m] = r * u / np.linalg.norm(u, axis=1, keepdims=True)
    return out


# The mixture-precision workload: 10,000,000 x 128 rows from 1,024 clusters
# (σ = 0.05, LogNormal(0, 0.5) norms), 100,000 queries from the same mixture.
C3_ROWS, C3_DIM, C3_QUERIES = 10_000_000, 128, 100_000
C3_ROW_SEED, C3_QUERY_SEED = 31, 32


def c3_rows(n=C3_ROWS, d=C3_DIM):
    return mixture(n, d, C3_ROW_SEED)


def c3_queries(n=C3_QUERIES, d=C3_DIM):
    return mixture(n, d, C3_QUERY_SEED)
