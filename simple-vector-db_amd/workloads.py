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


def _mixture_rows(out, C, rng, sigma, lognorm):
    m, d = out.shape
    j = rng.integers(0, C.shape[0], m)
    u = C[j] + sigma * rng.standard_normal((m, d))
    r = rng.lognormal(0.0, lognorm, (m, 1))
    out[:] = r * u / np.linalg.norm(u, axis=1, keepdims=True)


def mixture(n, d, seed, n_centers=1024, sigma=0.05, lognorm=0.5, center_seed=3,
            chunk=1_000_000, threads=None):
    """Rows of the cluster mixture.  Up to `chunk` rows: one PCG64 stream from
    `seed`.  Larger stores: blocks of `chunk` rows, block b drawn from the
    stream default_rng([seed, b]) (the same rows whatever the thread count),
    on `threads` threads (default: this process's cores, at most 16)."""
    C = centers(n_centers, d, center_seed)
    out = np.empty((n, d))
    if n <= chunk:
        _mixture_rows(out, C, np.random.default_rng(seed), sigma, lognorm)
        return out
    import concurrent.futures as cf
    import os
    nb = (n + chunk - 1) // chunk
    th = threads or min(16, len(os.sched_getaffinity(0)))

    def block(b):
        s = b * chunk
        # smaller pieces inside a block keep the temporaries bounded
        rng = np.random.default_rng([seed, b])
        e = min(n, s + chunk)
        step = max(1, (1 << 27) // (8 * d))
        for p in range(s, e, step):
            _mixture_rows(out[p:min(e, p + step)], C, rng, sigma, lognorm)

    with cf.ThreadPoolExecutor(max_workers=max(1, th)) as ex:
        list(ex.map(block, range(nb)))
    return out


# The mixture-precision workload: 10,000,000 x 128 rows from 1,024 clusters
# (σ = 0.05, LogNormal(0, 0.5) norms), 100,000 queries from the same mixture.
C3_ROWS, C3_DIM, C3_QUERIES = 10_000_000, 128, 100_000
C3_ROW_SEED, C3_QUERY_SEED = 31, 32


def c3_rows(n=C3_ROWS, d=C3_DIM):
    return mixture(n, d, C3_ROW_SEED)


def c3_queries(n=C3_QUERIES, d=C3_DIM):
    return mixture(n, d, C3_QUERY_SEED)
