"""ctypes binding of liboracle.so — the CPU oracle (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  See svdb_oracle.h for the contracts.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libsvdbref.so")

L2, COS, DOT = 0, 1, 2
METRICS = {"l2": L2, "euclidean": L2, "cos": COS, "cosine": COS, "dot": DOT}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB)
        vp, sz, dp, u8p, u64p = C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p
        L.orc_kd_create.restype = vp
        L.orc_kd_create.argtypes = [sz]
        L.orc_kd_build.restype = vp
        L.orc_kd_build.argtypes = [dp, sz, sz, sz]
        L.orc_kd_insert.argtypes = [vp, dp, sz]
        L.orc_kd_nearest.restype = sz
        L.orc_kd_nearest.argtypes = [vp, dp]
        L.orc_kd_nearest_batch.argtypes = [vp, dp, sz, sz, C.c_int, u64p]
        L.orc_kd_depth.restype = sz
        L.orc_kd_depth.argtypes = [vp]
        L.orc_kd_free.argtypes = [vp]
        for f in ("orc_ref_cos", "orc_ref_l2", "orc_ref_dot"):
            getattr(L, f).restype = C.c_float
            getattr(L, f).argtypes = [dp, dp, sz]
        L.orc_ref_pairs.argtypes = [dp, sz, sz, u64p, u64p, sz, C.c_int, vp]
        for f in ("orc_l2_key", "orc_dot_key"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [dp, dp, sz]
        L.orc_sumsq.restype = C.c_double
        L.orc_sumsq.argtypes = [dp, sz]
        L.orc_cos_score.restype = C.c_double
        L.orc_cos_score.argtypes = [C.c_double, C.c_double, C.c_double]
        for f in ("orc_topk", "orc_topk_scalar"):
            getattr(L, f).restype = C.c_int
        L.orc_topk.argtypes = [dp, sz, sz, u8p, dp, sz, sz, sz, C.c_int, sz, C.c_int, u64p, dp]
        L.orc_topk_scalar.argtypes = [dp, sz, sz, u8p, dp, sz, sz, sz, C.c_int, sz, u64p, dp]
        L.orc_pair_fp64.restype = C.c_double
        L.orc_pair_fp64.argtypes = [dp, dp, sz, C.c_int]
        L.orc_pairs_fp64.argtypes = [dp, sz, sz, u64p, u64p, sz, C.c_int, dp]
        L.orc_reftree_build.restype = vp
        L.orc_reftree_build.argtypes = [dp, sz, sz, sz, C.c_int]
        L.orc_reftree_handle.restype = vp
        L.orc_reftree_handle.argtypes = [vp]
        L.orc_reftree_at.restype = sz
        L.orc_reftree_at.argtypes = [vp, C.c_char_p, sz]
        L.orc_reftree_shape.argtypes = [vp, C.POINTER(sz), C.POINTER(sz)]
        L.orc_reftree_free.argtypes = [vp]
        L.orc_time_kd.restype = C.c_double
        L.orc_time_kd.argtypes = [vp, vp, dp, sz, sz, C.c_int, u64p, dp]
        L.orc_time_pairs.restype = C.c_double
        L.orc_time_pairs.argtypes = [vp, dp, sz, sz, u64p, u64p, sz, C.c_int, vp]
        _lib = L
    return _lib


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _metric(m):
    return METRICS[m] if isinstance(m, str) else int(m)


def topk(X, Q, k, metric, D=None, alive=None, threads=None, scalar=False):
    """Canonical FP64 exhaustive top-k.  Returns (ids uint64 [B,k], scores f64 [B,k])."""
    X = _f64(X)
    Q = _f64(np.atleast_2d(Q))
    N, d = X.shape
    B = Q.shape[0]
    D = d if D is None else D
    assert Q.shape[1] >= D and d >= D
    ids = np.empty((B, k), dtype=np.uint64)
    sc = np.empty((B, k), dtype=np.float64)
    al = None
    if alive is not None:
        al = np.ascontiguousarray(alive, dtype=np.uint8)
        assert al.shape[0] == N
    alp = al.ctypes.data if al is not None else None
    if scalar:
        rc = lib().orc_topk_scalar(X.ctypes.data, N, d, alp, Q.ctypes.data, B, Q.shape[1], D,
                                   _metric(metric), k, ids.ctypes.data, sc.ctypes.data)
    else:
        th = threads or os.cpu_count() or 1
        rc = lib().orc_topk(X.ctypes.data, N, d, alp, Q.ctypes.data, B, Q.shape[1], D,
                            _metric(metric), k, int(th), ids.ctypes.data, sc.ctypes.data)
    if rc != 0:
        raise ValueError("orc_topk failed")
    return ids, sc


def l2_key(q, x):
    q, x = _f64(q), _f64(x)
    return lib().orc_l2_key(q.ctypes.data, x.ctypes.data, q.shape[0])


def dot_key(q, x):
    q, x = _f64(q), _f64(x)
    return lib().orc_dot_key(q.ctypes.data, x.ctypes.data, q.shape[0])


def sumsq(x):
    x = _f64(x)
    return lib().orc_sumsq(x.ctypes.data, x.shape[0])


def pair_fp64(a, b, metric):
    a, b = _f64(a), _f64(b)
    return lib().orc_pair_fp64(a.ctypes.data, b.ctypes.data, a.shape[0], _metric(metric))


def pairs_fp64(A, ia, ib, metric):
    A = _f64(A)
    ia = np.ascontiguousarray(ia, dtype=np.uint64)
    ib = np.ascontiguousarray(ib, dtype=np.uint64)
    out = np.empty(ia.shape[0], dtype=np.float64)
    lib().orc_pairs_fp64(A.ctypes.data, A.shape[1], A.shape[1], ia.ctypes.data, ib.ctypes.data,
                         ia.shape[0], _metric(metric), out.ctypes.data)
    return out


def ref_pairs(A, ia, ib, metric):
    """Reference float metrics (vector_database.c:301-352) for pairs of rows of A."""
    A = _f64(A)
    ia = np.ascontiguousarray(ia, dtype=np.uint64)
    ib = np.ascontiguousarray(ib, dtype=np.uint64)
    out = np.empty(ia.shape[0], dtype=np.float32)
    lib().orc_ref_pairs(A.ctypes.data, A.shape[1], A.shape[1], ia.ctypes.data, ib.ctypes.data,
                        ia.shape[0], _metric(metric), out.ctypes.data)
    return out


class KDTree:
    """Restatement of the reference kd-tree (kdtree.c), printf-free."""

    def __init__(self, X=None, dimension=3):
        self.D = dimension
        if X is not None:
            X = _f64(X)
            self._keep = X
            self.h = lib().orc_kd_build(X.ctypes.data, X.shape[0], X.shape[1], dimension)
        else:
            self.h = lib().orc_kd_create(dimension)

    def insert(self, p, index):
        p = _f64(p)
        lib().orc_kd_insert(self.h, p.ctypes.data, index)

    def nearest(self, q):
        q = _f64(q)
        return lib().orc_kd_nearest(self.h, q.ctypes.data)

    def nearest_batch(self, Q, threads=1):
        Q = _f64(np.atleast_2d(Q))
        out = np.empty(Q.shape[0], dtype=np.uint64)
        lib().orc_kd_nearest_batch(self.h, Q.ctypes.data, Q.shape[0], Q.shape[1], int(threads),
                                   out.ctypes.data)
        return out

    def depth(self):
        return lib().orc_kd_depth(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_kd_free(self.h)
            self.h = None


# ---- the reference's own compiled CPU path (oracle/_ref), timed ----------

REF_BUILDS = {"as_shipped": REFLIB, "O2": os.path.join(HERE, "_ref", "libsvdbref_O2.so")}


def ref_lib(build="as_shipped"):
    """The reference's kdtree.c + vector_database.c built from its sources
    (oracle/Makefile `ref`), or None where that build is absent."""
    path = REF_BUILDS[build]
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.kdtree_nearest.restype = C.c_size_t
    L.kdtree_nearest.argtypes = [C.c_void_p, C.c_void_p]
    return L


def fn_addr(L, name):
    return C.cast(getattr(L, name), C.c_void_p).value


class RefTree:
    """The tree the reference's kdtree_insert grows from rows 0..n-1, laid out
    in its own KDTreeNode / KDTree structures (refkd_bench.c)."""

    def __init__(self, X, dimension, threads=1):
        X = _f64(X)
        self.h = lib().orc_reftree_build(X.ctypes.data, X.shape[0], X.shape[1], dimension,
                                         int(threads))
        if not self.h:
            raise MemoryError("orc_reftree_build failed")
        self.tree = lib().orc_reftree_handle(self.h)

    def at(self, path):
        return lib().orc_reftree_at(self.tree, bytes(path), len(path))

    def shape(self):
        c, h = C.c_size_t(0), C.c_size_t(0)
        lib().orc_reftree_shape(self.tree, C.byref(c), C.byref(h))
        return c.value, h.value

    def time_nearest(self, fn, Q, threads):
        """fn: address of a kdtree_nearest.  -> (ids, per-query ns, wall s)."""
        Q = _f64(np.atleast_2d(Q))
        ids = np.empty(Q.shape[0], dtype=np.uint64)
        ns = np.empty(Q.shape[0], dtype=np.float64)
        wall = lib().orc_time_kd(fn, self.tree, Q.ctypes.data, Q.shape[0], Q.shape[1],
                                 int(threads), ids.ctypes.data, ns.ctypes.data)
        return ids, ns, wall

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_reftree_free(self.h)
            self.h = None


def time_ref_pairs(fn, A, ia, ib, threads):
    """fn: address of a reference metric (cosine_similarity, ...).  -> (float out, wall s)."""
    A = _f64(A)
    ia = np.ascontiguousarray(ia, dtype=np.uint64)
    ib = np.ascontiguousarray(ib, dtype=np.uint64)
    out = np.empty(ia.shape[0], dtype=np.float32)
    wall = lib().orc_time_pairs(fn, A.ctypes.data, A.shape[1], A.shape[1], ia.ctypes.data,
                                ib.ctypes.data, ia.shape[0], int(threads), out.ctypes.data)
    return out, wall
