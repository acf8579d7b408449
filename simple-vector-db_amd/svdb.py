"""ctypes binding of libsvdb.so (the C ABI in include/*.h).

Mirrors the reference's C interface (vector_database.h, kdtree.h) and the
batch entry points (svdb.h) for tests and the benchmark.  Every call goes
through libsvdb.so; if the library is missing the import fails loudly —
there is no host fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libsvdb.so")

L2, COS, DOT = 0, 1, 2
METRICS = {"l2": L2, "euclidean": L2, "cos": COS, "cosine": COS, "dot": DOT}
PREC_FP64, PREC_BF16, PREC_FP8, PREC_AUTO = 0, 1, 2, 3
PRECISIONS = {"fp64": PREC_FP64, "bf16": PREC_BF16, "fp8": PREC_FP8, "auto": PREC_AUTO}
K_SCAN_FP64, K_SCAN_LOWP, K_SHORTLIST, K_RERANK, K_MERGE, K_PREPASS, K_SECOND_PASS, K_COUNT = range(8)
UUID_SIZE = 37
MAX_K = 128
NOID = np.uint64(2**64 - 1)

# Symbols include/*.h declares (checked by the CPU test suite).
EXPORTS = [
    "kdtree_create", "kdtree_insert", "kdtree_free", "kdtree_nearest",
    "vector_db_init", "vector_db_free", "vector_db_insert", "vector_db_read",
    "vector_db_read_by_uuid", "vector_db_update", "vector_db_delete", "vector_db_save",
    "vector_db_load", "cosine_similarity", "euclidean_distance", "dot_product",
    "svdb_search", "svdb_search_device", "svdb_pair_metrics", "svdb_bulk_insert",
    "svdb_set_shortlist", "svdb_set_budget_bytes", "svdb_enable_shadow", "svdb_get_stats",
    "svdb_set_profiling", "svdb_get_kernel_times", "svdb_last_error", "svdb_device_alloc",
    "svdb_device_free", "svdb_device_copy", "svdb_device_sync", "svdb_server_default_config",
    "svdb_server_start", "svdb_server_stop", "svdb_server_submit_search",
    "svdb_server_submit_insert", "svdb_server_submit_update", "svdb_server_submit_delete",
    "svdb_server_wait", "svdb_debug_fail_host_alloc", "svdb_nearest_request",
    "svdb_search_device_noalloc", "svdb_mfma_sustained",
]


class Vector(C.Structure):
    _fields_ = [("uuid", C.c_char * UUID_SIZE), ("dimension", C.c_size_t),
                ("data", C.POINTER(C.c_double))]


class KDTree(C.Structure):
    _fields_ = [("root", C.c_void_p), ("dimension", C.c_size_t), ("svdb_index", C.c_void_p)]


class VectorDatabase(C.Structure):
    _fields_ = [("vectors", C.POINTER(Vector)), ("size", C.c_size_t), ("capacity", C.c_size_t),
                ("kdtree", C.POINTER(KDTree)), ("mutex", C.c_byte * 40),
                ("svdb_engine", C.c_void_p), ("svdb_uuid", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("rows", C.c_size_t), ("dim", C.c_size_t), ("phys_rows", C.c_size_t),
                ("device_bytes", C.c_size_t), ("budget_bytes", C.c_size_t),
                ("searches", C.c_uint64), ("queries", C.c_uint64), ("certified", C.c_uint64),
                ("fallbacks", C.c_uint64), ("last_max_rel_err", C.c_double),
                ("widened", C.c_uint64), ("second_pass", C.c_uint64),
                ("compactions", C.c_uint64), ("graph_launches", C.c_uint64),
                ("graph_captures", C.c_uint64)]


class KernelTime(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("total_ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_size_t), ("budget_bytes", C.c_size_t), ("precision", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("submit_us", C.c_int64), ("start_us", C.c_int64), ("done_us", C.c_int64),
                ("tick", C.c_uint64), ("tick_queries", C.c_size_t)]


_lib = None
_libc = None


def lib():
    global _lib, _libc
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is not built: run __graft_entry__.build() or make -C {HERE}")
    L = C.CDLL(LIBPATH)
    DBP, TP, VP, sz = C.POINTER(VectorDatabase), C.POINTER(KDTree), C.POINTER(Vector), C.c_size_t
    vp = C.c_void_p
    L.kdtree_create.restype = TP
    L.kdtree_create.argtypes = [sz]
    L.kdtree_insert.argtypes = [TP, vp, sz]
    L.kdtree_free.argtypes = [TP]
    L.kdtree_nearest.restype = sz
    L.kdtree_nearest.argtypes = [TP, vp]
    L.vector_db_init.restype = DBP
    L.vector_db_init.argtypes = [sz, sz]
    L.vector_db_free.argtypes = [DBP]
    L.vector_db_insert.restype = sz
    L.vector_db_insert.argtypes = [DBP, Vector]
    L.vector_db_read.restype = VP
    L.vector_db_read.argtypes = [DBP, sz]
    L.vector_db_read_by_uuid.restype = VP
    L.vector_db_read_by_uuid.argtypes = [DBP, C.c_char_p]
    L.vector_db_update.argtypes = [DBP, sz, Vector]
    L.vector_db_delete.argtypes = [DBP, sz]
    L.vector_db_save.argtypes = [DBP, C.c_char_p]
    L.vector_db_load.restype = DBP
    L.vector_db_load.argtypes = [C.c_char_p, sz]
    for f in ("cosine_similarity", "euclidean_distance", "dot_product"):
        getattr(L, f).restype = C.c_float
        getattr(L, f).argtypes = [Vector, Vector]
    L.svdb_search.restype = C.c_int
    L.svdb_search.argtypes = [DBP, vp, sz, sz, vp, vp, C.c_int, sz, vp, vp]
    L.svdb_search_device.restype = C.c_int
    L.svdb_search_device.argtypes = [DBP, vp, sz, sz, sz, C.c_int, C.c_int, vp, vp]
    L.svdb_search_device_noalloc.restype = C.c_int
    L.svdb_search_device_noalloc.argtypes = [DBP, vp, sz, sz, sz, C.c_int, vp, vp]
    L.svdb_pair_metrics.restype = C.c_int
    L.svdb_pair_metrics.argtypes = [DBP, vp, vp, sz, C.c_int, vp]
    L.svdb_bulk_insert.restype = C.c_int
    L.svdb_bulk_insert.argtypes = [DBP, vp, sz, sz, C.c_char_p, C.POINTER(C.c_size_t)]
    for f in ("svdb_set_shortlist", "svdb_set_budget_bytes"):
        getattr(L, f).restype = C.c_int
        getattr(L, f).argtypes = [DBP, sz]
    L.svdb_enable_shadow.restype = C.c_int
    L.svdb_enable_shadow.argtypes = [DBP, C.c_int]
    L.svdb_get_stats.restype = C.c_int
    L.svdb_get_stats.argtypes = [DBP, C.POINTER(Stats)]
    L.svdb_set_profiling.restype = C.c_int
    L.svdb_set_profiling.argtypes = [DBP, C.c_int]
    L.svdb_get_kernel_times.restype = C.c_int
    L.svdb_get_kernel_times.argtypes = [DBP, C.POINTER(KernelTime)]
    L.svdb_last_error.restype = C.c_char_p
    L.svdb_device_alloc.restype = C.c_void_p
    L.svdb_device_alloc.argtypes = [C.c_size_t]
    L.svdb_device_free.argtypes = [C.c_void_p]
    L.svdb_device_copy.restype = C.c_int
    L.svdb_device_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.svdb_device_sync.restype = C.c_int
    L.svdb_server_default_config.argtypes = [C.POINTER(ServerConfig)]
    L.svdb_server_start.restype = C.c_void_p
    L.svdb_server_start.argtypes = [DBP, C.POINTER(ServerConfig)]
    L.svdb_server_stop.argtypes = [vp]
    L.svdb_server_submit_search.restype = C.c_int64
    L.svdb_server_submit_search.argtypes = [vp, vp, sz, sz, sz, C.c_int]
    L.svdb_server_submit_insert.restype = C.c_int64
    L.svdb_server_submit_insert.argtypes = [vp, vp, sz]
    L.svdb_server_submit_update.restype = C.c_int64
    L.svdb_server_submit_update.argtypes = [vp, sz, vp, sz]
    L.svdb_server_submit_delete.restype = C.c_int64
    L.svdb_server_submit_delete.argtypes = [vp, sz]
    L.svdb_server_wait.restype = C.c_int
    L.svdb_server_wait.argtypes = [vp, C.c_int64, vp, vp, C.POINTER(RequestInfo)]
    L.svdb_nearest_request.restype = C.c_int
    L.svdb_nearest_request.argtypes = [DBP, vp, sz, C.c_char_p, sz, C.c_char_p, C.c_char_p,
                                       C.c_int, C.POINTER(C.c_char_p), C.POINTER(sz)]
    L.svdb_debug_fail_host_alloc.restype = C.c_int64
    L.svdb_debug_fail_host_alloc.argtypes = [C.c_int64]
    _libc = C.CDLL(None)
    _libc.malloc.restype = C.c_void_p
    _libc.malloc.argtypes = [C.c_size_t]
    _libc.free.argtypes = [C.c_void_p]
    _lib = L
    return L


def free_response(p):
    """free() a response text returned by svdb_nearest_request."""
    lib()
    _libc.free(C.cast(p, C.c_void_p))


def nearest_request(db, body, expected_dim, number=None, metric=None, precision="auto",
                    server=None):
    """svdb_nearest_request -> (HTTP status, response text)."""
    out = C.c_char_p()
    n = C.c_size_t(0)
    raw = body.encode() if isinstance(body, str) else body
    st = lib().svdb_nearest_request(db.p if db is not None else None,
                                    server.p if server is not None else None, expected_dim,
                                    raw, len(raw),
                                    number.encode() if number is not None else None,
                                    metric.encode() if metric is not None else None,
                                    _prec(precision), C.byref(out), C.byref(n))
    text = C.string_at(out, n.value).decode()
    free_response(out)
    return st, text


def fail_host_alloc(n):
    """Test hook: the n-th host allocation point from now fails (0 disarms)."""
    return lib().svdb_debug_fail_host_alloc(n)


class SvdbError(RuntimeError):
    def __init__(self, code, where):
        msg = lib().svdb_last_error().decode(errors="replace")
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def _chk(rc, where):
    if rc != 0:
        raise SvdbError(rc, where)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _metric(m):
    return METRICS[m] if isinstance(m, str) else int(m)


def _prec(p):
    return PRECISIONS[p] if isinstance(p, str) else int(p)


def owned_vector(row, uuid=b""):
    """A Vector whose data is malloc'ed (ownership passes to the store)."""
    lib()
    row = _f64(row)
    buf = _libc.malloc(max(row.nbytes, 8))
    C.memmove(buf, row.ctypes.data, row.nbytes)
    v = Vector()
    v.uuid = uuid[:UUID_SIZE - 1] if isinstance(uuid, bytes) else uuid.encode()[:UUID_SIZE - 1]
    v.dimension = row.shape[0]
    v.data = C.cast(buf, C.POINTER(C.c_double))
    return v


def borrowed_vector(row, uuid=b""):
    """A Vector viewing a numpy array (for the metric functions)."""
    row = _f64(row)
    v = Vector()
    v.uuid = uuid
    v.dimension = row.shape[0]
    v.data = row.ctypes.data_as(C.POINTER(C.c_double))
    v._keep = row
    return v


class DeviceArray:
    """A device buffer from svdb_device_alloc (same HIP runtime as the engine)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = lib().svdb_device_alloc(self.nbytes)
        if not self.ptr:
            raise SvdbError(-2, "svdb_device_alloc")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _chk(lib().svdb_device_copy(d.ptr, a.ctypes.data, a.nbytes, 0), "svdb_device_copy")
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        _chk(lib().svdb_device_copy(out.ctypes.data, self.ptr, self.nbytes, 1), "svdb_device_copy")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().svdb_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_sync():
    _chk(lib().svdb_device_sync(), "svdb_device_sync")


MFMA_KINDS = {"fp64": 0, "bf16": 1, "fp8": 2}


def mfma_sustained(kind):
    """Sustained matrix-core TFLOP/s of one data type on this device
    (svdb_mfma_sustained): a loop of independent MFMAs, no memory traffic."""
    v = C.c_double()
    _chk(lib().svdb_mfma_sustained(MFMA_KINDS[kind], C.byref(v)), "svdb_mfma_sustained")
    return v.value


def legacy_metric(a, b, metric):
    """cosine_similarity / euclidean_distance / dot_product (float results)."""
    f = {L2: "euclidean_distance", COS: "cosine_similarity", DOT: "dot_product"}[_metric(metric)]
    va, vb = borrowed_vector(a), borrowed_vector(b)
    return getattr(lib(), f)(va, vb)


class Database:
    """A VectorDatabase* from vector_db_init / vector_db_load."""

    def __init__(self, kd_dimension=3, capacity=0, _ptr=None):
        L = lib()
        self.p = _ptr if _ptr is not None else L.vector_db_init(capacity, kd_dimension)
        if not self.p:
            raise SvdbError(-3, "vector_db_init")

    @classmethod
    def load(cls, filename, kd_dimension=3):
        p = lib().vector_db_load(os.fsencode(filename), kd_dimension)
        if not p:
            raise SvdbError(-1, "vector_db_load")
        return cls(_ptr=p)

    def close(self):
        if getattr(self, "p", None):
            lib().vector_db_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return self.p.contents.size

    # ---- reference store API ----
    def insert(self, row, uuid=b""):
        return lib().vector_db_insert(self.p, owned_vector(row, uuid))

    def update(self, index, row, uuid=b""):
        lib().vector_db_update(self.p, index, owned_vector(row, uuid))

    def delete(self, index):
        lib().vector_db_delete(self.p, index)

    def read(self, index):
        v = lib().vector_db_read(self.p, index)
        if not v:
            return None
        v = v.contents
        return v.uuid, np.ctypeslib.as_array(v.data, (v.dimension,)).copy()

    def read_by_uuid(self, uuid):
        v = lib().vector_db_read_by_uuid(self.p, uuid)
        if not v:
            return None
        v = v.contents
        return v.uuid, np.ctypeslib.as_array(v.data, (v.dimension,)).copy()

    def save(self, filename):
        lib().vector_db_save(self.p, os.fsencode(filename))

    def kd_nearest(self, q):
        q = _f64(q)
        return lib().kdtree_nearest(self.p.contents.kdtree, q.ctypes.data)

    def bulk_insert(self, rows, uuids=None):
        rows = _f64(rows)
        n, d = rows.shape
        first = C.c_size_t(0)
        ub = None
        if uuids is not None:
            ub = b"".join((u if isinstance(u, bytes) else u.encode())[:UUID_SIZE - 1]
                          .ljust(UUID_SIZE, b"\0") for u in uuids)
        _chk(lib().svdb_bulk_insert(self.p, rows.ctypes.data, n, d, ub, C.byref(first)),
             "svdb_bulk_insert")
        return first.value

    # ---- batch API ----
    def search(self, Q, k, metric="l2", precision="fp64"):
        Q = _f64(np.atleast_2d(Q))
        nq, d = Q.shape
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.uint64), (nq,)))
        ms = np.ascontiguousarray(np.broadcast_to(
            np.asarray([_metric(m) for m in np.atleast_1d(metric)], dtype=np.int32)
            if not isinstance(metric, (str, int)) else np.int32(_metric(metric)), (nq,)))
        kmax = int(ks.max()) if nq else 1
        ids = np.full((nq, kmax), NOID, dtype=np.uint64)
        sc = np.full((nq, kmax), np.nan)
        _chk(lib().svdb_search(self.p, Q.ctypes.data, nq, d, ks.ctypes.data, ms.ctypes.data,
                               _prec(precision), kmax, ids.ctypes.data, sc.ctypes.data),
             "svdb_search")
        return ids, sc

    def search_device(self, dq_ptr, nq, ldq, k, metric, precision, dids_ptr, dscores_ptr):
        _chk(lib().svdb_search_device(self.p, dq_ptr, nq, ldq, k, _metric(metric),
                                      _prec(precision), dids_ptr, dscores_ptr),
             "svdb_search_device")

    def pair_metrics(self, a, b, metric):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        out = np.empty(a.shape[0])
        _chk(lib().svdb_pair_metrics(self.p, a.ctypes.data, b.ctypes.data, a.shape[0],
                                     _metric(metric), out.ctypes.data), "svdb_pair_metrics")
        return out

    def set_shortlist(self, L):
        _chk(lib().svdb_set_shortlist(self.p, L), "svdb_set_shortlist")

    def set_budget(self, nbytes):
        _chk(lib().svdb_set_budget_bytes(self.p, nbytes), "svdb_set_budget_bytes")

    def enable_shadow(self, precision):
        _chk(lib().svdb_enable_shadow(self.p, _prec(precision)), "svdb_enable_shadow")

    def stats(self):
        s = Stats()
        _chk(lib().svdb_get_stats(self.p, C.byref(s)), "svdb_get_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def set_profiling(self, on=True):
        _chk(lib().svdb_set_profiling(self.p, int(on)), "svdb_set_profiling")

    def kernel_times(self):
        kt = (KernelTime * K_COUNT)()
        _chk(lib().svdb_get_kernel_times(self.p, kt), "svdb_get_kernel_times")
        return [(kt[i].launches, kt[i].total_ms) for i in range(K_COUNT)]


class StandaloneKDTree:
    """A KDTree* from kdtree_create (not attached to a database)."""

    def __init__(self, dimension):
        self.p = lib().kdtree_create(dimension)
        if not self.p:
            raise SvdbError(-3, "kdtree_create")

    def insert(self, point, index):
        point = _f64(point)
        lib().kdtree_insert(self.p, point.ctypes.data, index)

    def nearest(self, q):
        q = _f64(q)
        return lib().kdtree_nearest(self.p, q.ctypes.data)

    def close(self):
        if getattr(self, "p", None):
            lib().kdtree_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """A continuous-batching server over a Database (svdb_server_*)."""

    def __init__(self, db, max_batch=8192, budget_bytes=64 << 30, precision="auto"):
        cfg = ServerConfig()
        lib().svdb_server_default_config(C.byref(cfg))
        cfg.max_batch = max_batch
        cfg.budget_bytes = budget_bytes
        cfg.precision = _prec(precision)
        self.db = db
        self.p = lib().svdb_server_start(db.p, C.byref(cfg))
        if not self.p:
            raise SvdbError(-1, "svdb_server_start")
        self.pending = {}

    def submit_search(self, Q, k, metric):
        Q = _f64(np.atleast_2d(Q))
        t = lib().svdb_server_submit_search(self.p, Q.ctypes.data, Q.shape[0], Q.shape[1], k,
                                            _metric(metric))
        if t < 0:
            raise SvdbError(int(t), "svdb_server_submit_search")
        self.pending[t] = ("search", Q.shape[0], k)
        return t

    def submit_insert(self, row):
        row = _f64(row)
        t = lib().svdb_server_submit_insert(self.p, row.ctypes.data, row.shape[0])
        if t < 0:
            raise SvdbError(int(t), "svdb_server_submit_insert")
        self.pending[t] = ("insert", 1, 1)
        return t

    def submit_update(self, index, row):
        row = _f64(row)
        t = lib().svdb_server_submit_update(self.p, index, row.ctypes.data, row.shape[0])
        if t < 0:
            raise SvdbError(int(t), "svdb_server_submit_update")
        self.pending[t] = ("update", 0, 0)
        return t

    def submit_delete(self, index):
        t = lib().svdb_server_submit_delete(self.p, index)
        if t < 0:
            raise SvdbError(int(t), "svdb_server_submit_delete")
        self.pending[t] = ("delete", 0, 0)
        return t

    def wait(self, ticket):
        """-> (rc, ids, scores, info dict) ; ids/scores shaped (nq, k) for searches."""
        kind, nq, k = self.pending.pop(ticket)
        ids = np.full((max(nq, 1), max(k, 1)), NOID, dtype=np.uint64)
        sc = np.full((max(nq, 1), max(k, 1)), np.nan)
        info = RequestInfo()
        rc = lib().svdb_server_wait(self.p, ticket, ids.ctypes.data, sc.ctypes.data,
                                    C.byref(info))
        inf = {f: getattr(info, f) for f, _ in RequestInfo._fields_}
        if kind == "search":
            return rc, ids, sc, inf
        return rc, (int(ids[0, 0]) if kind == "insert" else None), None, inf

    def stop(self):
        if getattr(self, "p", None):
            lib().svdb_server_stop(self.p)
            self.p = None

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
