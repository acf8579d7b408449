"""Write the golden fixtures for the similarity / nearest-vector path.

Runs ONLY in the build container, where the reference tree is mounted: it
loads oracle/_ref/libsvdbref.so (the reference's kdtree.c + vector_database.c
compiled from source by oracle/Makefile) through ctypes and records its
outputs on seeded inputs.  Only the outputs (and the small inputs they need)
are committed, as tests/golden/c1_reference.npz and ref_save_small.bin.

Workload C1 (CPU parity): numpy default_rng(0) draws, in this order,
    DB    = uniform(-1, 1, (10000, 128))
    Q     = uniform(-1, 1, (256, 128))
    pairs = integers(0, 10000, (256, 2))
and the reference answers find-nearest with kd-tree dimension 3 and 128
(kdtree.c:87-91,171-178) and the three float metrics for the pairs
(vector_database.c:301-352).

Also recorded: a duplicate-row tie case (the reference keeps the first node
its depth-first search visits, kdtree.c:139), the stale-index case after
update/delete (vector_database.c:169-195 leave the tree untouched or add a
second node), special metric inputs, and a file written by vector_db_save
(vector_database.c:203-228) for the persistence-format tests.

Usage: python tests/golden/gen_golden.py   (after `make -C oracle`)
"""
import ctypes as C
import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libsvdbref.so")


class Vector(C.Structure):
    # vector_database.h:14-18 — uuid[37] @0, dimension @40, data @48
    _fields_ = [("uuid", C.c_char * 37), ("dimension", C.c_size_t),
                ("data", C.POINTER(C.c_double))]


def load_ref():
    if not os.path.exists(REFLIB):
        sys.exit("build the reference core first: make -C oracle")
    lib = C.CDLL(REFLIB)
    vp = C.c_void_p
    lib.kdtree_create.restype = vp
    lib.kdtree_create.argtypes = [C.c_size_t]
    lib.kdtree_insert.argtypes = [vp, C.POINTER(C.c_double), C.c_size_t]
    lib.kdtree_nearest.restype = C.c_size_t
    lib.kdtree_nearest.argtypes = [vp, C.POINTER(C.c_double)]
    lib.kdtree_free.argtypes = [vp]
    for f in ("cosine_similarity", "euclidean_distance", "dot_product"):
        getattr(lib, f).restype = C.c_float
        getattr(lib, f).argtypes = [Vector, Vector]
    lib.vector_db_init.restype = vp
    lib.vector_db_init.argtypes = [C.c_size_t, C.c_size_t]
    lib.vector_db_insert.restype = C.c_size_t
    lib.vector_db_insert.argtypes = [vp, Vector]
    lib.vector_db_update.argtypes = [vp, C.c_size_t, Vector]
    lib.vector_db_delete.argtypes = [vp, C.c_size_t]
    lib.vector_db_save.argtypes = [vp, C.c_char_p]
    lib.vector_db_free.argtypes = [vp]
    return lib


class Quiet:
    """Send fd 1 to /dev/null: the reference prints on every kd insert level."""

    def __enter__(self):
        sys.stdout.flush()
        self.libc = C.CDLL(None)
        self.libc.fflush(None)
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *a):
        self.libc.fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def kd_answers(lib, X, Q, D):
    with Quiet():
        t = lib.kdtree_create(D)
        for i in range(X.shape[0]):
            lib.kdtree_insert(t, dptr(X[i]), i)
        out = np.array([lib.kdtree_nearest(t, dptr(Q[j])) for j in range(Q.shape[0])],
                       dtype=np.uint64)
        lib.kdtree_free(t)
    return out


def vec(a, uuid=b"u"):
    v = Vector()
    v.uuid = uuid
    v.dimension = a.shape[0]
    v.data = dptr(a)
    return v


def owned_vec(lib_c, a, uuid):
    """A Vector whose data is malloc'ed, as vector_db_insert takes ownership."""
    buf = lib_c.malloc(a.nbytes)
    C.memmove(buf, a.ctypes.data, a.nbytes)
    v = Vector()
    v.uuid = uuid
    v.dimension = a.shape[0]
    v.data = C.cast(buf, C.POINTER(C.c_double))
    return v


def metric_bits(lib, A, B):
    out = np.zeros((A.shape[0], 3), dtype=np.float32)
    for p in range(A.shape[0]):
        va, vb = vec(A[p]), vec(B[p])
        out[p, 0] = lib.cosine_similarity(va, vb)
        out[p, 1] = lib.euclidean_distance(va, vb)
        out[p, 2] = lib.dot_product(va, vb)
    return out.view(np.uint32)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    lib = load_ref()
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    fx = {}

    # ---- C1 -------------------------------------------------------------
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (10000, 128))
    Q = rng.uniform(-1, 1, (256, 128))
    P = rng.integers(0, 10000, (256, 2))
    fx["c1_db_sha256"] = np.array(sha(X))
    fx["c1_q_sha256"] = np.array(sha(Q))
    fx["c1_pairs"] = P.astype(np.uint64)
    fx["c1_kd3_ids"] = kd_answers(lib, X, Q, 3)
    fx["c1_kd128_ids"] = kd_answers(lib, X, Q, 128)
    fx["c1_metric_bits"] = metric_bits(lib, X[P[:, 0]], X[P[:, 1]])  # cos, l2, dot

    # ---- duplicate rows: which of several equal-distance rows is returned -
    # Rows 0..5 of a 3-d set hold two copies each of three points; queries
    # hit each copy pair exactly.  The reference keeps the first node its DFS
    # visits, which need not be the lowest index.
    T = np.array([[0.5, 0.5, 0.5], [0.1, 0.9, 0.3], [0.5, 0.5, 0.5],
                  [0.7, 0.2, 0.8], [0.1, 0.9, 0.3], [0.7, 0.2, 0.8],
                  [0.3, 0.3, 0.3], [0.9, 0.1, 0.6]], dtype=np.float64)
    TQ = np.array([[0.5, 0.5, 0.5], [0.1, 0.9, 0.3], [0.7, 0.2, 0.8],
                   [0.52, 0.5, 0.5]], dtype=np.float64)
    fx["tie_db"] = T
    fx["tie_q"] = TQ
    fx["tie_ref_ids"] = kd_answers(lib, T, TQ, 3)
    # Equidistant distinct points (integer grid): the query is exactly
    # between rows 1 and 3 (and rows 0 and 2) in squared distance.
    G = np.array([[2.0, 0.0], [0.0, 1.0], [-2.0, 0.0], [0.0, -1.0], [5.0, 5.0]])
    GQ = np.array([[0.0, 0.0], [0.0, 0.5]])
    fx["grid_db"] = G
    fx["grid_q"] = GQ
    fx["grid_ref_ids"] = kd_answers(lib, G, GQ, 2)
    # Rows 1 and 2 are both at squared distance 9 from the query but lie on
    # opposite sides of the root; the search descends into row 2's side
    # first, so the reference answers 2 where the lowest-id rule answers 1.
    H = np.array([[0.0, 0.0], [-2.5, 8.0], [3.5, 8.0]])
    HQ = np.array([[0.5, 8.0]])
    fx["dfs_db"] = H
    fx["dfs_q"] = HQ
    fx["dfs_ref_ids"] = kd_answers(lib, H, HQ, 2)

    # ---- stale tree after update / delete --------------------------------
    S = np.array([[0.0, 0.0], [10.0, 10.0], [20.0, 20.0], [30.0, 30.0]])
    with Quiet():
        db = lib.vector_db_init(0, 2)
        for i in range(4):
            lib.vector_db_insert(db, owned_vec(libc, S[i], b"s%d" % i))
        lib.vector_db_update(db, 1, owned_vec(libc, np.array([100.0, 100.0]), b"upd"))
        lib.vector_db_delete(db, 0)
        kdt = C.cast(db + 24, C.POINTER(C.c_void_p)).contents.value  # VectorDatabase.kdtree @24
        sq = np.array([[10.0, 10.0], [100.0, 100.0], [30.0, 30.0]])
        stale = np.array([lib.kdtree_nearest(kdt, dptr(sq[j])) for j in range(3)],
                         dtype=np.uint64)
        lib.vector_db_free(db)
    fx["stale_rows"] = S
    fx["stale_q"] = sq
    fx["stale_ref_ids"] = stale

    # ---- special metric inputs ----------------------------------------------
    z = np.zeros(128)
    o = np.ones(128)
    w = np.linspace(-3, 3, 128)
    specA = np.stack([z, o, w, w, o])
    specB = np.stack([o, o, w, -w, z])
    fx["spec_a"] = specA
    fx["spec_b"] = specB
    fx["spec_metric_bits"] = metric_bits(lib, specA, specB)
    # Mismatched dimensions → -1.0f (vector_database.c:302-305,323-326,343-346)
    va, vb = vec(o[:5].copy()), vec(o[:4].copy())
    fx["mismatch_bits"] = np.array([lib.cosine_similarity(va, vb),
                                    lib.euclidean_distance(va, vb),
                                    lib.dot_product(va, vb)], dtype=np.float32).view(np.uint32)

    # ---- persistence format written by vector_db_save ----------------------
    rng2 = np.random.default_rng(7)
    R = rng2.standard_normal((5, 16))
    uu = [b"%08d-0000-4000-8000-%012d" % (i, i) for i in range(5)]
    with tempfile.TemporaryDirectory() as td, Quiet():
        db = lib.vector_db_init(0, 3)
        for i in range(5):
            lib.vector_db_insert(db, owned_vec(libc, R[i], uu[i]))
        fn = os.path.join(td, "save.db")
        lib.vector_db_save(db, fn.encode())
        lib.vector_db_free(db)
        blob = open(fn, "rb").read()
    fx["save_rows"] = R
    fx["save_uuids"] = np.array(uu)
    with open(os.path.join(HERE, "ref_save_small.bin"), "wb") as f:
        f.write(blob)

    np.savez_compressed(os.path.join(HERE, "c1_reference.npz"), **fx)
    print("kd3 ids[:8]", fx["c1_kd3_ids"][:8])
    print("kd128 ids[:8]", fx["c1_kd128_ids"][:8])
    print("tie ids", fx["tie_ref_ids"], "grid ids", fx["grid_ref_ids"],
          "dfs ids", fx["dfs_ref_ids"],
          "stale ids", fx["stale_ref_ids"])
    print("wrote", os.path.join(HERE, "c1_reference.npz"), "and ref_save_small.bin",
          len(blob), "bytes")


if __name__ == "__main__":
    main()
