"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/*.h declares, and keeps the reference's struct layouts."""
import ctypes as C
import os
import re

import pytest

import svdb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("vector_database.h", "kdtree.h", "svdb.h", "svdb_nearest.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", src, flags=re.M):
            name = m.group(1)
            if name not in ("if", "sizeof") and not name.isupper():
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    L = svdb.lib()
    decl = declared_functions()
    assert decl == set(svdb.EXPORTS), decl ^ set(svdb.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name


def test_struct_layouts_match_reference():
    # vector_database.h:14-18, :24-30; kdtree.h:21-24 (LP64)
    assert C.sizeof(svdb.Vector) == 56
    assert svdb.Vector.dimension.offset == 40 and svdb.Vector.data.offset == 48
    assert svdb.VectorDatabase.size.offset == 8
    assert svdb.VectorDatabase.capacity.offset == 16
    assert svdb.VectorDatabase.kdtree.offset == 24
    assert svdb.VectorDatabase.mutex.offset == 32
    assert svdb.KDTree.dimension.offset == 8


def test_no_gpu_fails_loudly():
    # In a container without a GPU the engine must refuse to start rather than
    # fall back to host code.
    p = svdb.lib().svdb_device_alloc(8)
    if p:
        svdb.lib().svdb_device_free(p)
        pytest.skip("GPU present")
    assert not svdb.lib().vector_db_init(0, 3)
    assert not svdb.lib().kdtree_create(3)
    assert b"HIP device" in svdb.lib().svdb_last_error() or svdb.lib().svdb_last_error()


def test_init_refuses_an_overflowing_capacity():
    # capacity * sizeof(Vector) would wrap: refused before any allocation
    # (vector_database.c:17-54 has the same multiplication unchecked)
    assert not svdb.lib().vector_db_init((1 << 64) // 32, 3)


def test_mfma_sustained_refuses_bad_arguments():
    # the kind and the output are checked before the device is touched
    v = C.c_double(-1.0)
    assert svdb.lib().svdb_mfma_sustained(3, C.byref(v)) == -1   # SVDB_EINVAL
    assert svdb.lib().svdb_mfma_sustained(-1, C.byref(v)) == -1
    assert svdb.lib().svdb_mfma_sustained(0, None) == -1
    assert b"svdb_mfma_sustained" in svdb.lib().svdb_last_error()
