"""A C translation unit against the boundary (the handlers' row of the scope).

CPU: every header in include/ compiles on its own as C99 with -Wall
-Werror, and tests/c_caller/handlers_replay.c — the reference handlers'
call sequences written as a C handler would write them — compiles and links
against libsvdb.so; without a GPU it stops where vector_db_init returns
NULL (exit 3), as the product must rather than compute on the host.

GPU: the prebuilt handlers_replay runs those sequences on the C1 workload
(10,000 x 128 uniform rows, seed 0, kd-tree dimension 3, 256 queries, 256
pairs) and its answers are compared with the reference's own
(tests/golden/c1_reference.npz: kd ids at D=3, the float /compare bits)
and, after PUT and DELETE, with the FP64 oracle on the mutated rows.
"""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "simple-vector-db_amd")
SRC = os.path.join(ROOT, "tests", "c_caller", "handlers_replay.c")
BIN = os.path.join(ROOT, "tests", "c_caller", "handlers_replay")

HEADERS = sorted(h for h in os.listdir(INC) if h.endswith(".h"))


@pytest.mark.parametrize("header", HEADERS)
def test_header_compiles_alone_as_c99(header, tmp_path):
    src = tmp_path / "tu.c"
    src.write_text(f'#include "{header}"\nint main(void) {{ return 0; }}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
                        "-fsyntax-only", "-I", INC, str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_handler_replay_compiles_and_links(tmp_path):
    out = tmp_path / "replay"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pthread", "-I", INC,
                        "-o", str(out), SRC, "-L", LIBDIR, "-lsvdb", f"-Wl,-rpath,{LIBDIR}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    gpu = subprocess.run(["python", "-c", "import svdb, sys; p = svdb.lib().svdb_device_alloc(8); "
                          "sys.exit(0 if p else 1)"], cwd=LIBDIR, capture_output=True)
    if gpu.returncode == 0:
        pytest.skip("GPU present: the GPU test runs the replay")
    inp = tmp_path / "in.bin"
    inp.write_bytes(np.array([4, 2, 2, 1, 0], dtype=np.uint64).tobytes() +
                    np.zeros(10).tobytes() + np.zeros(2, dtype=np.uint64).tobytes())
    r = subprocess.run([str(out), str(inp), str(tmp_path / "o"), str(tmp_path / "db")],
                       capture_output=True, text=True)
    assert r.returncode == 3  # vector_db_init refused: no GPU, no host fallback


def _write_input(path, X, Q, P, kd, updates, deletes):
    N, d = X.shape
    with open(path, "wb") as f:
        f.write(np.array([N, d, kd, Q.shape[0], P.shape[0]], dtype=np.uint64).tobytes())
        f.write(np.ascontiguousarray(X, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(Q, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(P, dtype=np.uint64).tobytes())
        f.write(np.uint64(len(updates)).tobytes())
        for i, row in updates:
            f.write(np.uint64(i).tobytes() + np.asarray(row, dtype=np.float64).tobytes())
        f.write(np.uint64(len(deletes)).tobytes())
        f.write(np.asarray(deletes, dtype=np.uint64).tobytes())


@pytest.mark.gpu
def test_handler_sequences_on_c1(c1_data, golden):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built: run __graft_entry__.build()")
    X, Q, P = c1_data
    rng = np.random.default_rng(11)
    updates = [(int(i), rng.uniform(-1, 1, 128)) for i in rng.integers(0, 10000, 8)]
    deletes = [int(i) for i in rng.integers(0, 9000, 8)] + [20000]  # the last is out of range
    model = X.copy()
    for i, row in updates:
        model[i] = row
    for i in deletes:
        if i < model.shape[0]:
            model = np.delete(model, i, axis=0)
    tmp = tempfile.mkdtemp()
    try:
        inp, outp, dbf = (os.path.join(tmp, n) for n in ("in.bin", "out.bin", "db.bin"))
        _write_input(inp, X, Q, P, 3, updates, deletes)
        print("handlers_replay on C1 (10000x128, kd 3, 256 queries, 256 pairs)", flush=True)
        r = subprocess.run([BIN, inp, outp, dbf], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
        raw = open(outp, "rb").read()
    finally:
        shutil.rmtree(tmp)
    nq, npairs = Q.shape[0], P.shape[0]
    o = 0
    kd_ids = np.frombuffer(raw, np.uint64, nq, o); o += 8 * nq
    mbits = np.frombuffer(raw, np.uint32, 3 * npairs, o).reshape(npairs, 3); o += 12 * npairs
    after = np.frombuffer(raw, np.uint64, nq, o); o += 8 * nq
    size_after, loaded, jlen = np.frombuffer(raw, np.uint64, 3, o); o += 24
    text = raw[o:o + int(jlen)].decode()
    # the reference's own answers on the same inputs
    assert np.array_equal(kd_ids, golden["c1_kd3_ids"])
    assert np.array_equal(mbits, golden["c1_metric_bits"])
    # after PUT and DELETE: the current rows (the reference's tree would be
    # stale here, a documented divergence)
    assert size_after == loaded == model.shape[0]
    rid, _ = O.topk(model, Q, 1, "l2", D=3)
    assert np.array_equal(after, rid[:, 0])
    # the batched /nearest layer: two queries, top-3 by dot product
    ans = json.loads(text)
    rid, rsc = O.topk(model, Q[:2], 3, "dot")
    assert len(ans) == 2
    for qi in range(2):
        assert [e["index"] for e in ans[qi]] == [int(v) for v in rid[qi]]
        for j, e in enumerate(ans[qi]):
            # numbers are printed as cJSON prints them: %1.15g when that reads
            # back within DBL_EPSILON relative, else %1.17g
            assert abs(e["dot_product"] - rsc[qi, j]) <= 2.3e-16 * abs(rsc[qi, j])
            np.testing.assert_allclose(e["vector"], model[e["index"]], rtol=2.3e-16, atol=0)
            assert e["uuid"].startswith(("row-", "upd-"))
