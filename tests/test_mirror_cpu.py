"""The host mirror's positional delete split over threads (svdb_mirror.h)
equals one memmove (vector_database.c:189-191) — a C++ model check on the
CPU, run with the split forced on for every size."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_parallel_shift_equals_memmove(tmp_path):
    exe = tmp_path / "mirror_shift_check"
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-Wall", "-Werror",
                        "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "simple-vector-db_amd", "csrc"),
                        "-o", str(exe), os.path.join(ROOT, "tests", "mirror_shift_check.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
