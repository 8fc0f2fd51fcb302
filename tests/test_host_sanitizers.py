"""The host-only exact layers (e-hgp_amd/host/exact.cpp: BigInt, ExactCloud,
ExactSupport; e-hgp_amd/host/certify.cpp: LBVH snapshot certification) built
with g++ under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer, and driven by tests/native/sanitize_exact.cpp: 128-bit
integer cross-checks of the multiprecision arithmetic, exact support
analysis against the lifted insphere determinant and the in-plane
circumcentre on an integer grid, and the same supports on eight threads
sharing one lazily converted cloud (the pattern of the host exact runs).
Any sanitizer report or failed check fails the test.  The reference
sanitizes its CPU targets the same way (CMakeLists.txt:32-71)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from helpers import PKG, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(HERE, "native", "sanitize_exact.cpp"),
           os.path.join(PKG, "host", "exact.cpp"), os.path.join(PKG, "host", "certify.cpp")]
FLAGS = {
    "asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
    "tsan": ["-fsanitize=thread"],
}

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    out = tmp_path_factory.mktemp("sanitize")
    exes = {}
    for name, flags in FLAGS.items():
        exe = str(out / name)
        subprocess.run(["g++", "-std=c++17", "-O1", "-g", *flags, *SOURCES, "-o", exe, "-pthread"],
                       check=True, capture_output=True)
        exes[name] = exe
    return exes


def _run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               TSAN_OPTIONS="halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]


def test_bigint_arithmetic_asan_ubsan(drivers):
    _run(drivers["asan"], "arith")


@pytest.mark.parametrize("threads", [1, 8])
def test_exact_supports_asan_ubsan(drivers, threads):
    _run(drivers["asan"], "supports", str(threads))


def test_exact_supports_threads_tsan(drivers):
    _run(drivers["tsan"], "supports", "8")


@pytest.mark.parametrize("name,pts", [
    ("uniform300", workloads.uniform_cloud(300, seed=5)),
    ("lattice2k", workloads.jittered_lattice(2000, side=13, seed=6)),
    ("collisions", np.array([[0, 0, 0], [0, 0, 1e-9], [0, 1e-9, 0], [1, 1, 1], [1, 1, 1 - 1e-12], [0.5, 0.5, 0.5]])),
])
def test_certify_snapshot_asan_ubsan(drivers, tmp_path, name, pts):
    c = oracle.canonicalize(pts)
    s = oracle.lbvh_snapshot(c)
    path = tmp_path / f"{name}.bin"
    with open(path, "wb") as fh:
        np.array([len(c)], dtype=np.int64).tofile(fh)
        for a, t in ((c, np.float64), (s.leaves, np.uint64), (s.nodes, np.uint64), (s.counters, np.uint64)):
            np.ascontiguousarray(a, dtype=t).tofile(fh)
    _run(drivers["asan"], "certify", str(path))
