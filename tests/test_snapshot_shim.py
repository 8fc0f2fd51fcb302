"""The reference-side binding of the LBVH snapshot: a C++ translation unit
(tests/native/snapshot_shim.cpp) that mirrors the reference's 16-byte leaf
and 80-byte node records with static_asserts and fills them straight from
mhgp_lbvh_snapshot.  It compiles and links against libmhgp.so on the CPU;
the GPU test runs it and it checks the records structurally."""

import os
import subprocess

import pytest

from helpers import PKG, ROOT

HERE = os.path.dirname(os.path.abspath(__file__))


def _build(out):
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "native", "snapshot_shim.cpp"), "-L", PKG, "-lmhgp", "-Wl,-rpath," + PKG,
           "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)


def test_shim_compiles_and_links(tmp_path):
    _build(tmp_path / "snapshot_shim")
    assert (tmp_path / "snapshot_shim").exists()


@pytest.mark.gpu
def test_shim_snapshot_records_on_gpu(tmp_path):
    exe = tmp_path / "snapshot_shim"
    _build(exe)
    for n in (2, 257, 50_000):
        print(f"snapshot shim n={n}", flush=True)
        r = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
