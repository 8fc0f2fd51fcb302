"""Level-class links of the sorted catalogue (e-hgp_amd/csrc/level_links.hpp):
a record starts a new exact class only when its interval lies above every
earlier interval, not only above its neighbour's."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from helpers import PKG

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("links") / "liblinks.so"
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-fPIC", "-shared", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
           "-x", "c++", "-I", os.path.join(PKG, "csrc"), "-I/opt/rocm/include",
           os.path.join(HERE, "native", "links_check.cpp"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    lib = ctypes.CDLL(str(out))
    lib.links_check.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p]
    return lib


def _links(lib, levels, errs):
    lv = np.ascontiguousarray(levels, dtype=np.float64)
    er = np.ascontiguousarray(errs, dtype=np.float32)
    out = np.zeros(len(lv), dtype=np.uint32)
    lib.links_check(lv.ctypes.data, er.ctypes.data, len(lv), out.ctypes.data)
    return out.tolist()


def test_wide_earlier_interval_keeps_a_later_record_in_its_run(lib):
    # x = 1.0 +- 2e-2 reaches 1.02; y = 1.001 and z = 1.01 (both +- 2e-12)
    # lie inside it.  z lies above its neighbour y, so a neighbour-only test
    # would start a new class at z although x may still be larger than z.
    links = _links(lib, [1.0, 1.001, 1.01], [1e-2, 1e-12, 1e-12])
    assert links == [1, 0, 0]


def test_disjoint_intervals_link(lib):
    links = _links(lib, [1.0, 2.0, 3.0], [1e-12, 1e-12, 1e-12])
    assert links == [1, 1, 1]


def test_prefix_maximum_moves_on(lib):
    # the wide record's reach ends before the fourth record
    links = _links(lib, [1.0, 1.001, 1.01, 1.5], [1e-2, 1e-12, 1e-12, 1e-12])
    assert links == [1, 0, 0, 1]


def test_randomised_against_quadratic_rule(lib):
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = int(rng.integers(2, 40))
        lv = np.sort(rng.random(n) + 1.0)
        er = (10.0 ** rng.uniform(-12, -1, size=n)).astype(np.float32)
        got = _links(lib, lv, er)
        up = lv * (1.0 + 2.0 * er.astype(np.float64))
        lo = lv * (1.0 - 2.0 * er.astype(np.float64))
        want = [1] + [int(all(lo[i] > up[j] for j in range(i))) for i in range(1, n)]
        assert got == want
