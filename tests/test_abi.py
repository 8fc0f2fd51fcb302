"""The C ABI library loads and exports every symbol declared in include/mhgp.h."""

import ctypes
import os
import re

from helpers import ROOT, mhgp


def test_header_symbols_exported():
    with open(os.path.join(ROOT, "include", "mhgp.h"), encoding="utf-8") as fh:
        text = fh.read()
    declared = set(re.findall(r"\b(mhgp_[a-z_]+)\s*\(", text))
    assert declared == set(mhgp.EXPORTED)
    lib = ctypes.CDLL(mhgp.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
