"""The C-ABI library loads and exports every entry point include/*.h declares.

Runs without a GPU: it only resolves symbols and calls functions that make no
HIP call.
"""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "raja_amd", "libraja_cdna4.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(raja_cdna4_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "raja_cdna4_sort_pairs" in names
    assert "raja_cdna4_exclusive_scan" in names
    assert "raja_cdna4_reduce" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libraja_cdna4.so is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    L.raja_cdna4_version.restype = ctypes.c_char_p
    assert b"gfx950" in L.raja_cdna4_version()
    L.raja_cdna4_last_error.restype = ctypes.c_char_p
    assert L.raja_cdna4_last_error() is not None


def test_library_is_gfx950_only():
    data = open(LIB, "rb").read()
    assert b"gfx950" in data
    for other in (b"gfx942", b"gfx90a", b"gfx1100"):
        assert other not in data
