"""The C-ABI library loads and exports every symbol include/pygw_hip.h
declares (no device calls)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "pygw_hip.h")
LIB = os.path.join(ROOT, "pygw_amd", "libpygw_hip.so")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pygw_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "pygw_tetra_weights" in names and "pygw_version" in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_all_declared():
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    from pygw_amd import _lib
    assert sorted(_lib.exported_symbols()) == declared()
