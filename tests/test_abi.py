"""The C-ABI library loads and exports every entry point its headers declare, and the ctypes
table binds each of them (no GPU needed: nothing is called)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in ("pylith_amd.h", "pylith_amd_debug.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(plx_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_entry_point():
    from pylith_amd import _lib
    so = os.path.join(ROOT, "pylith_amd", "libpylith_amd.so")
    if not os.path.exists(so):
        pytest.skip("library not built")
    lib = ctypes.CDLL(so)
    names = declared()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    unbound = [n for n in sorted(names) if n not in _lib.SIGNATURES]
    assert not unbound, unbound
