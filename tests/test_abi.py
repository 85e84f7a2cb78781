"""The C ABI: librbh.so loads and exports every entry point include/*.h declares."""
import os
import re

import pytest

from rainbow_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for name in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if name.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", name)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            syms.update(re.findall(r"\b(rbh_\w+)\s*\(", txt))
    return sorted(syms)


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "rbh_step" in syms and "rbh_sdf_build" in syms and len(syms) >= 20


def test_binding_covers_header():
    assert sorted(native.exported_symbols()) == declared_symbols()


@pytest.mark.skipif(not os.path.exists(native.LIB_PATH), reason="librbh.so not built")
def test_library_exports_every_symbol():
    import ctypes
    lib = ctypes.CDLL(native.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert native.load().rbh_version() == 1
