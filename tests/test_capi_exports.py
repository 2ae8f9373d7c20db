"""The C-ABI libraries load and export every symbol include/ttfort_amd.h
declares (CPU only: no compute calls are made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ttfort_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


def test_header_declares_core_entry_points():
    names = declared_functions()
    for n in ("ttf_init", "ttf_dbfun3", "ttf_dphi_left", "ttf_dphi_right", "dbfun3"):
        assert n in names


@pytest.mark.parametrize("kind", ["i4", "i8"])
def test_library_exports_all_declared_symbols(kind):
    import ttfort
    path = ttfort.lib_path(kind)
    assert os.path.exists(path), "library not built: " + path
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, "symbols declared in the header but not exported: %s" % missing


def test_binding_covers_header():
    import ttfort
    bound = set(ttfort.exported_symbols())
    missing = [n for n in declared_functions() if n not in bound]
    assert not missing, "header functions without a ctypes signature: %s" % missing
