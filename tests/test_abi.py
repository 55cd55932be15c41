"""The C-ABI library loads and exports every entry point include/*.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "seismic_tomography_3d_amd", "libst_hip.so")


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(st_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    assert "st_fim_solve" in names and "st_forward" in names and "st_optimize" in names
    assert len(names) >= 19


@pytest.mark.skipif(not os.path.exists(LIB), reason="libst_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libst_hip.so not built")
def test_python_binding_covers_header():
    from seismic_tomography_3d_amd import _lib
    assert set(declared()) <= set(_lib.exported_names())
