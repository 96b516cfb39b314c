"""The C-ABI library loads and exports every symbol include/ippi_lbm.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from tests import _pkg

HDR = os.path.join(_pkg.ROOT, "include", "ippi_lbm.h")


def declared():
    src = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ippi_\w+)\s*\(", src, re.M)))


def test_header_lists_entry_points():
    names = declared()
    assert "ippi_step" in names and "ippi_create" in names and len(names) >= 20


def test_library_exports_all_declared_symbols():
    ippi = _pkg.load()
    L = ctypes.CDLL(ippi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    ippi = _pkg.load()
    assert set(declared()) == set(ippi._SIGS)


def test_config_struct_layout():
    ippi = _pkg.load()
    # 7 ints (28 B, padded to 32), 4 doubles, 2 ints, 1 double
    assert ctypes.sizeof(ippi.Config) == 80
    assert ippi.library().ippi_version() == 1
