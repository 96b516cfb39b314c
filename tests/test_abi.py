"""The C-ABI library loads and exports every symbol include/ippi_lbm.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

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


def test_config_struct_layout(tmp_path):
    """The ctypes mirror of struct ippi_config has the C layout: sizes and
    field offsets from a C program compiled against the header."""
    ippi = _pkg.load()
    names = [f[0] for f in ippi.Config._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"ippi_lbm.h\"\nint main(void) {\n"
                   "  printf(\"%zu\\n\", sizeof(ippi_config));\n" +
                   "".join(f"  printf(\"%zu\\n\", offsetof(ippi_config, {n}));\n" for n in names) + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(_pkg.ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(ippi.Config)
    assert out[1:] == [getattr(ippi.Config, n).offset for n in names]
    assert ippi.library().ippi_version() == 1
