"""The C-ABI library loads and exports every function include/*.h declares."""

import ctypes
import glob
import os
import re

import clspmv_mic_amd as cs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(clspmv_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_functions():
    names = declared_functions()
    assert len(names) > 60
    assert "clspmv_spmv" in names and "clspmv_coo2sbell" in names


def test_library_exports_every_declared_symbol():
    missing = [n for n in sorted(declared_functions()) if not hasattr(cs.lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_struct_sizes_match_c_layout():
    # 3 ints of matinfo + pointers, as in the reference's templates.
    assert ctypes.sizeof(cs.MatInfo) == 12
    assert ctypes.sizeof(cs.CooMatrix) == 40
    assert ctypes.sizeof(cs.CsrMatrix) == 40
    assert ctypes.sizeof(cs.SellMatrix) == 48
    assert ctypes.sizeof(cs.B4ellMatrix) == 56


def test_error_text_is_reported():
    A = cs.Coo([5], [0], [1.0], height=3, width=3)  # row 5 outside 3 rows
    rc = cs.lib.clspmv_validate_coo(A.ref())
    assert rc == cs.ERR_INVALID
    assert b"outside" in cs.lib.clspmv_last_error()
