"""CPU oracle of the clspmv engine -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py (CPU baseline leg) as
the checker; the engine itself never imports it.

* ``Oracle`` wraps oracle/liboracle.so, the C restatement of the reference's
  serial SpMV (include/spmv_serial.h) plus an FP64 checker and the
  reference's comparison report (src/util.cpp:93-131).
* ``Reference`` wraps oracle/_ref/libclspmv_ref.so, the reference's own CPU
  code compiled from its sources (built only where the reference tree
  exists; the built library travels with the tree).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libclspmv_ref.so")

PI = C.POINTER(C.c_int)
PF = C.POINTER(C.c_float)
PD = C.POINTER(C.c_double)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(PI)


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(PF)


class Oracle:
    def __init__(self, path: str = ORACLE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run make -C oracle")
        self.lib = C.CDLL(path)

    # FP64 reference and tolerance bound
    def spmv_f64(self, row, col, val, x, height, y0=None):
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        y64 = np.zeros(height, np.float64)
        bound = np.zeros(height, np.float64)
        y0p = None
        if y0 is not None:
            y0 = np.ascontiguousarray(y0, np.float32)
            y0p = _f(y0)
        self.lib.oracle_spmv_f64(C.c_int(row.size), _i(row), _i(col), _f(val), _f(x), C.c_int(height),
                                 y0p, y64.ctypes.data_as(PD), bound.ctypes.data_as(PD))
        return y64, bound

    def scaled_error(self, y, y64, bound):
        y = np.ascontiguousarray(y, np.float32)
        worst = C.c_int()
        self.lib.oracle_scaled_error.restype = C.c_double
        e = self.lib.oracle_scaled_error(C.c_int(y.size), _f(y), y64.ctypes.data_as(PD),
                                         bound.ctypes.data_as(PD), C.byref(worst))
        return e, worst.value

    def two_vec_compare(self, a, b):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        out = np.zeros(5, np.float64)
        self.lib.oracle_two_vec_compare(_f(a), _f(b), C.c_int(a.size), out.ctypes.data_as(PD))
        return dict(distance=out[0], max_diff=out[1], max_ratio=out[2], max_id=int(out[3]),
                    differ=int(out[4]))

    # FP32 serial SpMV in the reference's loop order (y updated in place)
    def coo_spmv(self, row, col, val, x, y):
        self.lib.oracle_coo_spmv(C.c_int(len(row)), _i(row), _i(col), _f(val), _f(x), _f(y))

    def csr_spmv(self, rp, col, val, x, y):
        self.lib.oracle_csr_spmv(C.c_int(len(rp) - 1), _i(rp), _i(col), _f(val), _f(x), _f(y))

    def csr_spmv_threads(self, threads, rp, col, val, x, y):
        self.lib.oracle_csr_spmv_threads(C.c_int(threads), C.c_int(len(rp) - 1), _i(rp), _i(col), _f(val),
                                         _f(x), _f(y))

    def max_threads(self) -> int:
        return int(self.lib.oracle_max_threads())

    def ell_spmv(self, rows, hal, ellnum, col, val, x, y):
        self.lib.oracle_ell_spmv(rows, hal, ellnum, _i(col), _f(val), _f(x), _f(y))

    def sell_spmv(self, rows, h, ptr, col, val, x, y):
        self.lib.oracle_sell_spmv(rows, h, len(ptr) - 1, _i(ptr), _i(col), _f(val), _f(x), _f(y))

    def dia_spmv(self, length, off, data, x, y):
        self.lib.oracle_dia_spmv(length, len(off), _i(off), _f(data), _f(x), len(x), _f(y))

    def bdia_spmv(self, length, boff, bptr, data, x, y):
        self.lib.oracle_bdia_spmv(length, len(boff), _i(boff), _i(bptr), _f(data), _f(x), len(x), _f(y))

    def bcsr_spmv(self, rows, cols, bh, bw, aligned, rp, bcol, data, x, y):
        self.lib.oracle_bcsr_spmv(rows, cols, len(rp) - 1, bh, bw, aligned, _i(rp), _i(bcol), _f(data),
                                  _f(x), _f(y))

    def bell_spmv(self, brows, bh, bw, hal, ellnum, bcol, data, x, y):
        self.lib.oracle_bell_spmv(brows, bh, bw, hal, ellnum, _i(bcol), _f(data), _f(x), _f(y))

    def sbell_spmv(self, rows, cols, bh, bw, sh, sptr, col4, data, x, y):
        self.lib.oracle_sbell_spmv(rows, cols, bh, bw, sh, len(sptr) - 1, _i(sptr), _i(col4), _f(data),
                                   _f(x), _f(y))

    def b4csr_spmv(self, rows, cols, bh, bw, aligned, rp, col4, data, x, y):
        self.lib.oracle_b4csr_spmv(rows, cols, len(rp) - 1, bh, bw, aligned, _i(rp), _i(col4), _f(data),
                                   _f(x), _f(y))

    def b4ell_spmv(self, rows, cols, brows, bh, bw, hal, f4, ellnum, col4, data, x, y):
        self.lib.oracle_b4ell_spmv(rows, cols, brows, bh, bw, hal, f4, ellnum, _i(col4), _f(data), _f(x),
                                   _f(y))


def reference_available() -> bool:
    return os.path.exists(REF_LIB)


class Reference:
    """The reference's own CPU routines (include/matrix_storage.h,
    include/spmv_serial.h, src/analyze.cpp) behind oracle/ref_harness.cpp."""

    def __init__(self, path: str = REF_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing (needs the reference tree): make -C oracle ref")
        self.lib = C.CDLL(path)
        import clspmv_mic_amd as cs
        self.cs = cs
        P = C.POINTER
        L = self.lib
        L.ref_analyze_matrix.argtypes = [P(cs.CooMatrix), P(cs.CocktailMatrix), C.c_float, C.c_int,
                                         C.c_int, C.c_char_p]
        L.ref_free_cocktail.argtypes = [P(cs.CocktailMatrix)]

    def analyze(self, coo, ben_dir: str, mult=1.0, partial=True, full=False):
        """The reference's analyze_matrix; returns a CocktailMatrix whose arrays
        belong to the reference (release with free_cocktail)."""
        c = self.cs.CocktailMatrix()
        self.lib.ref_analyze_matrix(C.byref(coo.c), C.byref(c), C.c_float(mult), int(partial),
                                    int(full), ben_dir.encode())
        return c

    def free_cocktail(self, c):
        self.lib.ref_free_cocktail(C.byref(c))

    def convert(self, name: str, coo, *args):
        """ref_coo2<name>(coo, dst, *args); returns (status, struct)."""
        cs = self.cs
        T = {"csr": cs.CsrMatrix, "ell": cs.EllMatrix, "sell": cs.SellMatrix, "dia": cs.DiaMatrix,
             "diaext": cs.DiaExtMatrix,
             "bdia": cs.BdiaMatrix, "bcsr": cs.BcsrMatrix, "b4csr": cs.B4csrMatrix,
             "bell": cs.BellMatrix, "b4ell": cs.B4ellMatrix, "sbell": cs.SbellMatrix}[name]
        dst = T()
        f = getattr(self.lib, f"ref_coo2{name}")
        f.restype = C.c_int
        # Block shapes and slice heights are unsigned, alignments and widths int.
        unsigned = {"sell": 1, "bcsr": 2, "b4csr": 2, "bell": 2, "b4ell": 2, "sbell": 3}.get(name, 0)
        conv = [C.c_uint(a) if i < unsigned else C.c_int(a) for i, a in enumerate(args)]
        rc = f(C.byref(coo.c), C.byref(dst), *conv)
        return rc, dst

    def to_coo(self, name: str, struct):
        """ref_<name>2coo(struct, dst): the reference's inverse conversion;
        returns the COO arrays (copied) as (row, col, val, height, width)."""
        cs = self.cs
        dst = cs.CooMatrix()
        getattr(self.lib, f"ref_{name}2coo")(C.byref(struct), C.byref(dst))
        n = dst.matinfo.nnz
        row = np.ctypeslib.as_array(dst.coo_row_id, shape=(n,)).copy() if n else np.zeros(0, np.int32)
        col = np.ctypeslib.as_array(dst.coo_col_id, shape=(n,)).copy() if n else np.zeros(0, np.int32)
        val = np.ctypeslib.as_array(dst.coo_data, shape=(n,)).copy() if n else np.zeros(0, np.float32)
        self.lib.ref_free_coo(C.byref(dst))
        return row, col, val, dst.matinfo.height, dst.matinfo.width

    def spmv(self, name: str, struct, x, y):
        """The reference's serial <name>_spmv (include/spmv_serial.h) on a struct."""
        f = getattr(self.lib, f"ref_{name}_spmv")
        f.restype = None
        f(C.byref(struct), x.ctypes.data_as(PF), y.ctypes.data_as(PF))
