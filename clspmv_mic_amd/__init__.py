"""clspmv_mic_amd -- MI355X-native SpMV engine with the API surface of
clSpMV (bollig/clSpMV_MIC).

The engine is the C-ABI shared library ``libclspmv.so`` (declared in
``include/clspmv_formats.h``, ``include/clspmv_device.h`` and
``include/clspmv_cocktail.h``).  This module loads it with ctypes and offers
thin Python views of the same entry points for tests, benchmarks and
scripting:

* ``Coo`` and the other ``*Matrix`` ctypes structs are the C layout twins of
  the reference's ``coo_matrix<int,float>`` and friends
  (include/matrix_storage.h:15-365);
* ``coo2csr`` ... ``coo2sbell`` wrap the conversions (COO -> format);
* ``Context`` / ``DeviceMatrix`` / ``DeviceVector`` run SpMV on the GPU;
* ``analyze_matrix`` and ``Cocktail`` expose the multi-format partitioner
  and its runtime (include/analyze.h, include/spmv_cocktail.h).

There is no CPU fallback: every SpMV runs in the HIP library, and importing
this module fails when the library has not been built.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclspmv.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
        "or make -C clspmv_mic_amd)")

lib = C.CDLL(LIB_PATH)

OK = 0
ERR_INVALID, ERR_HIP, ERR_ALLOC, ERR_UNSUPPORTED, ERR_TOO_LARGE, ERR_IO = -1, -2, -3, -4, -5, -6
GPU_ALIGNMENT = 32
WARPSIZE = 32
SELL_GROUP_SIZE = 128
METHOD_AUTO = -1

FMT_CSR, FMT_COO, FMT_ELL, FMT_SELL, FMT_DIA, FMT_BDIA, FMT_BCSR, FMT_BELL, FMT_SBELL = range(9)
FORMAT_NAMES = ["csr", "coo", "ell", "sell", "dia", "bdia", "bcsr", "bell", "sbell"]
# Block shapes of the reference's blocked formats (src/analyze.cpp:545-560).
BLOCK_SHAPES = [(1, 4), (2, 4), (4, 4), (8, 4), (1, 8), (2, 8), (4, 8), (8, 8)]


class ClspmvError(RuntimeError):
    def __init__(self, code: int, what: str, msg: Optional[str] = None):
        if msg is None:  # the library's message for its own failures
            msg = lib.clspmv_last_error().decode(errors="replace")
        super().__init__(f"{what} failed with status {code}: {msg}")
        self.code = code


def check(code: int, what: str) -> int:
    if code < 0:
        raise ClspmvError(code, what)
    return code


# ---------------------------------------------------------------------------
# structs (include/clspmv_formats.h)

PI = C.POINTER(C.c_int)
PF = C.POINTER(C.c_float)


class MatInfo(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("nnz", C.c_int)]


class CooMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("coo_row_id", PI), ("coo_col_id", PI), ("coo_data", PF)]


class CsrMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("csr_row_ptr", PI), ("csr_col_id", PI), ("csr_data", PF)]


class EllMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("ell_num", C.c_int), ("ell_height_aligned", C.c_int),
                ("ell_col_id", PI), ("ell_data", PF)]


class SellMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("sell_slice_height", C.c_uint), ("sell_slice_num", C.c_int),
                ("sell_slice_ptr", PI), ("sell_col_id", PI), ("sell_data", PF)]


class DiaMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("dia_num", C.c_int), ("dia_length", C.c_int),
                ("dia_length_aligned", C.c_int), ("dia_offsets", PI), ("dia_data", PF)]


DiaExtMatrix = type("DiaExtMatrix", (C.Structure,), {"_fields_": DiaMatrix._fields_})


class BdiaMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("bdia_band_num", C.c_int), ("bdia_length", C.c_int),
                ("bdia_length_aligned", C.c_int), ("bdia_offsets", PI), ("bdia_bptr", PI),
                ("bdia_data", PF)]


class BcsrMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("bcsr_bwidth", C.c_uint), ("bcsr_bheight", C.c_uint),
                ("bcsr_row_num", C.c_int), ("bcsr_aligned_size", C.c_int), ("bcsr_block_num", C.c_int),
                ("bcsr_row_ptr", PI), ("bcsr_col_id", PI), ("bcsr_data", PF)]


class BellMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("bell_bwidth", C.c_uint), ("bell_bheight", C.c_uint),
                ("bell_row_num", C.c_int), ("bell_block_num", C.c_int), ("bell_height_aligned", C.c_int),
                ("bell_float4_aligned", C.c_int), ("bell_col_id", PI), ("bell_data", PF)]


class B4csrMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("b4csr_bwidth", C.c_uint), ("b4csr_bheight", C.c_uint),
                ("b4csr_row_num", C.c_int), ("b4csr_aligned_size", C.c_int), ("b4csr_block_num", C.c_int),
                ("b4csr_row_ptr", PI), ("b4csr_col_id", PI), ("b4csr_data", PF)]


class B4ellMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("b4ell_bwidth", C.c_uint), ("b4ell_bheight", C.c_uint),
                ("b4ell_row_num", C.c_int), ("b4ell_block_num", C.c_int), ("b4ell_height_aligned", C.c_int),
                ("b4ell_float4_aligned", C.c_int), ("b4ell_col_id", PI), ("b4ell_data", PF)]


class SbellMatrix(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("sbell_bwidth", C.c_uint), ("sbell_bheight", C.c_uint),
                ("sbell_slice_height", C.c_uint), ("sbell_slice_num", C.c_int), ("sbell_row_num", C.c_int),
                ("sbell_slice_ptr", PI), ("sbell_col_id", PI), ("sbell_data", PF)]


class CocktailMatrix(C.Structure):
    _fields_ = ([(f"ifuse{n}", C.c_bool) for n in
                 ["bdia", "dia", "sbell", "bell", "bcsr", "sell", "ell", "csr", "coo"]] +
                [(f"{n}_meth_num", C.c_int) for n in
                 ["bdia", "dia", "sbell", "bell", "bcsr", "sell", "ell", "csr", "coo"]] +
                [("csr_group_num", C.c_int), ("coo_group_num", C.c_int),
                 ("mat_width", C.c_int), ("mat_height", C.c_int), ("mat_nnz", C.c_int),
                 ("bdia", BdiaMatrix), ("dia", DiaMatrix), ("sbell", SbellMatrix),
                 ("bell", B4ellMatrix), ("bcsr", B4csrMatrix), ("sell", SellMatrix),
                 ("ell", EllMatrix), ("csr", CsrMatrix), ("coo", CooMatrix)])


class DmatInfo(C.Structure):
    _fields_ = [("format", C.c_int), ("rows", C.c_int), ("cols", C.c_int), ("nnz", C.c_int),
                ("stored", C.c_longlong), ("device_bytes", C.c_size_t), ("bheight", C.c_int),
                ("bwidth", C.c_int), ("slice_height", C.c_int), ("panel_width", C.c_int),
                ("panel_active", C.c_int), ("plain", C.c_int), ("dtype", C.c_int),
                ("panel_groups", C.c_int), ("panel_subchunks", C.c_int), ("panel_blocks", C.c_int),
                ("block_slots", C.c_int), ("spmm_window", C.c_int)]


class DriverReport(C.Structure):
    _fields_ = [("format", C.c_int), ("variants", C.c_int), ("best_method", C.c_int),
                ("best_bheight", C.c_int), ("best_bwidth", C.c_int), ("best_ms", C.c_double),
                ("best_gflops", C.c_double), ("ms", C.c_double * 64), ("method", C.c_int * 64),
                ("bheight", C.c_int * 64), ("bwidth", C.c_int * 64), ("max_rel_diff", C.c_double),
                ("check_distance", C.c_double), ("check_max_diff", C.c_double), ("check_max_ratio", C.c_double),
                ("max_scaled_error", C.c_double), ("variants_over_tol", C.c_int)]


class PartTiming(C.Structure):
    _fields_ = [("format", C.c_int), ("method", C.c_int), ("nnz", C.c_int),
                ("stored", C.c_longlong), ("ms", C.c_double)]


class EvalReport(C.Structure):
    _fields_ = [("ms_per_spmv", C.c_double), ("gflops", C.c_double), ("l2_distance", C.c_double),
                ("max_diff", C.c_double), ("max_ratio", C.c_double)]


class ErrorReport(C.Structure):
    """clspmv_error_report (include/clspmv_device.h)."""
    _fields_ = [("max_scaled_error", C.c_double), ("rms_scaled_error", C.c_double),
                ("max_abs_error", C.c_double), ("worst_row", C.c_int), ("rows", C.c_int),
                ("rows_over_tol", C.c_longlong), ("tol", C.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


# ---------------------------------------------------------------------------
# signatures

def _sig(name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


P = C.POINTER
VP = C.c_void_p
_sig("clspmv_last_error", C.c_char_p)
for _n, _T in [("coo", CooMatrix), ("csr", CsrMatrix), ("ell", EllMatrix), ("sell", SellMatrix),
               ("dia", DiaMatrix), ("dia_ext", DiaExtMatrix), ("bdia", BdiaMatrix),
               ("bcsr", BcsrMatrix), ("bell", BellMatrix), ("b4csr", B4csrMatrix),
               ("b4ell", B4ellMatrix), ("sbell", SbellMatrix)]:
    _sig(f"clspmv_init_{_n}_matrix", None, P(_T))
    _sig(f"clspmv_free_{_n}_matrix", None, P(_T))
_sig("clspmv_if_sorted_coo", C.c_int, P(CooMatrix))
_sig("clspmv_sort_coo", C.c_int, P(CooMatrix))
_sig("clspmv_aligned_length", C.c_int, C.c_int, C.c_int)
_sig("clspmv_validate_coo", C.c_int, P(CooMatrix))
_sig("clspmv_coo2csr", C.c_int, P(CooMatrix), P(CsrMatrix))
_sig("clspmv_coo2ell", C.c_int, P(CooMatrix), P(EllMatrix), C.c_int, C.c_int)
_sig("clspmv_coo2sell", C.c_int, P(CooMatrix), P(SellMatrix), C.c_uint)
_sig("clspmv_coo2dia", C.c_int, P(CooMatrix), P(DiaMatrix), C.c_int)
_sig("clspmv_coo2diaext", C.c_int, P(CooMatrix), P(DiaExtMatrix), C.c_int)
_sig("clspmv_coo2bdia", C.c_int, P(CooMatrix), P(BdiaMatrix), C.c_int)
_sig("clspmv_coo2bcsr", C.c_int, P(CooMatrix), P(BcsrMatrix), C.c_uint, C.c_uint, C.c_int)
_sig("clspmv_coo2b4csr", C.c_int, P(CooMatrix), P(B4csrMatrix), C.c_uint, C.c_uint, C.c_int)
_sig("clspmv_coo2bell", C.c_int, P(CooMatrix), P(BellMatrix), C.c_uint, C.c_uint, C.c_int, C.c_int)
_sig("clspmv_coo2b4ell", C.c_int, P(CooMatrix), P(B4ellMatrix), C.c_uint, C.c_uint, C.c_int, C.c_int)
_sig("clspmv_coo2sbell", C.c_int, P(CooMatrix), P(SbellMatrix), C.c_uint, C.c_uint, C.c_uint)
_sig("clspmv_coo2b4csr_nocap", C.c_int, P(CooMatrix), P(B4csrMatrix), C.c_uint, C.c_uint, C.c_int)
_sig("clspmv_coo2b4ell_nocap", C.c_int, P(CooMatrix), P(B4ellMatrix), C.c_uint, C.c_uint, C.c_int, C.c_int)
_sig("clspmv_coo2sbell_nocap", C.c_int, P(CooMatrix), P(SbellMatrix), C.c_uint, C.c_uint, C.c_uint)
for _n, _T in [("csr", CsrMatrix), ("ell", EllMatrix), ("dia", DiaMatrix), ("diaext", DiaExtMatrix),
               ("bdia", BdiaMatrix), ("bcsr", BcsrMatrix), ("bell", BellMatrix)]:
    _sig(f"clspmv_{_n}2coo", C.c_int, P(_T), P(CooMatrix))
_sig("clspmv_read_mmf", C.c_int, C.c_char_p, P(CooMatrix))

_sig("clspmv_ctx_create", C.c_int, C.c_int, P(VP))
_sig("clspmv_ctx_destroy", C.c_int, VP)
_sig("clspmv_ctx_synchronize", C.c_int, VP)
_sig("clspmv_ctx_stream", VP, VP)
_sig("clspmv_ctx_device", C.c_int, VP)
_sig("clspmv_device_info", C.c_int, VP, C.c_char_p, C.c_int, P(C.c_int))
_sig("clspmv_device_malloc", C.c_int, VP, C.c_size_t, P(VP))
_sig("clspmv_device_free", C.c_int, VP, VP)
_sig("clspmv_memcpy_h2d", C.c_int, VP, VP, VP, C.c_size_t)
_sig("clspmv_memcpy_d2h", C.c_int, VP, VP, VP, C.c_size_t)
_sig("clspmv_memcpy_d2d", C.c_int, VP, VP, VP, C.c_size_t)
_sig("clspmv_memset_zero", C.c_int, VP, VP, C.c_size_t)
_sig("clspmv_method_count", C.c_int, C.c_int)
for _n, _T in [("csr", CsrMatrix), ("coo", CooMatrix), ("ell", EllMatrix), ("sell", SellMatrix),
               ("dia", DiaMatrix), ("bdia", BdiaMatrix), ("b4csr", B4csrMatrix),
               ("b4ell", B4ellMatrix), ("sbell", SbellMatrix)]:
    _sig(f"clspmv_upload_{_n}", C.c_int, VP, P(_T), P(VP))
_sig("clspmv_dmat_free", C.c_int, VP)
_sig("clspmv_dmat_get_info", C.c_int, VP, P(DmatInfo))
class UploadOpts(C.Structure):
    _fields_ = [("panel_width", C.c_int), ("dtype", C.c_int), ("panel_group", C.c_int)]


_sig("clspmv_upload_coo_ex", C.c_int, VP, P(CooMatrix), P(UploadOpts), P(VP))
_sig("clspmv_upload_csr_ex", C.c_int, VP, P(CsrMatrix), P(UploadOpts), P(VP))
_sig("clspmv_upload_dia_ex", C.c_int, VP, P(DiaMatrix), P(UploadOpts), P(VP))
_sig("clspmv_dmat_set_panel", C.c_int, VP, C.c_int)
_sig("clspmv_upload_bcsr_lowp", C.c_int, VP, P(BcsrMatrix), C.c_int, P(VP))
_sig("clspmv_spmv_multi", C.c_int, VP, VP, C.c_int, VP, VP, C.c_int)
_sig("clspmv_verify_spmv", C.c_int, VP, VP, VP, VP, VP, C.c_double, P(ErrorReport))
_sig("clspmv_dmat_error_report", C.c_int, VP, VP, VP, C.c_ulonglong, C.c_double, P(ErrorReport))
_sig("clspmv_verify_spmv_noalloc", C.c_int, VP, VP, VP, VP, VP, C.c_double, VP)
NOALLOC_MAX_ROW_BLOCKS = 131072
_sig("clspmv_spmm", C.c_int, VP, VP, C.c_int, VP, C.c_int, VP, C.c_int)
_sig("clspmv_spmm_tiled", C.c_int, VP, VP, C.c_int, VP, VP, C.c_int)
_sig("clspmv_spmm_tiled_ex", C.c_int, VP, VP, C.c_int, VP, VP, C.c_int, C.c_int)
_sig("clspmv_spmm_pack_x", C.c_int, VP, C.c_int, C.c_int, VP, C.c_int, VP)
_sig("clspmv_convert_from_f32", C.c_int, PF, VP, C.c_size_t, C.c_int)
_sig("clspmv_convert_to_f32", C.c_int, VP, PF, C.c_size_t, C.c_int)
_sig("clspmv_server_create", C.c_int, VP, P(VP))
_sig("clspmv_server_destroy", C.c_int, VP)
_sig("clspmv_server_add_matrix", C.c_int, VP, VP, C.c_int, C.c_int, P(C.c_int))
_sig("clspmv_server_submit", C.c_int, VP, VP, P(C.c_longlong))
_sig("clspmv_server_wait", C.c_int, VP, C.c_longlong, P(C.c_double))
_sig("clspmv_server_clock", C.c_double, VP)
_sig("clspmv_server_replay", C.c_int, VP, C.c_int, VP, C.c_int, P(C.c_double), P(C.c_double))
_sig("clspmv_request_x", None, C.c_ulonglong, C.c_int, C.c_int, PF)
_sig("clspmv_spmv", C.c_int, VP, VP, C.c_int, VP, VP)
_sig("clspmv_time_spmv", C.c_int, VP, VP, C.c_int, VP, VP, C.c_int, C.c_int, P(C.c_double))
_sig("clspmv_time_spmv_multi", C.c_int, VP, VP, C.c_int, VP, VP, C.c_int, C.c_int, C.c_int, P(C.c_double))
_sig("clspmv_bandwidth_probe", C.c_int, VP, C.c_size_t, C.c_int, P(C.c_double), P(C.c_double))
_sig("clspmv_gather_probe", C.c_int, VP, C.c_int, C.c_longlong, C.c_int, P(C.c_double), P(C.c_double))
_sig("clspmv_read_probe", C.c_int, VP, C.c_size_t, C.c_int, P(C.c_double), P(C.c_double))
DRIVERS = ["csr_scalar", "csr_vector", "coo", "ell", "sell", "dia", "bdia", "bell", "sbell", "bcsr"]
for _n in DRIVERS:
    _sig(f"clspmv_spmv_{_n}", C.c_int, C.c_char_p, P(CooMatrix), C.c_int, C.c_int, C.c_int,
         P(DriverReport), PF)

if hasattr(lib, "clspmv_analyze_matrix"):
    _sig("clspmv_init_cocktail", None, P(CocktailMatrix))
    _sig("clspmv_free_cocktail", None, P(CocktailMatrix))
    _sig("clspmv_analyze_matrix", C.c_int, P(CooMatrix), P(CocktailMatrix), C.c_float, C.c_int,
         C.c_int, C.c_char_p)
_sig("clspmv_autotune", C.c_int, C.c_int, C.c_char_p, C.c_int, C.c_int)
if hasattr(lib, "clspmv_init_mat_kernels"):
    _sig("clspmv_init_mat_kernels", C.c_int, P(CocktailMatrix), PF, PF, P(VP), P(VP), C.c_int)
    _sig("clspmv_init_mat_kernels_on", C.c_int, C.c_int, P(CocktailMatrix), PF, PF, P(VP), P(VP))
    _sig("clspmv_do_spmv", C.c_int, VP, C.c_int)
    _sig("clspmv_cocktail_time_cold", C.c_int, VP, C.c_int, C.c_size_t, P(C.c_double))
    _sig("clspmv_do_spmv_async", C.c_int, VP, C.c_int)
    for _n in ["vector_from_cpu", "result_from_cpu"]:
        _sig(f"clspmv_cpy_{_n}", C.c_int, VP, VP, PF, C.c_int)
    for _n in ["vector_from_gpu", "result_from_gpu", "result_to_gpu"]:
        _sig(f"clspmv_cpy_{_n}", C.c_int, VP, VP, VP, C.c_int)
    _sig("clspmv_cpy_result_to_cpu", C.c_int, VP, VP, PF, C.c_int)
    _sig("clspmv_free_mat_kernels", C.c_int, VP, VP)
    _sig("clspmv_cocktail_gpu_vec", VP, VP)
    _sig("clspmv_cocktail_gpu_res", VP, VP)
    _sig("clspmv_cocktail_kernels_ctx", VP, VP)
    _sig("clspmv_cocktail_profile", C.c_int, VP, C.c_int, P(PartTiming), C.c_int, P(C.c_int),
         P(C.c_double))
    _sig("clspmv_cocktail_tune", C.c_int, VP, C.c_int, P(C.c_int))
    _sig("clspmv_cocktail_set_graph", C.c_int, VP, C.c_int)
    _sig("clspmv_cocktail_set_overlap", C.c_int, VP, C.c_int)
    _sig("clspmv_cocktail_set_fuse", C.c_int, VP, C.c_int)
    _sig("clspmv_cocktail_launch_modes", C.c_int, VP, P(C.c_int), P(C.c_int), P(C.c_int))
    _sig("clspmv_evaluate", C.c_int, P(CocktailMatrix), PF, PF, PF, P(EvalReport))


# ---------------------------------------------------------------------------
# host helpers

def _ip(a: np.ndarray):
    return a.ctypes.data_as(PI)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(PF)


def arr(ptr, n: int, dtype) -> np.ndarray:
    """Copy n elements behind a ctypes pointer into a numpy array."""
    if n <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Coo:
    """A COO matrix whose arrays live in numpy; ``.c`` is the C struct
    pointing at them (valid while this object lives)."""

    def __init__(self, rows, cols, vals, height: int, width: int):
        self.row = np.ascontiguousarray(rows, dtype=np.int32)
        self.col = np.ascontiguousarray(cols, dtype=np.int32)
        self.val = np.ascontiguousarray(vals, dtype=np.float32)
        assert self.row.shape == self.col.shape == self.val.shape
        self.height, self.width = int(height), int(width)
        self.c = CooMatrix()
        self.c.matinfo = MatInfo(self.width, self.height, int(self.row.size))
        self.c.coo_row_id = _ip(self.row)
        self.c.coo_col_id = _ip(self.col)
        self.c.coo_data = _fp(self.val)

    @property
    def nnz(self) -> int:
        return int(self.row.size)

    def ref(self):
        return C.byref(self.c)

    def copy(self) -> "Coo":
        return Coo(self.row.copy(), self.col.copy(), self.val.copy(), self.height, self.width)

    def sorted(self) -> "Coo":
        order = np.lexsort((self.col, self.row))
        return Coo(self.row[order], self.col[order], self.val[order], self.height, self.width)


def struct_arrays(m) -> dict:
    """numpy copies of every array of a format struct, sized from its header."""
    t = type(m)
    mi = m.matinfo
    if t is CooMatrix:
        n = mi.nnz
        return dict(row=arr(m.coo_row_id, n, np.int32), col=arr(m.coo_col_id, n, np.int32),
                    val=arr(m.coo_data, n, np.float32))
    if t is CsrMatrix:
        return dict(rp=arr(m.csr_row_ptr, mi.height + 1, np.int32),
                    col=arr(m.csr_col_id, mi.nnz, np.int32), val=arr(m.csr_data, mi.nnz, np.float32))
    if t is EllMatrix:
        n = m.ell_num * m.ell_height_aligned
        return dict(col=arr(m.ell_col_id, n, np.int32), val=arr(m.ell_data, n, np.float32))
    if t is SellMatrix:
        ptr = arr(m.sell_slice_ptr, m.sell_slice_num + 1, np.int32)
        n = int(ptr[-1]) if ptr.size else 0
        return dict(ptr=ptr, col=arr(m.sell_col_id, n, np.int32), val=arr(m.sell_data, n, np.float32))
    if t in (DiaMatrix, DiaExtMatrix):
        return dict(off=arr(m.dia_offsets, m.dia_num, np.int32),
                    val=arr(m.dia_data, m.dia_num * m.dia_length_aligned, np.float32))
    if t is BdiaMatrix:
        bptr = arr(m.bdia_bptr, m.bdia_band_num + 1, np.int32)
        nd = int(bptr[-1]) if bptr.size else 0
        return dict(off=arr(m.bdia_offsets, m.bdia_band_num, np.int32), bptr=bptr,
                    val=arr(m.bdia_data, nd * m.bdia_length_aligned, np.float32))
    if t is BcsrMatrix:
        bs = m.bcsr_bwidth * m.bcsr_bheight
        return dict(rp=arr(m.bcsr_row_ptr, m.bcsr_row_num + 1, np.int32),
                    col=arr(m.bcsr_col_id, m.bcsr_block_num, np.int32),
                    val=arr(m.bcsr_data, m.bcsr_aligned_size * bs, np.float32))
    if t is BellMatrix:
        n = m.bell_height_aligned * m.bell_block_num
        bs = m.bell_bwidth * m.bell_bheight
        return dict(col=arr(m.bell_col_id, n, np.int32), val=arr(m.bell_data, n * bs, np.float32))
    if t is B4csrMatrix:
        return dict(rp=arr(m.b4csr_row_ptr, m.b4csr_row_num + 1, np.int32),
                    col=arr(m.b4csr_col_id, m.b4csr_block_num, np.int32),
                    val=arr(m.b4csr_data, m.b4csr_aligned_size * m.b4csr_bheight * (m.b4csr_bwidth // 4),
                            np.float32))
    if t is B4ellMatrix:
        n = m.b4ell_height_aligned * m.b4ell_block_num
        nd = m.b4ell_float4_aligned * m.b4ell_bheight * (m.b4ell_bwidth // 4) * m.b4ell_block_num
        return dict(col=arr(m.b4ell_col_id, n, np.int32), val=arr(m.b4ell_data, nd, np.float32))
    if t is SbellMatrix:
        ptr = arr(m.sbell_slice_ptr, m.sbell_slice_num + 1, np.int32)
        n = int(ptr[-1]) if ptr.size else 0
        return dict(ptr=ptr, col=arr(m.sbell_col_id, n, np.int32),
                    val=arr(m.sbell_data, n * m.sbell_bwidth * m.sbell_bheight, np.float32))
    raise TypeError(t)


def struct_header(m) -> dict:
    """Scalar fields of a format struct (no pointers)."""
    out = {}
    for name, typ in m._fields_:
        v = getattr(m, name)
        if name == "matinfo":
            out.update(width=v.width, height=v.height, nnz=v.nnz)
        elif typ in (C.c_int, C.c_uint, C.c_bool):
            out[name] = int(v)
    return out


_FREE = {CooMatrix: "coo", CsrMatrix: "csr", EllMatrix: "ell", SellMatrix: "sell", DiaMatrix: "dia",
         DiaExtMatrix: "dia_ext", BdiaMatrix: "bdia", BcsrMatrix: "bcsr", BellMatrix: "bell",
         B4csrMatrix: "b4csr", B4ellMatrix: "b4ell", SbellMatrix: "sbell"}


class Owned:
    """A format struct filled by the library; frees its arrays on delete."""

    def __init__(self, struct):
        self.s = struct

    def __del__(self):
        try:
            getattr(lib, f"clspmv_free_{_FREE[type(self.s)]}_matrix")(C.byref(self.s))
        except Exception:
            pass

    def arrays(self) -> dict:
        return struct_arrays(self.s)

    def header(self) -> dict:
        return struct_header(self.s)


def _convert(fn: str, T, coo: Coo, *args, expect_bool=False):
    dst = T()
    getattr(lib, f"clspmv_init_{_FREE[T]}_matrix")(C.byref(dst))
    rc = getattr(lib, fn)(coo.ref(), C.byref(dst), *args)
    check(rc, fn)
    if expect_bool and rc == 0:
        return None
    return Owned(dst)


def coo2csr(coo: Coo):
    return _convert("clspmv_coo2csr", CsrMatrix, coo)


def coo2ell(coo: Coo, alignment=GPU_ALIGNMENT, ellnum=0):
    return _convert("clspmv_coo2ell", EllMatrix, coo, alignment, ellnum)


def coo2sell(coo: Coo, slice_height=WARPSIZE):
    return _convert("clspmv_coo2sell", SellMatrix, coo, slice_height)


def coo2dia(coo: Coo, alignment=GPU_ALIGNMENT):
    return _convert("clspmv_coo2dia", DiaMatrix, coo, alignment, expect_bool=True)


def coo2diaext(coo: Coo, alignment=GPU_ALIGNMENT):
    return _convert("clspmv_coo2diaext", DiaExtMatrix, coo, alignment, expect_bool=True)


def coo2bdia(coo: Coo, alignment=GPU_ALIGNMENT):
    return _convert("clspmv_coo2bdia", BdiaMatrix, coo, alignment, expect_bool=True)


def coo2bcsr(coo: Coo, bw: int, bh: int, alignment=GPU_ALIGNMENT):
    return _convert("clspmv_coo2bcsr", BcsrMatrix, coo, bw, bh, alignment)


def coo2b4csr(coo: Coo, bw: int, bh: int, alignment=GPU_ALIGNMENT, cap=True):
    fn = "clspmv_coo2b4csr" if cap else "clspmv_coo2b4csr_nocap"
    return _convert(fn, B4csrMatrix, coo, bw, bh, alignment, expect_bool=True)


def coo2bell(coo: Coo, bw: int, bh: int, alignment=GPU_ALIGNMENT, ellnum=0):
    return _convert("clspmv_coo2bell", BellMatrix, coo, bw, bh, alignment, ellnum)


def coo2b4ell(coo: Coo, bw: int, bh: int, alignment=GPU_ALIGNMENT, ellnum=0, cap=True):
    fn = "clspmv_coo2b4ell" if cap else "clspmv_coo2b4ell_nocap"
    return _convert(fn, B4ellMatrix, coo, bw, bh, alignment, ellnum, expect_bool=True)


def coo2sbell(coo: Coo, bw: int, bh: int, slice_height=WARPSIZE, cap=True):
    fn = "clspmv_coo2sbell" if cap else "clspmv_coo2sbell_nocap"
    return _convert(fn, SbellMatrix, coo, bw, bh, slice_height, expect_bool=True)


def to_coo(owned: Owned) -> Coo:
    """X -> COO through the library's inverse converter."""
    name = {CsrMatrix: "csr", EllMatrix: "ell", DiaMatrix: "dia", DiaExtMatrix: "diaext",
            BdiaMatrix: "bdia", BcsrMatrix: "bcsr", BellMatrix: "bell"}[type(owned.s)]
    dst = CooMatrix()
    check(getattr(lib, f"clspmv_{name}2coo")(C.byref(owned.s), C.byref(dst)), f"{name}2coo")
    o = Owned(dst)
    a = o.arrays()
    return Coo(a["row"], a["col"], a["val"], dst.matinfo.height, dst.matinfo.width)


def read_mmf(path: str) -> Coo:
    dst = CooMatrix()
    check(lib.clspmv_read_mmf(path.encode(), C.byref(dst)), "read_mmf")
    o = Owned(dst)
    a = o.arrays()
    return Coo(a["row"], a["col"], a["val"], dst.matinfo.height, dst.matinfo.width)


# ---------------------------------------------------------------------------
# device side

class Context:
    """One GPU + one HIP stream (replaces cl_context + cl_command_queue)."""

    def __init__(self, device: int = 0):
        h = VP()
        check(lib.clspmv_ctx_create(device, C.byref(h)), "ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib.clspmv_ctx_destroy(self.h)
            self.h = VP()

    def __del__(self):
        self.close()

    def sync(self):
        check(lib.clspmv_ctx_synchronize(self.h), "ctx_synchronize")

    def bandwidth_probe(self, nbytes: int = 1 << 30, ntimes: int = 20):
        """(ms per copy, GB/s counting read + write) of a float4 copy."""
        ms, gbps = C.c_double(), C.c_double()
        check(lib.clspmv_bandwidth_probe(self.h, nbytes, ntimes, C.byref(ms), C.byref(gbps)), "probe")
        return ms.value, gbps.value

    def read_probe(self, nbytes: int = 1 << 30, ntimes: int = 20):
        """(ms per pass, GB/s read) of a read-only float4 stream."""
        ms, gbps = C.c_double(), C.c_double()
        check(lib.clspmv_read_probe(self.h, nbytes, ntimes, C.byref(ms), C.byref(gbps)), "read_probe")
        return ms.value, gbps.value

    def gather_probe(self, nx: int, n: int = 1 << 24, ntimes: int = 20):
        """(ms per call, G gathers/s) of random reads from an nx-float vector."""
        ms, g = C.c_double(), C.c_double()
        check(lib.clspmv_gather_probe(self.h, nx, n, ntimes, C.byref(ms), C.byref(g)), "gather_probe")
        return ms.value, g.value

    def info(self):
        buf = C.create_string_buffer(256)
        cu = C.c_int()
        check(lib.clspmv_device_info(self.h, buf, 256, C.byref(cu)), "device_info")
        return buf.value.decode(), cu.value


class DeviceVector:
    """A float32 device buffer owned through the C-ABI."""

    def __init__(self, ctx: Context, n: int, host: Optional[np.ndarray] = None):
        self.ctx, self.n = ctx, int(n)
        p = VP()
        check(lib.clspmv_device_malloc(ctx.h, max(self.n, 1) * 4, C.byref(p)), "device_malloc")
        self.p = p
        if host is not None:
            self.upload(host)

    def upload(self, host: np.ndarray):
        h = np.ascontiguousarray(host, dtype=np.float32)
        assert h.size == self.n
        check(lib.clspmv_memcpy_h2d(self.ctx.h, self.p, h.ctypes.data_as(VP), self.n * 4), "h2d")

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float32)
        check(lib.clspmv_memcpy_d2h(self.ctx.h, out.ctypes.data_as(VP), self.p, self.n * 4), "d2h")
        return out

    def zero(self):
        check(lib.clspmv_memset_zero(self.ctx.h, self.p, self.n * 4), "memset")

    def view(self, offset: int) -> "DeviceVector":
        """The floats from `offset` on, sharing this buffer (not owned)."""
        v = DeviceVector.__new__(DeviceVector)
        v.ctx, v.n, v._base = self.ctx, self.n - int(offset), self
        v.p = VP((self.p.value or 0) + 4 * int(offset))
        v._owned = False
        return v

    def __del__(self):
        try:
            if self.p and getattr(self, "_owned", True):
                lib.clspmv_device_free(self.ctx.h, self.p)
        except Exception:
            pass


_UPLOAD = {CsrMatrix: "csr", CooMatrix: "coo", EllMatrix: "ell", SellMatrix: "sell",
           DiaMatrix: "dia", BdiaMatrix: "bdia", B4csrMatrix: "b4csr", B4ellMatrix: "b4ell",
           SbellMatrix: "sbell"}


PANEL_AUTO = -1
PANEL_BOTH = -2
DT_F32, DT_BF16, DT_F16 = 0, 1, 2
FMT_BSPMM = 9


def convert_from_f32(a: np.ndarray, dtype: int) -> np.ndarray:
    """float32 -> 16-bit storage (round to nearest even), as uint16 bits."""
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty(a.shape, np.uint16)
    check(lib.clspmv_convert_from_f32(_fp(a.reshape(-1)), out.ctypes.data_as(VP), a.size, dtype),
          "convert_from_f32")
    return out


def convert_to_f32(bits: np.ndarray, dtype: int) -> np.ndarray:
    bits = np.ascontiguousarray(bits, np.uint16)
    out = np.empty(bits.shape, np.float32)
    check(lib.clspmv_convert_to_f32(bits.ctypes.data_as(VP), _fp(out.reshape(-1)), bits.size, dtype),
          "convert_to_f32")
    return out


class DeviceBuffer:
    """Raw device bytes (16-bit operands of clspmv_spmm)."""

    def __init__(self, ctx: Context, host: np.ndarray):
        h = np.ascontiguousarray(host)
        self.ctx, self.nbytes, self.dtype, self.shape = ctx, h.nbytes, h.dtype, h.shape
        p = VP()
        check(lib.clspmv_device_malloc(ctx.h, max(self.nbytes, 16), C.byref(p)), "device_malloc")
        self.p = p
        check(lib.clspmv_memcpy_h2d(ctx.h, p, h.ctypes.data_as(VP), self.nbytes), "h2d")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(lib.clspmv_memcpy_d2h(self.ctx.h, out.ctypes.data_as(VP), self.p, self.nbytes), "d2h")
        return out

    def __del__(self):
        try:
            if self.p:
                lib.clspmv_device_free(self.ctx.h, self.p)
        except Exception:
            pass


def _nbytes(buf) -> int:
    """Size in bytes of a DeviceBuffer or DeviceVector."""
    return buf.nbytes if hasattr(buf, "nbytes") else 4 * buf.n


def spmm_pack_x(ctx: Context, k: int, cols: int, X: DeviceBuffer, ldx: int) -> DeviceBuffer:
    """Column-major 16-bit X (k columns, leading dimension ldx) into the
    4-row tiles of clspmv_spmm_tiled, on the GPU (clspmv_spmm_pack_x)."""
    need = 2 * (ldx * (k - 1) + cols) if k > 0 else 0
    if _nbytes(X) < need:
        raise ClspmvError(ERR_INVALID, "spmm_pack_x", f"X holds {_nbytes(X)} bytes, k={k} ldx={ldx} need {need}")
    X4 = DeviceBuffer(ctx, np.zeros(max(1, (cols + 3) // 4 * k * 4), np.uint16))
    check(lib.clspmv_spmm_pack_x(ctx.h, k, cols, X.p, ldx, X4.p), "spmm_pack_x")
    return X4


class ServerConfig(C.Structure):
    _fields_ = [("device_ordinal", C.c_int), ("streams", C.c_int), ("max_batch", C.c_int),
                ("hbm_budget", C.c_size_t), ("use_graphs", C.c_int), ("host_staging", C.c_size_t)]


class Request(C.Structure):
    _fields_ = [("matrix", C.c_int), ("k", C.c_int), ("seed", C.c_ulonglong), ("x", PF),
                ("x_on_device", C.c_int), ("host_y", PF), ("device_y", VP), ("arrival", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_longlong), ("passes", C.c_longlong), ("columns", C.c_longlong),
                ("evictions", C.c_longlong), ("pool_bytes", C.c_size_t), ("pool_used", C.c_size_t)]


_sig("clspmv_server_get_stats", C.c_int, VP, P(ServerStats))


def request_x(seed: int, j: int, n: int) -> np.ndarray:
    """x_j of a served request, as the device generates it."""
    out = np.empty(n, np.float32)
    lib.clspmv_request_x(C.c_ulonglong(seed), j, n, _fp(out))
    return out


def _request(matrix, k, seed, arrival=0.0, host_y=None, x=None, device_y=None) -> Request:
    """A clspmv_request; x is a host array (cols * k floats, column-major) or
    a DeviceVector / DeviceBuffer on the server's GPU."""
    if x is None:
        xp, xdev = None, 0
    elif isinstance(x, np.ndarray):
        assert x.dtype == np.float32 and x.flags.c_contiguous
        xp, xdev = _fp(x), 0
    else:
        xp, xdev = C.cast(x.p, PF), 1
    return Request(matrix, k, seed, xp, xdev, _fp(host_y) if host_y is not None else None,
                   device_y.p if device_y is not None else None, arrival)


class Server:
    """The request-serving loop (include/clspmv_serve.h)."""

    def __init__(self, device: int = 0, streams: int = 4, max_batch: int = 64,
                 hbm_budget: int = 8 << 30, use_graphs: bool = True, host_staging: int = 0):
        cfg = ServerConfig(device, streams, max_batch, hbm_budget, int(use_graphs), host_staging)
        h = VP()
        check(lib.clspmv_server_create(C.byref(cfg), C.byref(h)), "server_create")
        self.h = h
        self._keep = []  # matrices must outlive the server

    def add(self, dmat, method: int = METHOD_AUTO, kmax: int = 16) -> int:
        i = C.c_int()
        check(lib.clspmv_server_add_matrix(self.h, dmat.h, method, kmax, C.byref(i)), "server_add_matrix")
        self._keep.append(dmat)
        return i.value

    def submit(self, matrix: int, k: int, seed: int, host_y: Optional[np.ndarray] = None, x=None,
               device_y=None) -> int:
        r = _request(matrix, k, seed, 0.0, host_y, x, device_y)
        t = C.c_longlong()
        check(lib.clspmv_server_submit(self.h, C.byref(r), C.byref(t)), "server_submit")
        return t.value

    def wait(self, ticket: int) -> float:
        lat = C.c_double()
        check(lib.clspmv_server_wait(self.h, C.c_longlong(ticket), C.byref(lat)), "server_wait")
        return lat.value

    def stats(self) -> dict:
        st = ServerStats()
        check(lib.clspmv_server_get_stats(self.h, C.byref(st)), "server_get_stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def replay(self, reqs, mode: int = 0):
        """reqs: list of (matrix, k, seed, arrival_s, host_y or None[, x[, device_y]]).
        Returns (latencies, wall seconds)."""
        arr = (Request * len(reqs))()
        for i, r in enumerate(reqs):
            m, k, seed, t, y = r[:5]
            arr[i] = _request(m, k, seed, t, y, r[5] if len(r) > 5 else None, r[6] if len(r) > 6 else None)
        lat = np.zeros(len(reqs), np.float64)
        wall = C.c_double()
        check(lib.clspmv_server_replay(self.h, len(reqs), arr, mode, lat.ctypes.data_as(P(C.c_double)),
                                       C.byref(wall)), "server_replay")
        return lat, wall.value

    def close(self):
        if self.h:
            lib.clspmv_server_destroy(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BlockSpmm:
    """A 4x4-block BCSR matrix with BF16 / FP16 values on the device and its
    MFMA SpMM: Y += A X (clspmv_upload_bcsr_lowp / clspmv_spmm)."""

    def __init__(self, ctx: Context, bcsr: "Owned", dtype: int):
        h = VP()
        check(lib.clspmv_upload_bcsr_lowp(ctx.h, C.byref(bcsr.s), dtype, C.byref(h)), "upload_bcsr_lowp")
        self.ctx, self.h, self.dtype = ctx, h, dtype
        self.info = DmatInfo()
        check(lib.clspmv_dmat_get_info(h, C.byref(self.info)), "dmat_info")

    def _check_y(self, k: int, Y, ldy: int):
        need = 4 * (ldy * (k - 1) + self.info.rows) if k > 0 else 0
        if _nbytes(Y) < need:
            raise ClspmvError(ERR_INVALID, "spmm", f"Y holds {_nbytes(Y)} bytes, k={k} ldy={ldy} need {need}")

    def spmm(self, k: int, X, ldx: int, Y, ldy: int):
        # The C entry point takes raw pointers and cannot see buffer sizes:
        # a short X or Y is refused here instead of read or written past
        # its end on the GPU.
        need = 2 * (ldx * (k - 1) + self.info.cols) if k > 0 else 0
        if _nbytes(X) < need:
            raise ClspmvError(ERR_INVALID, "spmm", f"X holds {_nbytes(X)} bytes, k={k} ldx={ldx} need {need}")
        self._check_y(k, Y, ldy)
        check(lib.clspmv_spmm(self.ctx.h, self.h, k, X.p, ldx, Y.p, ldy), "spmm")

    def spmm_tiled(self, k: int, X4, Y, ldy: int, accumulate: bool = True):
        """Y += A X (accumulate) or Y = A X, Y not read, with X in 4-row
        tiles (clspmv_spmm_tiled_ex; spmm_pack_x builds them from a
        column-major X)."""
        need = 8 * ((self.info.cols + 3) // 4) * k
        if _nbytes(X4) < need:
            raise ClspmvError(ERR_INVALID, "spmm_tiled", f"X4 holds {_nbytes(X4)} bytes, k={k} need {need}")
        self._check_y(k, Y, ldy)
        check(lib.clspmv_spmm_tiled_ex(self.ctx.h, self.h, k, X4.p, Y.p, ldy, int(accumulate)), "spmm_tiled")

    def error_report(self, ref: "DeviceMatrix", seed: int = 1, tol: float = 1e-5) -> dict:
        """Error of one SpMM column (x rounded to the matrix's type) against
        the FP64 SpMV of the FP32 CSR `ref` (clspmv_dmat_error_report)."""
        return error_report(self, ref, seed, tol)

    def __del__(self):
        try:
            if self.h:
                lib.clspmv_dmat_free(self.h)
        except Exception:
            pass


class DeviceMatrix:
    """A device-resident matrix of any storage format.  For COO and CSR,
    ``panel`` selects the device encoding (0 own arrays, a width, PANEL_AUTO
    or PANEL_BOTH); for COO, CSR and DIA, ``dtype`` the stored value type
    (DT_F32, DT_BF16, DT_F16) -- clspmv_upload_{coo,csr,dia}_ex.  Both None:
    the plain upload call."""

    def __init__(self, ctx: Context, m, panel: Optional[int] = None, dtype: Optional[int] = None,
                 group: int = 0):
        s = m.s if isinstance(m, Owned) else (m.c if isinstance(m, Coo) else m)
        h = VP()
        name = _UPLOAD[type(s)]
        if panel is not None or dtype is not None or group:
            if name not in ("coo", "csr", "dia"):
                raise ValueError("encoding options exist for COO, CSR and DIA only")
            default_pw = PANEL_AUTO if name == "coo" else 0
            if name == "csr" and dtype not in (None, DT_F32) and panel is None:
                default_pw = PANEL_AUTO
            opts = UploadOpts(default_pw if panel is None else int(panel), DT_F32 if dtype is None else dtype,
                              int(group))
            check(getattr(lib, f"clspmv_upload_{name}_ex")(ctx.h, C.byref(s), C.byref(opts), C.byref(h)),
                  "upload")
        else:
            check(getattr(lib, f"clspmv_upload_{name}")(ctx.h, C.byref(s), C.byref(h)), "upload")
        self.ctx, self.h = ctx, h
        self.info = DmatInfo()
        self._refresh()

    def _refresh(self):
        check(lib.clspmv_dmat_get_info(self.h, C.byref(self.info)), "dmat_info")

    def spmv_multi(self, k: int, X, Y, ldy: int):
        """Y[:, q] += A X[:, q] for q < k; X row-major (cols*k), Y column-major."""
        check(lib.clspmv_spmv_multi(self.ctx.h, self.h, k, X.p, Y.p, ldy), "spmv_multi")

    def time_multi(self, k: int, X, Y, ldy: int, warmup: int = 2, ntimes: int = 5) -> float:
        """ms per spmv_multi pass over k right-hand sides (clspmv_time_spmv_multi;
        Y keeps accumulating)."""
        ms = C.c_double()
        check(lib.clspmv_time_spmv_multi(self.ctx.h, self.h, k, X.p, Y.p, ldy, warmup, ntimes, C.byref(ms)),
              "time_spmv_multi")
        return ms.value

    def set_panel(self, use: bool):
        check(lib.clspmv_dmat_set_panel(self.h, int(use)), "dmat_set_panel")
        self._refresh()

    def error_report(self, ref: "DeviceMatrix", seed: int = 1, tol: float = 1e-5) -> dict:
        """Row-scaled error of this matrix's SpMV against the FP64 SpMV of the
        FP32 CSR `ref` on x = request_x(seed, 0) (clspmv_dmat_error_report)."""
        return error_report(self, ref, seed, tol)

    @property
    def format(self) -> int:
        return self.info.format

    def methods(self) -> int:
        return lib.clspmv_method_count(self.info.format)

    def spmv(self, x: DeviceVector, y: DeviceVector, method: int = METHOD_AUTO):
        """y += A*x (asynchronous on the context's stream)."""
        check(lib.clspmv_spmv(self.ctx.h, self.h, method, x.p, y.p), "spmv")

    def time(self, x: DeviceVector, y: DeviceVector, method=METHOD_AUTO, warmup=3, ntimes=20) -> float:
        ms = C.c_double()
        check(lib.clspmv_time_spmv(self.ctx.h, self.h, method, x.p, y.p, warmup, ntimes, C.byref(ms)),
              "time_spmv")
        return ms.value

    def __del__(self):
        try:
            if self.h:
                lib.clspmv_dmat_free(self.h)
        except Exception:
            pass


def error_report(A, ref: "DeviceMatrix", seed: int = 1, tol: float = 1e-5) -> dict:
    """clspmv_dmat_error_report: A (DeviceMatrix or BlockSpmm) against the
    FP64 SpMV of the FP32 CSR matrix `ref` on the GPU."""
    rep = ErrorReport()
    check(lib.clspmv_dmat_error_report(A.ctx.h, A.h, ref.h, seed, tol, C.byref(rep)), "dmat_error_report")
    return rep.as_dict()


def verify_spmv(ctx: Context, ref: "DeviceMatrix", x: DeviceVector, y: DeviceVector,
                y0: Optional[DeviceVector] = None, tol: float = 1e-5) -> dict:
    """clspmv_verify_spmv: scores y (= y0 + A x, computed by the caller)
    against the FP64 SpMV of the FP32 CSR matrix `ref` on the GPU."""
    rep = ErrorReport()
    check(lib.clspmv_verify_spmv(ctx.h, ref.h, x.p, y0.p if y0 is not None else None, y.p, tol,
                                 C.byref(rep)), "verify_spmv")
    return rep.as_dict()


def verify_spmv_noalloc(ctx: Context, ref: "DeviceMatrix", x: DeviceVector, y: DeviceVector,
                        report: "DeviceBuffer", y0: Optional[DeviceVector] = None, tol: float = 1e-5) -> None:
    """clspmv_verify_spmv_noalloc: clspmv_verify_spmv's check without any
    allocation, asynchronous on the context's stream; the report lands in
    `report`, a device buffer of at least one clspmv_error_report (read it
    with report_from_device after a sync)."""
    if report.nbytes < C.sizeof(ErrorReport):
        raise ValueError(f"report buffer holds {report.nbytes} bytes, needs {C.sizeof(ErrorReport)}")
    check(lib.clspmv_verify_spmv_noalloc(ctx.h, ref.h, x.p, y0.p if y0 is not None else None, y.p, tol,
                                         report.p), "verify_spmv_noalloc")


def report_from_device(report: "DeviceBuffer") -> dict:
    """The clspmv_error_report at the start of a device buffer."""
    raw = report.download().view(np.uint8)[:C.sizeof(ErrorReport)]
    return ErrorReport.from_buffer_copy(raw.tobytes()).as_dict()


def run_driver(name: str, coo: Coo, ntimes: int = 10, device: int = 0, want_y: bool = True):
    """Per-format driver spmv_<name> (include/spmv_<fmt>.h:7)."""
    rep = DriverReport()
    y = np.zeros(coo.height, dtype=np.float32) if want_y else None
    check(getattr(lib, f"clspmv_spmv_{name}")(None, coo.ref(), 1, ntimes, device, C.byref(rep),
                                               _fp(y) if want_y else None), f"spmv_{name}")
    return rep, y


# ---------------------------------------------------------------------------
# Cocktail (include/clspmv_cocktail.h)

class Cocktail:
    """A partitioned matrix (cocktail<int,int,float>) owned by the library."""

    def __init__(self):
        self.c = CocktailMatrix()
        lib.clspmv_init_cocktail(C.byref(self.c))

    def __del__(self):
        try:
            lib.clspmv_free_cocktail(C.byref(self.c))
        except Exception:
            pass

    def parts(self) -> list:
        """[(format name, part struct)] of the parts in use, in launch order."""
        out = []
        for n in ["bdia", "dia", "sbell", "bell", "bcsr", "sell", "ell", "csr", "coo"]:
            if getattr(self.c, "ifuse" + n):
                out.append((n, getattr(self.c, n)))
        return out

    def summary(self) -> dict:
        return {n: int(getattr(p, "matinfo").nnz) for n, p in self.parts()}


def autotune(out_dir: str, device: int = 0, level: int = 0, verbose: bool = False) -> str:
    """Measure all kernel variants on this GPU and write <out_dir>/benchmark/*.ben."""
    check(lib.clspmv_autotune(device, out_dir.encode(), level, int(verbose)), "autotune")
    return out_dir


def evaluate(ck: Cocktail, x: np.ndarray, res: np.ndarray, coores: Optional[np.ndarray] = None) -> dict:
    """clspmv_evaluate (evaluate, src/spmv_cocktail.cpp:1431): one SpMV
    res += A x (res updated in place), the reference's comparison with
    coores, then 1000 timed passes."""
    assert res.dtype == np.float32 and res.flags.c_contiguous
    rep = EvalReport()
    xx = np.ascontiguousarray(x, np.float32)
    check(lib.clspmv_evaluate(C.byref(ck.c), _fp(xx), _fp(res),
                              _fp(coores) if coores is not None else None, C.byref(rep)), "evaluate")
    return {f: getattr(rep, f) for f, _ in EvalReport._fields_}


# Tables measured on an MI355X by autotune(), shipped with the package.
TABLES_DIR = os.path.join(_HERE, "tables_mi355x")


def analyze_matrix(coo: Coo, ben_dir: Optional[str] = None, dia_priority_multiplier: float = 1.0,
                   estimate_flat_partial: bool = True, estimate_flat_full: bool = False) -> Cocktail:
    """analyze_matrix (include/analyze.h:257); tables from <ben_dir>/benchmark."""
    ck = Cocktail()
    check(lib.clspmv_analyze_matrix(coo.ref(), C.byref(ck.c), C.c_float(dia_priority_multiplier),
                                    int(estimate_flat_partial), int(estimate_flat_full),
                                    ben_dir.encode() if ben_dir else None), "analyze_matrix")
    return ck


class CocktailRuntime:
    """init_mat_kernels ... free_mat_kernels (include/spmv_cocktail.h:166-175)
    over a partitioned matrix; one HIP graph per pass over the parts."""

    def __init__(self, ck: Cocktail, x: np.ndarray, y: Optional[np.ndarray] = None, device: int = 0):
        x = np.ascontiguousarray(x, np.float32)
        self.width, self.height, self.nnz = ck.c.mat_width, ck.c.mat_height, ck.c.mat_nnz
        y = np.zeros(self.height, np.float32) if y is None else np.ascontiguousarray(y, np.float32)
        g, k = VP(), VP()
        check(lib.clspmv_init_mat_kernels_on(device, C.byref(ck.c), _fp(x), _fp(y), C.byref(g), C.byref(k)),
              "init_mat_kernels")
        self.g, self.k = g, k

    def do_spmv(self, ntimes: int = 1):
        check(lib.clspmv_do_spmv(self.k, ntimes), "do_spmv")

    def do_spmv_async(self, ntimes: int = 1):
        check(lib.clspmv_do_spmv_async(self.k, ntimes), "do_spmv_async")

    def sync(self):
        check(lib.clspmv_ctx_synchronize(lib.clspmv_cocktail_kernels_ctx(self.k)), "sync")

    def set_x(self, x: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        check(lib.clspmv_cpy_vector_from_cpu(self.g, self.k, _fp(x), x.size), "cpy_vector_from_cpu")

    def set_y(self, y: np.ndarray):
        y = np.ascontiguousarray(y, np.float32)
        check(lib.clspmv_cpy_result_from_cpu(self.g, self.k, _fp(y), y.size), "cpy_result_from_cpu")

    def get_y(self) -> np.ndarray:
        y = np.empty(self.height, np.float32)
        check(lib.clspmv_cpy_result_to_cpu(self.g, self.k, _fp(y), self.height), "cpy_result_to_cpu")
        return y

    def y_to_x(self):
        """cpy_vector_from_gpu(gpumat, ..., gpumat.res): next x = current y."""
        check(lib.clspmv_cpy_vector_from_gpu(self.g, self.k, lib.clspmv_cocktail_gpu_res(self.g),
                                             self.height), "cpy_vector_from_gpu")

    def profile(self, ntimes: int = 20):
        parts = (PartTiming * 16)()
        n = C.c_int()
        gms = C.c_double()
        check(lib.clspmv_cocktail_profile(self.k, ntimes, parts, 16, C.byref(n), C.byref(gms)), "profile")
        return [dict(format=FORMAT_NAMES[parts[i].format], method=parts[i].method, nnz=parts[i].nnz,
                     stored=parts[i].stored, ms=parts[i].ms) for i in range(n.value)], gms.value

    def time_cold(self, ntimes: int = 20, scrub_bytes: int = 1 << 30) -> float:
        """ms per pass with the caches scrubbed before each pass
        (clspmv_cocktail_time_cold)."""
        ms = C.c_double()
        check(lib.clspmv_cocktail_time_cold(self.k, ntimes, C.c_size_t(scrub_bytes), C.byref(ms)),
              "cocktail_time_cold")
        return ms.value

    def tune(self, ntimes: int = 20) -> bool:
        """Time every allowed method (and lane count) per part, keep the
        fastest, and pick graph or plain launches; returns True for graph."""
        g = C.c_int()
        check(lib.clspmv_cocktail_tune(self.k, ntimes, C.byref(g)), "cocktail_tune")
        return bool(g.value)

    def set_graph(self, graph: bool):
        """One graph launch per pass (True) or plain kernel launches."""
        check(lib.clspmv_cocktail_set_graph(self.k, int(graph)), "cocktail_set_graph")

    def set_overlap(self, overlap: bool):
        """Run the x-panel parts' first pass on a second stream beside the
        earlier parts (True, the default) or everything on one stream."""
        check(lib.clspmv_cocktail_set_overlap(self.k, int(overlap)), "cocktail_set_overlap")

    def set_fuse(self, fuse: bool):
        """Sum an FP32 DIA part directly ahead of an x-panel part in that
        part's first-pass launch (True) or in its own launch (False, the
        default after init)."""
        check(lib.clspmv_cocktail_set_fuse(self.k, int(fuse)), "cocktail_set_fuse")

    def launch_modes(self) -> dict:
        """{"graph", "overlap", "fuse"}: overlap / fuse are True only when a
        pass actually uses them."""
        g, o, f = C.c_int(), C.c_int(), C.c_int()
        check(lib.clspmv_cocktail_launch_modes(self.k, C.byref(g), C.byref(o), C.byref(f)),
              "cocktail_launch_modes")
        return {"graph": bool(g.value), "overlap": bool(o.value), "fuse": bool(f.value)}

    def close(self):
        if self.k:
            lib.clspmv_free_mat_kernels(self.g, self.k)
            self.g, self.k = VP(), VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
