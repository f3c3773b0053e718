"""GPU tests of the FP64 verification SpMV, the per-matrix error report and
the FP64 accumulation of rows longer than 4096 entries.

Pinned by the FP64 oracle (oracle/spmv_oracle.c: spmv_f64 and the bound
sum |a x|) and by bounds the arithmetic fixes:
  - a row summed in FP64 from exact products and rounded once into FP32 is
    the correctly rounded sum: within half an ulp of itself, so within
    U = 2^-24 relative to |y0| + sum |a x|;
  - the panel path sums at most 512 entries in FP32 (8 in order per lane,
    then a segmented scan of depth 6: at most 15 roundings) and block SpMM
    16 blocks before FP64 takes over, so their long rows stay within 16 U /
    17 U whatever the row length (bounds; these two checks do not separate
    FP64 from FP32 carries on random data);
  - 16-bit storage moves each value by at most 2^-8 (BF16: 8 significant
    bits) or 2^-11 (FP16: 11) of itself (round to nearest, normal range),
    so the row-scaled error of the whole SpMV stays below that (twice that
    when x is rounded too).
"""

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices

pytestmark = pytest.mark.gpu

U = 2.0 ** -24
TOL = 1e-5


def _host_errors(y, y64, bound):
    """Row-scaled errors as clspmv_verify_spmv defines them (0 / inf when
    the bound is 0, NaN -> inf)."""
    d = np.abs(y.astype(np.float64) - y64)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.where(bound > 0, d / np.where(bound > 0, bound, 1.0), np.where(d > 0, np.inf, 0.0))
    e[np.isnan(e)] = np.inf
    return e


def _ref(ctx, A):
    return cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=0)


def _stored(vals, dtype):
    if dtype in (None, cs.DT_F32):
        return vals
    return cs.convert_to_f32(cs.convert_from_f32(vals, dtype), dtype)


MATS = {
    "rand": matrices.random_sparse(3000, 2500, 0.004, seed=11, long_rows=3, long_len=900),
    "rand_empty_rows": matrices.random_sparse(1000, 1000, 0.01, seed=3, empty_rows=0.3),
    "very_long_rows": matrices.random_sparse(2000, 50000, 0.0004, seed=21, long_rows=6, long_len=9000),
    "stencil27_16": matrices.stencil_27pt(16, seed=3),
}


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("with_y0", [False, True])
def test_verify_matches_oracle(ctx, oracle, name, with_y0):
    """clspmv_verify_spmv against the oracle's FP64 SpMV on a result with
    errors planted at known rows: same worst row, same errors."""
    A = MATS[name]
    x = matrices.uniform_vector(A.width, seed=5)
    y0 = matrices.uniform_vector(A.height, seed=6) if with_y0 else None
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    y = y64.astype(np.float32)
    rng = np.random.default_rng(7)
    planted = rng.choice(np.flatnonzero(bound > 0), size=2, replace=False)
    y[planted[0]] += np.float32(3e-4 * max(bound[planted[0]], 1.0))
    y[planted[1]] -= np.float32(1e-6 * max(bound[planted[1]], 1.0))
    e = _host_errors(y, y64, bound)
    ref = _ref(ctx, A)
    rep = cs.verify_spmv(ctx, ref, cs.DeviceVector(ctx, A.width, x), cs.DeviceVector(ctx, A.height, y),
                         cs.DeviceVector(ctx, A.height, y0) if with_y0 else None, tol=TOL)
    assert rep["rows"] == A.height
    assert rep["worst_row"] == planted[0] == int(np.argmax(e))
    assert rep["max_scaled_error"] == pytest.approx(e.max(), rel=1e-9)
    assert rep["rms_scaled_error"] == pytest.approx(np.sqrt(np.mean(e * e)), rel=1e-6)
    assert rep["rows_over_tol"] == int(np.sum(e > TOL))
    assert rep["max_abs_error"] == pytest.approx(np.abs(y.astype(np.float64) - y64).max(), rel=1e-9)


def test_verify_exact_and_zero_bound(ctx, oracle):
    """Dyadic data: the FP32 result is exact, every error is 0 (worst row 0);
    an empty row with y != y0 scores +inf."""
    A = matrices.random_sparse(700, 500, 0.01, seed=2, dyadic=True, empty_rows=0.2)
    x = matrices.dyadic_vector(A.width, seed=3)
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height)
    y = y64.astype(np.float32)
    ref = _ref(ctx, A)
    dx = cs.DeviceVector(ctx, A.width, x)
    rep = cs.verify_spmv(ctx, ref, dx, cs.DeviceVector(ctx, A.height, y))
    assert rep["max_scaled_error"] == 0.0 and rep["worst_row"] == 0 and rep["rows_over_tol"] == 0
    empty = int(np.flatnonzero(bound == 0)[0])
    y[empty] = 1.0
    rep = cs.verify_spmv(ctx, ref, dx, cs.DeviceVector(ctx, A.height, y))
    assert rep["max_scaled_error"] == np.inf and rep["worst_row"] == empty and rep["rows_over_tol"] == 1


def _report_cases():
    out = []
    for dt in (cs.DT_F32, cs.DT_BF16, cs.DT_F16):
        out += [("rand", "csr", dt), ("rand", "coo", dt), ("stencil27_16", "dia", dt),
                ("very_long_rows", "csr", dt)]
    out += [("block", "bspmm", cs.DT_BF16), ("block", "bspmm", cs.DT_F16)]
    return out


BLOCK = matrices.block_stencil(700, 4, 4, seed=6, offsets=(0, 1, -1, 32, -33, 5))


@pytest.mark.parametrize("name,fmt,dtype", _report_cases())
def test_error_report(ctx, oracle, name, fmt, dtype):
    """clspmv_dmat_error_report: equals verify_spmv on the same x and y,
    stays under the bound of the value type, and for 16-bit storage shows
    the rounding (above a floor no FP32 result reaches)."""
    A = BLOCK if fmt == "bspmm" else MATS[name]
    ref = _ref(ctx, A)
    if fmt == "csr":
        dA = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=0, dtype=dtype)
    elif fmt == "coo":
        dA = cs.DeviceMatrix(ctx, A.sorted(), panel=512, dtype=dtype)
    elif fmt == "dia":
        dA = cs.DeviceMatrix(ctx, cs.coo2dia(A.copy()), dtype=dtype)
    else:
        dA = cs.BlockSpmm(ctx, cs.coo2bcsr(A.copy(), 4, 4), dtype)
    rep = dA.error_report(ref, seed=9, tol=TOL)
    assert rep["rows"] == A.height and rep["tol"] == TOL
    # The same x through the caller-facing pieces, scored by the oracle.
    x = cs.request_x(9, 0, A.width)
    xs = _stored(x, dtype) if fmt == "bspmm" else x
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height)
    if fmt != "bspmm":
        y = cs.DeviceVector(ctx, A.height, np.zeros(A.height, np.float32))
        dA.spmv(cs.DeviceVector(ctx, A.width, x), y)
        ctx.sync()
        e = _host_errors(y.download(), y64, bound)
        assert rep["max_scaled_error"] == pytest.approx(e.max(), rel=1e-9)
        assert rep["worst_row"] == int(np.argmax(e))
        assert rep["rows_over_tol"] == int(np.sum(e > TOL))
    if dtype == cs.DT_F32:
        assert rep["max_scaled_error"] <= TOL
        return
    step = 2.0 ** -8 if dtype == cs.DT_BF16 else 2.0 ** -11
    bound_t = (2 * step if fmt == "bspmm" else step) + 1e-6
    assert rep["max_scaled_error"] <= bound_t
    assert rep["max_scaled_error"] > step / 16
    if fmt != "bspmm":
        # Against the values as stored the same result is within the FP32
        # tolerance: the reported error is the storage rounding.
        y64s, bounds = oracle.spmv_f64(A.row, A.col, _stored(A.val, dtype), xs, A.height)
        assert _host_errors(y.download(), y64s, bounds).max() <= TOL


LONG = MATS["very_long_rows"]
LONG_ROWS = np.flatnonzero(np.bincount(LONG.row, minlength=LONG.height) > 4096)


def _long_check(y, y64, bound, limit, what):
    e = _host_errors(y, y64, bound)
    assert LONG_ROWS.size >= 4
    assert e[LONG_ROWS].max() <= limit, f"{what}: long rows {e[LONG_ROWS].max():.3e} > {limit:.3e}"
    assert e.max() <= TOL, f"{what}: {e.max():.3e}"


def _correctly_rounded(y, y64, bound, what):
    """Long rows equal y64 rounded to FP32: within half an ulp of the result
    (plus 1e-9 of the bound for the FP64 sums' own order).  With the signs
    mixed, an FP32 accumulation of these rows misses by up to ~18 ulps."""
    r = LONG_ROWS
    miss = np.abs(y[r].astype(np.float64) - y64[r]) - 0.5 * np.spacing(np.abs(y[r])).astype(np.float64)
    assert np.all(miss <= 1e-9 * bound[r]), f"{what}: rows {r[miss > 1e-9 * bound[r]]}"


@pytest.mark.parametrize("dtype", [cs.DT_F32, cs.DT_BF16, cs.DT_F16])
@pytest.mark.parametrize("k", [1, 4, 12, 16])
def test_long_rows_fp64_csr(ctx, oracle, dtype, k):
    """CSR rows longer than 4096 entries (16-bit values, and every
    multi-vector pass) are summed in FP64 and rounded once: the FP32 result
    is the correctly rounded exact sum."""
    A = LONG
    vals = _stored(A.val, dtype)
    dA = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=0, dtype=dtype)
    rng = np.random.default_rng(k)
    X = rng.uniform(-1, 1, size=(A.width, k)).astype(np.float32)
    Y0 = rng.uniform(-1, 1, size=(k, A.height)).astype(np.float32)
    dY = cs.DeviceVector(ctx, Y0.size, Y0.reshape(-1))
    if k == 1 and dtype != cs.DT_F32:
        dA.spmv(cs.DeviceVector(ctx, A.width, X[:, 0].copy()), dY)  # the 16-bit CSR SpMV
    else:
        dA.spmv_multi(k, cs.DeviceVector(ctx, X.size, X.reshape(-1)), dY, A.height)
    ctx.sync()
    Y = dY.download().reshape(k, A.height)
    for q in range(k):
        y64, bound = oracle.spmv_f64(A.row, A.col, vals, X[:, q].copy(), A.height, Y0[q].copy())
        _long_check(Y[q], y64, bound, U * (1 + 1e-6), f"rhs {q}")
        _correctly_rounded(Y[q], y64, bound, f"rhs {q}")


@pytest.mark.parametrize("fmt,dtype", [("csr", cs.DT_F32), ("coo", cs.DT_F32), ("coo", cs.DT_BF16),
                                       ("csr", cs.DT_F16)])
def test_long_rows_fp64_panel(ctx, oracle, fmt, dtype):
    """Panel encoding: a row's segment sums (each at most 512 entries, at
    most 15 FP32 roundings deep) are added in FP64 for rows with more than
    8 segments (every row over 4096 entries among them), so long rows stay
    within 16 U whatever their length."""
    A = LONG
    vals = _stored(A.val, dtype)
    owned = cs.coo2csr(A.copy()) if fmt == "csr" else A.sorted()
    dA = cs.DeviceMatrix(ctx, owned, panel=4096, dtype=dtype)
    assert dA.info.panel_active == 1
    x = matrices.uniform_vector(A.width, seed=3)
    y0 = matrices.uniform_vector(A.height, seed=4)
    y = cs.DeviceVector(ctx, A.height, y0)
    dA.spmv(cs.DeviceVector(ctx, A.width, x), y)
    ctx.sync()
    y64, bound = oracle.spmv_f64(A.row, A.col, vals, x, A.height, y0)
    _long_check(y.download(), y64, bound, 16 * U, f"{fmt} panel")


def _long_block_matrix():
    """4x4 dense blocks; block rows 0..3 hold 1500 blocks (6000 entries per
    row), the others 10."""
    rng = np.random.default_rng(31)
    nbr, nbc = 64, 3000
    rows, cols = [], []
    for br in range(nbr):
        nb = 1500 if br < 4 else 10
        bc = rng.choice(nbc, size=nb, replace=False)
        r = (br * 4 + np.arange(4))[:, None, None]
        c = (bc * 4)[None, :, None] + np.arange(4)[None, None, :]
        rr, cc = np.broadcast_arrays(r, c)
        rows.append(rr.reshape(-1))
        cols.append(cc.reshape(-1))
    row = np.concatenate(rows).astype(np.int32)
    col = np.concatenate(cols).astype(np.int32)
    val = rng.standard_normal(row.size).astype(np.float32)
    return cs.Coo(row, col, val, nbr * 4, nbc * 4).sorted()


@pytest.mark.parametrize("dtype", [cs.DT_BF16, cs.DT_F16])
def test_long_rows_fp64_spmm(ctx, oracle, dtype):
    """Block SpMM with block rows longer than 1024 blocks: the MFMA sums are
    carried into FP64 every 16 blocks, so a row's error is at most 16
    roundings of a chunk plus one into y: within 17 U."""
    A = _long_block_matrix()
    dA = cs.BlockSpmm(ctx, cs.coo2bcsr(A.copy(), 4, 4), dtype)
    k = 4
    ldx, ldy = A.width, A.height
    rng = np.random.default_rng(8)
    X = _stored(rng.uniform(-1, 1, size=k * ldx).astype(np.float32), dtype)
    dX = cs.DeviceBuffer(ctx, cs.convert_from_f32(X, dtype))
    dY = cs.DeviceVector(ctx, k * ldy, np.zeros(k * ldy, np.float32))
    dA.spmm(k, dX, ldx, dY, ldy)
    ctx.sync()
    Y = dY.download().reshape(k, ldy)
    vals = _stored(A.val, dtype)
    long_rows = np.arange(16)
    for q in range(k):
        y64, bound = oracle.spmv_f64(A.row, A.col, vals, X[q * ldx:(q + 1) * ldx].copy(), A.height)
        e = _host_errors(Y[q], y64, bound)
        assert e[long_rows].max() <= 17 * U, f"rhs {q}: {e[long_rows].max():.3e}"
        assert e.max() <= TOL
