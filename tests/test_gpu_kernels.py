"""GPU parity of every storage format and kernel variant against the oracle.

Bar: |y_i - y64_i| <= 1e-5 * (|y0_i| + sum_j |a_ij x_j|) against the FP64
oracle, and bit-exact equality on dyadic inputs,
where every FP32 partial sum is exact so the summation order cannot matter.
All calls go through the C-ABI (libclspmv.so).
"""

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _cases():
    return {
        "rand_small": matrices.random_sparse(300, 257, 0.03, seed=1),
        "rand_dyadic": matrices.random_sparse(513, 700, 0.02, seed=2, dyadic=True),
        "rand_empty_rows": matrices.random_sparse(1000, 1000, 0.01, seed=3, empty_rows=0.3),
        "rand_long_rows": matrices.random_sparse(2000, 3000, 0.002, seed=4, long_rows=5,
                                                 long_len=2500),
        "rand_long_dyadic": matrices.random_sparse(1500, 1500, 0.003, seed=5, dyadic=True,
                                                   long_rows=3, long_len=1400),
        # rows of up to ~1000 entries inside the short-row blocks (spanning
        # many stream chunks), empty rows, dyadic values
        "rand_mid_dyadic": matrices.random_sparse(3000, 2000, 0.002, seed=14, dyadic=True,
                                                  empty_rows=0.2, long_rows=6, long_len=1000),
        "lap2d_256": matrices.laplacian_2d(256),
        "stencil27_16": matrices.stencil_27pt(16, seed=3),
        "stencil27_dyadic": matrices.stencil_27pt(12, seed=5, dyadic=True),
        "block_small": matrices.block_stencil(700, 4, 4, seed=6, offsets=(0, 1, -1, 32, -33, 5)),
        "tall": matrices.random_sparse(3001, 97, 0.05, seed=7),
        "wide": matrices.random_sparse(97, 3001, 0.05, seed=8),
        "banded": matrices.banded(5000, (0, 1, -1, 2, 3, -7, 300, -300, 301, 302), seed=9),
        "banded_dyadic": matrices.banded(4099, (-2, -1, 0, 1, 5, 6, 7, 8, 1000), seed=10,
                                         dyadic=True),
    }


CASES = _cases()


def _is_dyadic(a: np.ndarray) -> bool:
    return bool(np.all(a * 64 == np.round(a * 64)))


def _run(ctx, oracle, A: cs.Coo, owned, method, panel=None, x_offset=0, dtype=None):
    dyadic = _is_dyadic(A.val)
    if dyadic:
        x = matrices.dyadic_vector(A.width, seed=11)
        y0 = matrices.dyadic_vector(A.height, seed=12)
    else:
        x = matrices.uniform_vector(A.width, seed=11)
        y0 = matrices.uniform_vector(A.height, seed=12)
    dA = cs.DeviceMatrix(ctx, owned, panel=panel, dtype=dtype)
    if panel is not None and panel > 0:
        assert dA.info.panel_width == panel and dA.info.panel_active == 1 and dA.info.plain == 0
    vals = A.val
    if dtype not in (None, cs.DT_F32):
        # The reference for 16-bit storage: the FP64 product of the values
        # as stored.
        assert dA.info.dtype == dtype
        vals = cs.convert_to_f32(cs.convert_from_f32(A.val, dtype), dtype)
    if x_offset:
        # x starting x_offset floats into its buffer: unaligned for 128-bit loads
        dx = cs.DeviceVector(ctx, A.width + x_offset, np.concatenate([np.zeros(x_offset, np.float32), x]))
        dx = dx.view(x_offset)
    else:
        dx = cs.DeviceVector(ctx, A.width, x)
    dy = cs.DeviceVector(ctx, A.height, y0)
    dA.spmv(dx, dy, method)
    ctx.sync()
    y = dy.download()
    y64, bound = oracle.spmv_f64(A.row, A.col, vals, x, A.height, y0)
    err, worst = oracle.scaled_error(y, y64, bound)
    assert err <= TOL, f"scaled error {err:.3e} at row {worst}: got {y[worst]} want {y64[worst]}"
    if dyadic:
        np.testing.assert_array_equal(y, y64.astype(np.float32))
    return y


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("method", [0, 1])
def test_csr(ctx, oracle, name, method):
    A = CASES[name]
    _run(ctx, oracle, A, cs.coo2csr(A.copy()), method)


@pytest.mark.parametrize("name", list(CASES))
def test_coo(ctx, oracle, name):
    A = CASES[name]
    _run(ctx, oracle, A, A.copy(), 0)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("method", [0, 1])
def test_ell(ctx, oracle, name, method):
    A = CASES[name]
    if name.startswith("rand_long"):
        pytest.skip("ELL of a matrix with 2500-entry rows is all padding")
    _run(ctx, oracle, A, cs.coo2ell(A.copy()), method)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("slice_height", [32, 128])
def test_sell(ctx, oracle, name, slice_height):
    A = CASES[name]
    for method in (0, 1):
        _run(ctx, oracle, A, cs.coo2sell(A, slice_height), method)


DIA_CASES = ["lap2d_256", "stencil27_16", "stencil27_dyadic", "block_small", "banded",
             "banded_dyadic"]


@pytest.mark.parametrize("name", DIA_CASES)
@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_dia(ctx, oracle, name, method):
    A = CASES[name]
    d = cs.coo2dia(A)
    if d is None:
        pytest.skip("DIA refused: padded size >= 2*nnz (include/matrix_storage.h:972)")
    offs = d.arrays()["off"]
    span = int(offs.max() - offs.min())
    if method >= 2 and (1024 if method == 3 else 256) + span > 8192:
        with pytest.raises(cs.ClspmvError):
            _run(ctx, oracle, A, d, method)
        return
    _run(ctx, oracle, A, d, method)


@pytest.mark.parametrize("name", DIA_CASES)
@pytest.mark.parametrize("method", [0, 1, 2, 3, 4])
def test_bdia(ctx, oracle, name, method):
    A = CASES[name]
    b = cs.coo2bdia(A)
    if b is None:
        pytest.skip("BDIA refused: padded size >= 2*nnz (include/matrix_storage.h:972)")
    _run(ctx, oracle, A, b, method)


BLOCK_CASES = ["block_small", "rand_small", "rand_dyadic", "lap2d_256", "tall", "wide"]


def _bcsr_slots(b) -> int:
    """Slots per block row of the slot-major BCSR encoding the upload picks
    (every block row at most 64 blocks, padding to the longest at most 10 %
    of the blocks), or 0 for the row-pointer encoding."""
    h, rp = b.header(), b.arrays()["rp"]
    nbr, nb = h["b4csr_row_num"], h["b4csr_block_num"]
    w = int(np.diff(rp).max()) if nbr else 0
    return w if nb > 0 and w <= 64 and w * nbr * 10 <= nb * 11 else 0


@pytest.mark.parametrize("name", BLOCK_CASES)
@pytest.mark.parametrize("shape", cs.BLOCK_SHAPES)
def test_bcsr(ctx, oracle, name, shape):
    """Both BCSR encodings (block_small and most lap2d shapes slot-major,
    the random matrices row-pointer) against the oracle."""
    A = CASES[name]
    bh, bw = shape
    b = cs.coo2b4csr(A.copy(), bw, bh)
    assert cs.DeviceMatrix(ctx, b).info.block_slots == _bcsr_slots(b)
    _run(ctx, oracle, A, b, 0)


def _ragged_blocks(bh, bw):
    """Dyadic block rows of 6 blocks, every 20th block row missing one (a
    padding slot in the slot-major encoding)."""
    A = matrices.block_stencil(900, bh, bw, seed=13, offsets=(0, 1, -1, 32, -33, 5), dyadic=True)
    brow, bcol = A.row // bh, A.col // bw
    drop = (brow % 20 == 7) & (bcol == (brow + 5) % 900)
    return cs.Coo(A.row[~drop], A.col[~drop], A.val[~drop], A.height, A.width)


@pytest.mark.parametrize("shape", [(4, 4), (2, 8), (8, 4), (1, 4)])
@pytest.mark.parametrize("x_offset", [0, 1])
def test_bcsr_slot_major_ragged(ctx, oracle, shape, x_offset):
    """Slot-major BCSR with padding slots: bit for bit the oracle's sum on
    dyadic data, with x aligned and unaligned for float4 loads."""
    bh, bw = shape
    A = _ragged_blocks(bh, bw)
    b = cs.coo2b4csr(A.copy(), bw, bh)
    hdr = b.header()
    assert 0 < _bcsr_slots(b) and _bcsr_slots(b) * hdr["b4csr_row_num"] > hdr["b4csr_block_num"]
    assert cs.DeviceMatrix(ctx, b).info.block_slots == _bcsr_slots(b)
    _run(ctx, oracle, A, b, 0, x_offset=x_offset)


@pytest.mark.parametrize("name", BLOCK_CASES)
@pytest.mark.parametrize("shape", cs.BLOCK_SHAPES)
def test_bell(ctx, oracle, name, shape):
    A = CASES[name]
    bh, bw = shape
    owned = cs.coo2b4ell(A.copy(), bw, bh)
    for method in (0, 1):
        _run(ctx, oracle, A, owned, method)


@pytest.mark.parametrize("name", BLOCK_CASES)
@pytest.mark.parametrize("shape", cs.BLOCK_SHAPES)
def test_sbell(ctx, oracle, name, shape):
    A = CASES[name]
    bh, bw = shape
    _run(ctx, oracle, A, cs.coo2sbell(A.copy(), bw, bh), 0)


def test_auto_method_every_format(ctx, oracle):
    A = CASES["stencil27_16"]
    for owned in [cs.coo2csr(A.copy()), A.copy(), cs.coo2ell(A.copy()), cs.coo2sell(A),
                  cs.coo2dia(A), cs.coo2bdia(A), cs.coo2b4csr(A.copy(), 4, 4),
                  cs.coo2b4ell(A.copy(), 4, 4), cs.coo2sbell(A.copy(), 4, 4)]:
        _run(ctx, oracle, A, owned, cs.METHOD_AUTO)


def test_empty_matrix(ctx):
    A = cs.Coo(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 10, 10)
    dA = cs.DeviceMatrix(ctx, cs.coo2csr(A))
    y0 = np.arange(10, dtype=np.float32)
    dx = cs.DeviceVector(ctx, 10, np.ones(10, np.float32))
    dy = cs.DeviceVector(ctx, 10, y0)
    dA.spmv(dx, dy, 0)
    ctx.sync()
    np.testing.assert_array_equal(dy.download(), y0)


def test_bad_method_rejected(ctx):
    A = CASES["rand_small"]
    dA = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()))
    dx = cs.DeviceVector(ctx, A.width, np.ones(A.width, np.float32))
    dy = cs.DeviceVector(ctx, A.height, np.zeros(A.height, np.float32))
    with pytest.raises(cs.ClspmvError):
        dA.spmv(dx, dy, 7)


def test_repeat_is_deterministic(ctx):
    """Same inputs, same kernel: bitwise identical results (no atomics)."""
    A = CASES["rand_long_rows"]
    x = matrices.uniform_vector(A.width, seed=3)
    for owned, m in [(cs.coo2csr(A.copy()), 1), (A.copy(), 0)]:
        dA = cs.DeviceMatrix(ctx, owned)
        dx = cs.DeviceVector(ctx, A.width, x)
        outs = []
        for _ in range(3):
            dy = cs.DeviceVector(ctx, A.height, np.zeros(A.height, np.float32))
            dA.spmv(dx, dy, m)
            ctx.sync()
            outs.append(dy.download())
        np.testing.assert_array_equal(outs[0], outs[1])
        np.testing.assert_array_equal(outs[0], outs[2])


PANEL_CASES = ["rand_small", "rand_dyadic", "rand_empty_rows", "rand_long_rows", "rand_long_dyadic",
               "lap2d_256", "tall", "wide", "banded_dyadic"]


@pytest.mark.parametrize("name", PANEL_CASES)
@pytest.mark.parametrize("pw", [8, 64, 1024, 32760])
@pytest.mark.parametrize("fmt", ["coo", "csr"])
def test_panel_encoding(ctx, oracle, name, pw, fmt):
    """x-panel encoding (16-bit in-panel columns, LDS-staged x) at widths from
    one to many panels, COO and CSR, both methods of CSR."""
    A = CASES[name]
    owned = A.sorted() if fmt == "coo" else cs.coo2csr(A.copy())
    if -(-A.width // pw) > 2048:  # more panels than the encoding holds
        with pytest.raises(cs.ClspmvError) as e:
            cs.DeviceMatrix(ctx, owned, panel=pw)
        assert e.value.code == cs.ERR_UNSUPPORTED
        return
    for method in ([0] if fmt == "coo" else [0, 1]):
        _run(ctx, oracle, A, owned, method, panel=pw)


@pytest.mark.parametrize("fmt", ["coo", "csr"])
def test_panel_unaligned_x(ctx, oracle, fmt):
    A = CASES["rand_long_rows"]
    owned = A.sorted() if fmt == "coo" else cs.coo2csr(A.copy())
    _run(ctx, oracle, A, owned, 0, panel=64, x_offset=1)


def test_panel_both_encodings_switch(ctx, oracle):
    """PANEL_BOTH keeps both encodings; each gives the oracle's result."""
    A = matrices.random_sparse(3000, 70000, 0.002, seed=21, long_rows=4, long_len=3000)
    for fmt in ("coo", "csr"):
        owned = A.sorted() if fmt == "coo" else cs.coo2csr(A.copy())
        x = matrices.uniform_vector(A.width, seed=11)
        y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height)
        dA = cs.DeviceMatrix(ctx, owned, panel=cs.PANEL_BOTH)
        assert dA.info.panel_width > 0 and dA.info.plain == 1
        for use in (True, False):
            dA.set_panel(use)
            dx = cs.DeviceVector(ctx, A.width, x)
            dy = cs.DeviceVector(ctx, A.height, np.zeros(A.height, np.float32))
            dA.spmv(dx, dy, 0)
            ctx.sync()
            err, worst = oracle.scaled_error(dy.download(), y64, bound)
            assert err <= TOL, f"{fmt} panel={use}: {err:.3e} at {worst}"


def test_panel_bad_width_rejected(ctx):
    A = CASES["rand_small"].sorted()
    for bad in (4, 12, 32768, -3):
        with pytest.raises(cs.ClspmvError):
            cs.DeviceMatrix(ctx, A, panel=bad)


@pytest.mark.parametrize("name", DIA_CASES)
@pytest.mark.parametrize("method", [0, 1, 2, 3])
@pytest.mark.parametrize("dtype", [cs.DT_BF16, cs.DT_F16])
def test_dia_16bit_values(ctx, oracle, name, method, dtype):
    """DIA with values stored in BF16 / FP16, accumulated in FP32."""
    A = CASES[name]
    d = cs.coo2dia(A)
    if d is None:
        pytest.skip("DIA refused: padded size >= 2*nnz (include/matrix_storage.h:972)")
    offs = d.arrays()["off"]
    if method >= 2 and (1024 if method == 3 else 256) + int(offs.max() - offs.min()) > 8192:
        pytest.skip("offset span too wide for the LDS window")
    _run(ctx, oracle, A, d, method, dtype=dtype)


@pytest.mark.parametrize("name", ["rand_small", "rand_dyadic", "rand_long_dyadic", "lap2d_256", "wide"])
@pytest.mark.parametrize("dtype", [cs.DT_BF16, cs.DT_F16])
@pytest.mark.parametrize("fmt", ["coo", "csr"])
def test_panel_16bit_values(ctx, oracle, name, dtype, fmt):
    A = CASES[name]
    owned = A.sorted() if fmt == "coo" else cs.coo2csr(A.copy())
    _run(ctx, oracle, A, owned, 0, panel=64, dtype=dtype)
    _run(ctx, oracle, A, owned, 0, dtype=dtype)  # automatic panel width


def test_16bit_needs_panel(ctx):
    A = CASES["rand_small"]
    for panel in (0, cs.PANEL_BOTH):
        with pytest.raises(cs.ClspmvError):
            cs.DeviceMatrix(ctx, A.sorted(), panel=panel, dtype=cs.DT_BF16)


MULTI_CASES = ["rand_small", "rand_dyadic", "rand_long_rows", "rand_long_dyadic", "rand_mid_dyadic", "lap2d_256",
               "stencil27_dyadic", "banded", "tall", "wide"]


@pytest.mark.parametrize("name", MULTI_CASES)
@pytest.mark.parametrize("fmt", ["dia", "csr"])
@pytest.mark.parametrize("dtype", [cs.DT_F32, cs.DT_BF16, cs.DT_F16])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 8, 12, 16])
def test_spmv_multi(ctx, oracle, name, fmt, dtype, k):
    """One pass for k right-hand sides (X row-major, Y column-major) equals k
    oracle SpMVs; 16-bit values are compared with the values as stored."""
    A = CASES[name]
    if fmt == "dia" and A.height != A.width:
        pytest.skip("DIA conversion is defined for square matrices")
    owned = cs.coo2dia(A) if fmt == "dia" else cs.coo2csr(A.copy())
    if owned is None:
        pytest.skip("DIA refused: padded size >= 2*nnz (include/matrix_storage.h:972)")
    dyadic = _is_dyadic(A.val)
    rng = np.random.default_rng(k)
    if dyadic:
        X = (rng.integers(-64, 65, size=(A.width, k)) / 64.0).astype(np.float32)
        Y0 = (rng.integers(-64, 65, size=(k, A.height + 3)) / 64.0).astype(np.float32)
    else:
        X = rng.uniform(-1, 1, size=(A.width, k)).astype(np.float32)
        Y0 = rng.uniform(-1, 1, size=(k, A.height + 3)).astype(np.float32)
    ldy = A.height + 3
    dA = cs.DeviceMatrix(ctx, owned, panel=0 if fmt == "csr" else None, dtype=dtype)
    vals = A.val if dtype == cs.DT_F32 else cs.convert_to_f32(cs.convert_from_f32(A.val, dtype), dtype)
    dX = cs.DeviceVector(ctx, X.size, X.reshape(-1))
    dY = cs.DeviceVector(ctx, Y0.size, Y0.reshape(-1))
    dA.spmv_multi(k, dX, dY, ldy)
    ctx.sync()
    Y = dY.download().reshape(k, ldy)
    for q in range(k):
        y64, bound = oracle.spmv_f64(A.row, A.col, vals, X[:, q].copy(), A.height, Y0[q, :A.height].copy())
        err, worst = oracle.scaled_error(Y[q, :A.height].copy(), y64, bound)
        assert err <= TOL, f"rhs {q}: {err:.3e} at row {worst}"
        if dyadic:
            np.testing.assert_array_equal(Y[q, :A.height], y64.astype(np.float32))
        np.testing.assert_array_equal(Y[q, A.height:], Y0[q, A.height:])


@pytest.mark.parametrize("name", ["rand_small", "rand_dyadic", "rand_long_rows"])
@pytest.mark.parametrize("dtype", [cs.DT_BF16, cs.DT_F16])
def test_csr_16bit_values(ctx, oracle, name, dtype):
    """CSR's own arrays with 16-bit values through clspmv_spmv (every method)."""
    A = CASES[name]
    for method in (0, 1):
        _run(ctx, oracle, A, cs.coo2csr(A.copy()), method, panel=0, dtype=dtype)


@pytest.mark.parametrize("dtype", [cs.DT_BF16, cs.DT_F16])
def test_csr_16bit_both_encodings(ctx, oracle, dtype):
    """16-bit CSR values in both encodings: the panel SpMV, the plain SpMV
    after clspmv_dmat_set_panel(0), and a multi-vector pass each match the
    oracle on the values as stored."""
    A = matrices.random_sparse(20000, 30000, 0.0005, seed=12, long_rows=4, long_len=5000)
    vals = cs.convert_to_f32(cs.convert_from_f32(A.val, dtype), dtype)
    dA = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=cs.PANEL_BOTH, dtype=dtype)
    assert dA.info.plain == 1 and dA.info.panel_active == 1 and dA.info.dtype == dtype
    x = matrices.uniform_vector(A.width, seed=3)
    y0 = matrices.uniform_vector(A.height, seed=4)
    y64, bound = oracle.spmv_f64(A.row, A.col, vals, x, A.height, y0)
    dx = cs.DeviceVector(ctx, A.width, x)
    for panel in (True, False):
        dA.set_panel(panel)
        assert dA.info.panel_active == int(panel)
        y = cs.DeviceVector(ctx, A.height, y0)
        dA.spmv(dx, y)
        ctx.sync()
        err, worst = oracle.scaled_error(y.download(), y64, bound)
        assert err <= TOL, f"panel={panel}: {err:.3e} at row {worst}"
    k = 4
    X = np.random.default_rng(5).uniform(-1, 1, size=(A.width, k)).astype(np.float32)
    dY = cs.DeviceVector(ctx, k * A.height, np.zeros(k * A.height, np.float32))
    dA.spmv_multi(k, cs.DeviceVector(ctx, X.size, X.reshape(-1)), dY, A.height)
    ctx.sync()
    Y = dY.download().reshape(k, A.height)
    for q in range(k):
        y64q, bq = oracle.spmv_f64(A.row, A.col, vals, X[:, q].copy(), A.height)
        err, worst = oracle.scaled_error(Y[q].copy(), y64q, bq)
        assert err <= TOL, f"rhs {q}: {err:.3e} at row {worst}"


def test_panel_block_with_many_long_rows(ctx, oracle):
    """Pass-2 blocks holding more than 1024 rows of more than 8 slots (the
    long-row list then takes more than one round of the work-group)."""
    # Every row: one entry in each of 10 distinct panels (of 16), so 10 slots.
    rng = np.random.default_rng(17)
    n, pw = 700_000, 4096
    r = np.repeat(np.arange(n, dtype=np.int64), 10)
    j = np.tile(np.arange(10, dtype=np.int64), n)
    cols = ((r + 3 * j) % 16) * pw + rng.integers(0, pw, size=r.size)
    A = cs.Coo(r, cols, rng.standard_normal(r.size).astype(np.float32), n, 16 * pw).sorted()
    dA = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=pw)
    assert dA.info.panel_active == 1
    assert n / dA.info.panel_blocks > 1024  # more than 1024 long rows per block
    x = matrices.uniform_vector(A.width, seed=5)
    y0 = matrices.uniform_vector(A.height, seed=6)
    y = cs.DeviceVector(ctx, A.height, y0)
    dA.spmv(cs.DeviceVector(ctx, A.width, x), y)
    ctx.sync()
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    err, worst = oracle.scaled_error(y.download(), y64, bound)
    assert err <= TOL, f"{err:.3e} at row {worst}"


def test_panel_row_too_long_falls_back(ctx, oracle):
    """A row whose segments exceed one pass-2 block: an explicit panel width
    is refused, the automatic choice keeps the plain arrays; both give the
    oracle's result where they run."""
    rng = np.random.default_rng(23)
    # About n_long / 512 segments (one per sub-chunk) against 16384 per
    # pass-2 block.
    n_long = 9_000_000
    rows = np.concatenate([np.zeros(n_long, np.int64), rng.integers(1, 2000, size=20000)])
    cols = rng.integers(0, 4_000_000, size=rows.size)
    A = cs.Coo(rows, cols, rng.standard_normal(rows.size).astype(np.float32), 2000, 4_000_000).sorted()
    with pytest.raises(cs.ClspmvError):
        cs.DeviceMatrix(ctx, A, panel=32760)
    dA = cs.DeviceMatrix(ctx, A, panel=cs.PANEL_AUTO)
    assert dA.info.panel_active == 0 and dA.info.plain == 1
    x = matrices.uniform_vector(A.width, seed=5)
    y = cs.DeviceVector(ctx, A.height, np.zeros(A.height, np.float32))
    dA.spmv(cs.DeviceVector(ctx, A.width, x), y)
    ctx.sync()
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height)
    err, worst = oracle.scaled_error(y.download(), y64, bound)
    assert err <= TOL, f"{err:.3e} at row {worst}"


def test_spmv_multi_refuses_other_formats(ctx):
    A = CASES["rand_small"]
    dA = cs.DeviceMatrix(ctx, cs.coo2ell(A.copy()))
    dX = cs.DeviceVector(ctx, A.width * 2)
    dY = cs.DeviceVector(ctx, A.height * 2)
    with pytest.raises(cs.ClspmvError):
        dA.spmv_multi(2, dX, dY, A.height)


@pytest.mark.parametrize("fmt", ["dia", "csr"])
def test_time_spmv_multi_accumulates(ctx, oracle, fmt):
    """clspmv_time_spmv_multi runs warmup + ntimes passes into the same Y: on
    dyadic data Y = Y0 + (warmup + ntimes) A X exactly, and the time is
    positive."""
    A = CASES["stencil27_dyadic"]
    owned = cs.coo2dia(A) if fmt == "dia" else cs.coo2csr(A.copy())
    k, warmup, ntimes = 5, 2, 3
    rng = np.random.default_rng(8)
    X = (rng.integers(-8, 9, size=(A.width, k)) / 8.0).astype(np.float32)
    Y0 = (rng.integers(-8, 9, size=(k, A.height)) / 8.0).astype(np.float32)
    dA = cs.DeviceMatrix(ctx, owned, panel=0 if fmt == "csr" else None)
    dY = cs.DeviceVector(ctx, Y0.size, Y0.reshape(-1))
    ms = dA.time_multi(k, cs.DeviceVector(ctx, X.size, X.reshape(-1)), dY, A.height, warmup, ntimes)
    assert ms > 0
    Y = dY.download().reshape(k, A.height)
    for q in range(k):
        y64, _ = oracle.spmv_f64(A.row, A.col, A.val, X[:, q].copy(), A.height)
        np.testing.assert_array_equal(Y[q], (Y0[q] + (warmup + ntimes) * y64).astype(np.float32))
    with pytest.raises(cs.ClspmvError):
        dA.time_multi(k, cs.DeviceVector(ctx, X.size, X.reshape(-1)), dY, A.height, 0, 0)
