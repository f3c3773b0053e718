"""GPU parity of the mixed-precision MFMA block SpMM (clspmv_spmm): Y += A X
with A's 4x4 blocks and X in BF16 or FP16, accumulated in FP32.

Reference: the FP64 oracle applied to the values as stored (A and X rounded
to 16 bits, exactly representable in FP32), one right-hand side at a time;
bar |Y - Y64| <= 1e-5 (|Y0| + |A||X|) per element, and bit-exact on dyadic
inputs, where every product and partial sum is exact.  The rounding error
of storing A and X in 16 bits is reported by bench.py, not bounded here."""

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _cases():
    return {
        "block_small": matrices.block_stencil(700, 4, 4, seed=6, offsets=(0, 1, -1, 32, -33, 5)),
        "rand": matrices.random_sparse(501, 333, 0.02, seed=41),       # rows/cols not multiples of 4
        "rand_dyadic": matrices.random_sparse(300, 260, 0.05, seed=42, dyadic=True),
        "wide": matrices.random_sparse(37, 1201, 0.03, seed=43),
        "long_rows": matrices.random_sparse(64, 3000, 0.002, seed=44, long_rows=3, long_len=2500),
        # X windows of 101 tiles (two staging rounds at 16 right-hand sides)
        # and of 168 (windows at k <= 8 only, two or three rounds; the LDS
        # of 16 right-hand sides holds at most 108 tiles)
        "win_rounds": matrices.random_sparse(256, 4000, 0.0004, seed=50),
        "win_wide": matrices.random_sparse(256, 4000, 0.0007, seed=45),
    }


CASES = _cases()


def _run(ctx, oracle, A, dtype, k, pad=0):
    dyadic = bool(np.all(A.val * 64 == np.round(A.val * 64)))
    rng = np.random.default_rng(k * 7 + dtype)
    H, W = A.height, A.width
    ldx = (W + 3) // 4 * 4 + pad
    ldy = (H + 3) // 4 * 4 + pad
    if dyadic:
        X = (rng.integers(-64, 65, size=(k, ldx)) / 64.0).astype(np.float32)
        Y0 = (rng.integers(-64, 65, size=(k, ldy)) / 64.0).astype(np.float32)
    else:
        X = rng.uniform(-1, 1, size=(k, ldx)).astype(np.float32)
        Y0 = rng.uniform(-1, 1, size=(k, ldy)).astype(np.float32)
    Xb = cs.convert_from_f32(X, dtype)               # column j of X = row j here
    Xs = cs.convert_to_f32(Xb, dtype)                # the values the GPU sees
    vals_s = cs.convert_to_f32(cs.convert_from_f32(A.val, dtype), dtype)
    bc = cs.coo2bcsr(A.copy(), 4, 4)
    M = cs.BlockSpmm(ctx, bc, dtype)
    assert M.info.format == cs.FMT_BSPMM
    dX = cs.DeviceBuffer(ctx, Xb)
    dY = cs.DeviceBuffer(ctx, Y0)
    M.spmm(k, dX, ldx, dY, ldy)
    ctx.sync()
    Y = dY.download()
    for j in range(k):
        y64, bound = oracle.spmv_f64(A.row, A.col, vals_s, Xs[j, :W].copy(), H, Y0[j, :H].copy())
        err, worst = oracle.scaled_error(Y[j, :H].copy(), y64, bound)
        assert err <= TOL, f"rhs {j}: scaled error {err:.3e} at row {worst}"
        if dyadic:
            np.testing.assert_array_equal(Y[j, :H], y64.astype(np.float32))
        # rows past the matrix (leading-dimension padding) are untouched
        np.testing.assert_array_equal(Y[j, H:], Y0[j, H:])


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("dtype", [cs.DT_BF16, cs.DT_F16])
@pytest.mark.parametrize("k", [1, 3, 4, 8, 16, 32, 40])
def test_spmm_matches_oracle(ctx, oracle, name, dtype, k):
    _run(ctx, oracle, CASES[name], dtype, k)


def test_spmm_leading_dimensions(ctx, oracle):
    _run(ctx, oracle, CASES["rand"], cs.DT_BF16, 8, pad=12)


@pytest.mark.parametrize("k", [3, 8, 20])
def test_spmm_column_major_padding_unused(ctx, k):
    """Column-major X: the rows past the matrix's width (333 columns, so the
    last 4-row tile is partial) are never used, even when they hold NaN --
    the X windows stage them as zeros; Y is bit for bit the one of X with
    zeros there."""
    A = CASES["rand"]
    rng = np.random.default_rng(300 + k)
    H, W = A.height, A.width
    ldx, ldy = (W + 3) // 4 * 4 + 4, (H + 3) // 4 * 4
    X = rng.uniform(-1, 1, size=(k, ldx)).astype(np.float32)
    X[:, W:] = 0.0
    Xn = X.copy()
    Xn[:, W:] = np.nan
    M = cs.BlockSpmm(ctx, cs.coo2bcsr(A.copy(), 4, 4), cs.DT_BF16)
    assert M.info.spmm_window > 0
    ys = []
    for Xi in (X, Xn):
        dX = cs.DeviceBuffer(ctx, cs.convert_from_f32(Xi, cs.DT_BF16))
        dY = cs.DeviceBuffer(ctx, np.zeros((k, ldy), np.float32))
        M.spmm(k, dX, ldx, dY, ldy)
        ctx.sync()
        ys.append(dY.download())
    assert not np.any(np.isnan(ys[1]))
    np.testing.assert_array_equal(ys[1].view(np.uint32), ys[0].view(np.uint32))


def test_spmm_argument_checks(ctx):
    A = CASES["rand"]
    M = cs.BlockSpmm(ctx, cs.coo2bcsr(A.copy(), 4, 4), cs.DT_BF16)
    ldx, ldy = (A.width + 3) // 4 * 4, (A.height + 3) // 4 * 4
    dX = cs.DeviceBuffer(ctx, np.zeros((4, ldx), np.uint16))
    dY = cs.DeviceBuffer(ctx, np.zeros((4, ldy), np.float32))
    for bad in [(4, ldx + 2, ldy), (4, ldx, ldy - 4), (4, A.width - 4, ldy)]:
        with pytest.raises(cs.ClspmvError):
            M.spmm(bad[0], dX, bad[1], dY, bad[2])
    with pytest.raises(cs.ClspmvError):  # 2x4 blocks are not MFMA-shaped
        cs.BlockSpmm(ctx, cs.coo2bcsr(A.copy(), 4, 2), cs.DT_BF16)
    with pytest.raises(cs.ClspmvError):  # SpMV entry refuses the SpMM-only format
        x = cs.DeviceVector(ctx, A.width, np.zeros(A.width, np.float32))
        y = cs.DeviceVector(ctx, A.height)
        check = cs.check
        check(cs.lib.clspmv_spmv(ctx.h, M.h, 0, x.p, y.p), "spmv")


def _tiles(Xb: np.ndarray, cols: int) -> np.ndarray:
    """Host restatement of the 4-row tile layout: value r of rows 4c..4c+3
    of right-hand side j at [(c * k + j) * 4 + r], zeros past `cols`."""
    k = Xb.shape[0]
    nt = (cols + 3) // 4
    full = np.zeros((k, nt * 4), np.uint16)
    full[:, :cols] = Xb[:, :cols]
    return np.ascontiguousarray(full.reshape(k, nt, 4).transpose(1, 0, 2)).reshape(-1)


def _window_max(bc) -> int:
    """Most distinct block columns in one slice of 16 block rows: the X
    window the tiled kernel stages in LDS.  0 (no windows built) when a slice
    has more than 254, when padding every slice's list to the longest would
    more than double the tiles staged, or when a block row is long enough
    for FP64 carries."""
    a = bc.arrays()
    rp, cols = a["rp"], a["col"]
    nbr = len(rp) - 1
    if np.any(np.diff(rp) * 4 > 4096):
        return 0
    sizes = [len(np.unique(cols[rp[s]:rp[min(nbr, s + 16)]])) for s in range(0, nbr, 16)]
    if not sizes or max(sizes) > 254 or len(sizes) * max(sizes) > 2 * sum(sizes) + 64:
        return 0
    return max(sizes)


@pytest.mark.parametrize("name", list(CASES))
def test_spmm_windows(ctx, name):
    """The upload builds each slice's X window (its distinct block columns)
    exactly when every slice has at most 254 and no row needs FP64 carries;
    the cases hold both kinds."""
    bc = cs.coo2bcsr(CASES[name].copy(), 4, 4)
    M = cs.BlockSpmm(ctx, bc, cs.DT_BF16)
    assert M.info.spmm_window == _window_max(bc)
    if name in ("block_small", "rand", "rand_dyadic"):
        assert M.info.spmm_window > 0
    if name == "win_rounds":
        assert 96 < M.info.spmm_window <= 108
    if name == "win_wide":
        assert M.info.spmm_window > 108
    if name in ("wide", "long_rows"):
        assert M.info.spmm_window == 0


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("dtype", [cs.DT_BF16, cs.DT_F16])
@pytest.mark.parametrize("k", [1, 3, 6, 8, 12, 18, 32, 40])
def test_spmm_tiled_matches_column_major(ctx, name, dtype, k):
    """The tiled-X entry point runs the same products in the same order as
    clspmv_spmm: identical Y, bit for bit, with X read through the LDS
    windows (even k, every case with windows; chunks narrower than 16 at
    k = 6, 12, 18, 40) or from the tiles directly (odd k, cases without
    windows); the GPU packer builds the tiles the host restatement does (odd
    column counts pad with zeros)."""
    A = CASES[name]
    rng = np.random.default_rng(100 + k)
    H, W = A.height, A.width
    ldx, ldy = (W + 3) // 4 * 4 + 4, (H + 3) // 4 * 4
    X = rng.uniform(-1, 1, size=(k, ldx)).astype(np.float32)
    Y0 = rng.uniform(-1, 1, size=(k, ldy)).astype(np.float32)
    Xb = cs.convert_from_f32(X, dtype)
    M = cs.BlockSpmm(ctx, cs.coo2bcsr(A.copy(), 4, 4), dtype)
    dX = cs.DeviceBuffer(ctx, Xb)
    dX4 = cs.spmm_pack_x(ctx, k, W, dX, ldx)
    ctx.sync()
    np.testing.assert_array_equal(dX4.download(), _tiles(Xb, W))
    dY = cs.DeviceBuffer(ctx, Y0)
    dY4 = cs.DeviceBuffer(ctx, Y0)
    M.spmm(k, dX, ldx, dY, ldy)
    M.spmm_tiled(k, dX4, dY4, ldy)
    ctx.sync()
    np.testing.assert_array_equal(dY4.download(), dY.download())


def test_spmm_tiled_argument_checks(ctx):
    A = CASES["rand"]
    M = cs.BlockSpmm(ctx, cs.coo2bcsr(A.copy(), 4, 4), cs.DT_BF16)
    dX4 = cs.DeviceBuffer(ctx, np.zeros((A.width + 3) // 4 * 4 * 4, np.uint16))
    dY = cs.DeviceBuffer(ctx, np.zeros(4 * ((A.height + 3) // 4 * 4), np.float32))
    with pytest.raises(cs.ClspmvError):
        M.spmm_tiled(4, dX4, dY, A.height + 1)       # ldy not a multiple of 4
    dm = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()))
    with pytest.raises(cs.ClspmvError):
        lib_spmm = cs.lib.clspmv_spmm_tiled(ctx.h, dm.h, 4, dX4.p, dY.p, (A.height + 3) // 4 * 4)
        cs.check(lib_spmm, "spmm_tiled")             # not a BSPMM matrix


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("k", [1, 8, 16, 32])
def test_spmm_tiled_overwrite(ctx, name, k):
    """clspmv_spmm_tiled_ex with accumulate = 0: Y = A X without reading Y
    (Y starts as NaN) equals, bit for bit, a zeroed Y plus A X."""
    A = CASES[name]
    rng = np.random.default_rng(200 + k)
    H, W = A.height, A.width
    ldx, ldy = (W + 3) // 4 * 4, (H + 3) // 4 * 4 + 4
    X = rng.uniform(-1, 1, size=(k, ldx)).astype(np.float32)
    Xb = cs.convert_from_f32(X, cs.DT_BF16)
    M = cs.BlockSpmm(ctx, cs.coo2bcsr(A.copy(), 4, 4), cs.DT_BF16)
    dX4 = cs.spmm_pack_x(ctx, k, W, cs.DeviceBuffer(ctx, Xb), ldx)
    nan = np.full((k, ldy), np.nan, np.float32)
    dY = cs.DeviceBuffer(ctx, nan)
    dZ = cs.DeviceBuffer(ctx, np.zeros((k, ldy), np.float32))
    M.spmm_tiled(k, dX4, dY, ldy, accumulate=False)
    M.spmm_tiled(k, dX4, dZ, ldy)
    ctx.sync()
    y, z = dY.download().reshape(k, ldy), dZ.download().reshape(k, ldy)
    np.testing.assert_array_equal(y[:, :H].view(np.uint32), z[:, :H].view(np.uint32))
    assert np.all(np.isnan(y[:, H:]))  # past the rows: untouched
