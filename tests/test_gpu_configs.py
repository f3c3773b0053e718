"""GPU parity at the benchmark workloads' full sizes.

Each test builds one benchmark matrix at the size its workload is quoted on
and checks the engine's result against the FP64 oracle of the whole matrix
(row-scaled error |y - y64| <= 1e-5 (|y0| + |A||x|)), and bit for bit on a
dyadic copy (values and x of the form k/64, so every product and partial
sum is exact in FP32 and the result cannot depend on summation order):

* power-law, 1,000,000 rows (Zipf(1.8) lengths in [1, 5000], uniform random
  columns) + band {0, +-1, +-1000}: analyze_matrix on the shipped MI355X
  tables (DIA + COO), measured tuning, the DIA rows summed inside the
  x-panel first-pass launch and separately, graph and plain launches; the
  COO part alone exercises the encoder's sizing at full size (31 x panels,
  pass-1 work-groups and pass-2 row blocks sized to one round on the CUs);
* 27-point stencil on a 128^3 grid (2,097,152 rows, 55.7 M nonzeros):
  analyze_matrix (DIA, 27 diagonals), tuned;
* 4x4 dense blocks, 524,288 block rows x 12 (100.7 M nonzeros):
  analyze_matrix (BCSR 4x4), tuned;
* BF16 / FP16 block SpMM on that matrix with 8, 16 and 32 right-hand
  sides: tiled and column-major X give identical Y, and sampled columns
  match the oracle applied to the values as stored.
"""

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _dyadic(A):
    v = (np.round(A.val * 4) / 64.0).astype(np.float32)
    v[v == 0] = 1.0 / 64
    return cs.Coo(A.row, A.col, v, A.height, A.width)


def _check(oracle, A, x, y0, y, exact=False, what=""):
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    err, worst = oracle.scaled_error(y, y64, bound)
    print(f"{what}: scaled error {err:.3e} (row {worst})", flush=True)
    assert err <= TOL, f"{what}: scaled error {err:.3e} at row {worst}"
    if exact:
        np.testing.assert_array_equal(y, y64.astype(np.float32), err_msg=what)


def _cocktail(oracle, A, exact, what, fuse_modes=(None,)):
    ck = cs.analyze_matrix(A.copy(), cs.TABLES_DIR)
    if exact:
        x = matrices.dyadic_vector(A.width, seed=11)
        y0 = matrices.dyadic_vector(A.height, seed=12)
    else:
        x = matrices.uniform_vector(A.width, seed=11)
        y0 = matrices.uniform_vector(A.height, seed=12)
    rt = cs.CocktailRuntime(ck, x, y0)
    rt.tune(3)
    np.testing.assert_array_equal(rt.get_y(), y0)  # tuning keeps y
    modes = rt.launch_modes()
    for fuse in fuse_modes:
        if fuse is not None:
            rt.set_fuse(fuse)
        for graph in (True, False):
            rt.set_y(y0)
            rt.set_graph(graph)
            rt.do_spmv(1)
            _check(oracle, A, x, y0, rt.get_y(), exact,
                   f"{what} {ck.summary()} tuned {modes} fuse={fuse} graph={graph}")
    rt.close()
    return ck


@pytest.fixture(scope="module")
def powerlaw():
    return matrices.power_law(1_000_000, seed=2)


def test_powerlaw_1m_cocktail(oracle, powerlaw):
    ck = _cocktail(oracle, powerlaw, False, "power-law 1M", fuse_modes=(False, True))
    summ = ck.summary()
    assert summ.get("dia", 0) > 0 and summ.get("coo", 0) + summ.get("csr", 0) > 0, summ


def test_powerlaw_1m_cocktail_dyadic_exact(oracle, powerlaw):
    _cocktail(oracle, _dyadic(powerlaw), True, "power-law 1M dyadic", fuse_modes=(False, True))


def test_powerlaw_1m_panel_part_sizing(ctx, oracle, powerlaw):
    """The COO remainder of the power-law matrix (15.6 M entries on 1 M
    columns) in the x-panel encoding at full size: every panel staged, the
    pass-1 work-groups and pass-2 row blocks cut to fill the CUs in one
    round, and the result within tolerance of the oracle."""
    ck = cs.analyze_matrix(powerlaw.copy(), cs.TABLES_DIR)
    parts = dict(ck.parts())
    part = parts.get("coo") or parts.get("csr")
    name = "coo" if "coo" in parts else "csr"
    cus = ctx.info()[1]
    dA = cs.DeviceMatrix(ctx, part, panel=cs.PANEL_AUTO if name == "coo" else cs.PANEL_BOTH)
    info = {f: getattr(dA.info, f) for f, _ in cs.DmatInfo._fields_}
    assert info["panel_active"] == 1 and info["panel_width"] > 0, info
    panels = -(-powerlaw.width // info["panel_width"])
    assert panels >= 31, info
    assert 0 < info["panel_groups"] <= cus, info       # pass 1: one round
    assert panels <= info["panel_groups"], info
    assert 0 < info["panel_blocks"] <= 2 * cus, info   # pass 2: one round, two per CU
    arr = cs.struct_arrays(part)
    if name == "coo":
        R, Cc, V = arr["row"], arr["col"], arr["val"]
    else:
        R = np.repeat(np.arange(part.matinfo.height, dtype=np.int32), np.diff(arr["rp"]))
        Cc, V = arr["col"], arr["val"]
    P = cs.Coo(R, Cc, V, part.matinfo.height, part.matinfo.width)
    x = matrices.uniform_vector(P.width, seed=13)
    y0 = matrices.uniform_vector(P.height, seed=14)
    dx, dy = cs.DeviceVector(ctx, P.width, x), cs.DeviceVector(ctx, P.height, y0)
    dA.spmv(dx, dy)
    ctx.sync()
    _check(oracle, P, x, y0, dy.download(), False, f"{name} part in panels {info}")


def test_stencil27_128_cocktail(oracle):
    A = matrices.stencil_27pt(128, seed=3)
    ck = _cocktail(oracle, A, False, "27-point stencil 128^3")
    assert ck.summary() == {"dia": A.nnz}, ck.summary()


def test_stencil27_128_cocktail_dyadic_exact(oracle):
    A = matrices.stencil_27pt(128, seed=3, dyadic=True)
    _cocktail(oracle, A, True, "27-point stencil 128^3 dyadic")


@pytest.fixture(scope="module")
def blocks4x4():
    return matrices.block_stencil(524_288, 4, 4, seed=4)


def test_block4x4_524288_cocktail(oracle, blocks4x4):
    # One block format takes the whole matrix: BCSR or BELL, whichever the
    # tables rate faster at 12 blocks per block row (both run a lane per
    # block row over contiguous slot planes here, within noise of each other).
    ck = _cocktail(oracle, blocks4x4, False, "4x4 blocks 524,288 x 12")
    summ = ck.summary()
    assert len(summ) == 1 and set(summ) <= {"bcsr", "bell"}, summ


def test_block4x4_524288_bcsr(ctx, oracle, blocks4x4):
    """The matrix as BCSR on its own (clspmv_upload_b4csr): the slot-major
    encoding, 12 slots per block row and no padding, against the oracle."""
    A = blocks4x4
    dA = cs.DeviceMatrix(ctx, cs.coo2b4csr(A.copy(), 4, 4, cap=False))
    assert dA.info.block_slots == 12
    x = matrices.uniform_vector(A.width, seed=41)
    y0 = matrices.uniform_vector(A.height, seed=42)
    dy = cs.DeviceVector(ctx, A.height, y0)
    dA.spmv(cs.DeviceVector(ctx, A.width, x), dy)
    ctx.sync()
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    err, worst = oracle.scaled_error(dy.download(), y64, bound)
    assert err <= TOL, f"scaled error {err:.3e} at row {worst}"


@pytest.mark.parametrize("dtype", [cs.DT_BF16, cs.DT_F16])
def test_block_spmm_524288(ctx, oracle, blocks4x4, dtype):
    A = blocks4x4
    H, W = A.height, A.width
    vals_s = cs.convert_to_f32(cs.convert_from_f32(A.val, dtype), dtype)
    M = cs.BlockSpmm(ctx, cs.coo2bcsr(A.copy(), 4, 4), dtype)
    for k in (8, 16, 32):
        rng = np.random.default_rng(50 + k)
        X = rng.uniform(-1, 1, size=(k, W)).astype(np.float32)
        Xb = cs.convert_from_f32(X, dtype)
        Xs = cs.convert_to_f32(Xb, dtype)
        dX = cs.DeviceBuffer(ctx, Xb)
        dX4 = cs.spmm_pack_x(ctx, k, W, dX, W)
        dY = cs.DeviceBuffer(ctx, np.zeros((k, H), np.float32))
        dYt = cs.DeviceBuffer(ctx, np.zeros((k, H), np.float32))
        M.spmm(k, dX, W, dY, H)
        M.spmm_tiled(k, dX4, dYt, H)
        ctx.sync()
        Y, Yt = dY.download(), dYt.download()
        np.testing.assert_array_equal(Y, Yt, err_msg=f"k={k}: tiled X differs from column-major X")
        for j in sorted({0, k // 2, k - 1}):
            y64, bound = oracle.spmv_f64(A.row, A.col, vals_s, Xs[j].copy(), H)
            err, worst = oracle.scaled_error(Y[j].copy(), y64, bound)
            print(f"spmm dtype {dtype} k={k} rhs {j}: scaled error {err:.3e} (row {worst})", flush=True)
            assert err <= TOL, f"k={k} rhs {j}: scaled error {err:.3e} at row {worst}"
        del dX, dX4, dY, dYt
