"""GPU test of clspmv_verify_spmv_noalloc, the verification entry point that
allocates nothing (its per-work-group results live in a fixed device array
inside the library).  Its report must equal clspmv_verify_spmv's on the same
inputs: the same worst row, counts and maxima bit for bit (the same kernel
and the same combining order), the RMS to within an ulp (FP64 division and
square root on the GPU instead of the host).  Two calls in a row, the second
on a smaller matrix, show that nothing of the first call's scratch leaks
into the second; the oracle's FP64 SpMV pins the planted worst rows.
"""

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _case(ctx, oracle, A, seed, planted_rel):
    """Device reference, x, y0 and a y with one error planted at a known row."""
    x = matrices.uniform_vector(A.width, seed=seed)
    y0 = matrices.uniform_vector(A.height, seed=seed + 1)
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    y = y64.astype(np.float32)
    rng = np.random.default_rng(seed)
    row = int(rng.choice(np.flatnonzero(bound > 0)))
    y[row] += np.float32(planted_rel * bound[row])
    ref = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=0)
    return (ref, cs.DeviceVector(ctx, A.width, x), cs.DeviceVector(ctx, A.height, y0),
            cs.DeviceVector(ctx, A.height, y), row)


def _same(rep, want):
    for f in ("max_scaled_error", "max_abs_error", "worst_row", "rows", "rows_over_tol", "tol"):
        assert rep[f] == want[f], (f, rep[f], want[f])
    assert rep["rms_scaled_error"] == pytest.approx(want["rms_scaled_error"], rel=1e-15, abs=0)


def test_noalloc_matches_verify_two_calls(ctx, oracle):
    big = matrices.power_law(60_000, seed=12)
    small = matrices.random_sparse(700, 650, 0.02, seed=13, empty_rows=0.2)
    report = cs.DeviceBuffer(ctx, np.full(64, 0xFF, np.uint8))  # poisoned: every field must be written
    for A, seed, rel in ((big, 21, 4e-4), (small, 31, 2e-5)):
        ref, x, y0, y, row = _case(ctx, oracle, A, seed, rel)
        want = cs.verify_spmv(ctx, ref, x, y, y0, tol=TOL)
        cs.verify_spmv_noalloc(ctx, ref, x, y, report, y0, tol=TOL)
        ctx.sync()
        rep = cs.report_from_device(report)
        print(f"noalloc {A.height}x{A.width} nnz={A.nnz}: {rep}", flush=True)
        assert rep["worst_row"] == row
        assert rep["rows_over_tol"] >= 1
        _same(rep, want)


def test_noalloc_without_y0_and_empty(ctx, oracle):
    A = matrices.stencil_27pt(12, seed=3)
    ref, x, _, y, _ = _case(ctx, oracle, A, 41, 0.0)
    report = cs.DeviceBuffer(ctx, np.full(64, 0xFF, np.uint8))
    y.zero()
    dA = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=0)
    dA.spmv(x, y)
    want = cs.verify_spmv(ctx, ref, x, y, None, tol=TOL)
    cs.verify_spmv_noalloc(ctx, ref, x, y, report, None, tol=TOL)
    ctx.sync()
    _same(cs.report_from_device(report), want)
    assert want["max_scaled_error"] <= TOL
    # A matrix without rows: the report says so (worst row -1, zeros).
    E = cs.Coo(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 0, 5)
    eref = cs.DeviceMatrix(ctx, cs.coo2csr(E), panel=0)
    ex = cs.DeviceVector(ctx, 5, np.ones(5, np.float32))
    ey = cs.DeviceVector(ctx, 0)
    cs.verify_spmv_noalloc(ctx, eref, ex, ey, report, None, tol=TOL)
    ctx.sync()
    _same(cs.report_from_device(report), cs.verify_spmv(ctx, eref, ex, ey, None, tol=TOL))


def test_noalloc_rejects_too_many_row_blocks(ctx):
    """A reference with more row blocks than the scratch holds is refused
    before any launch: 256 empty rows make one block."""
    rows = cs.NOALLOC_MAX_ROW_BLOCKS * 256 + 1
    E = cs.Coo(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), rows, 4)
    eref = cs.DeviceMatrix(ctx, cs.coo2csr(E), panel=0)
    x = cs.DeviceVector(ctx, 4, np.ones(4, np.float32))
    y = cs.DeviceVector(ctx, rows)
    report = cs.DeviceBuffer(ctx, np.zeros(64, np.uint8))
    with pytest.raises(cs.ClspmvError) as e:
        cs.verify_spmv_noalloc(ctx, eref, x, y, report, None, tol=TOL)
    assert e.value.code == cs.ERR_TOO_LARGE
