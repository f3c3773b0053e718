"""GPU tests of the per-format drivers (spmv_<fmt>, include/spmv_<fmt>.h:7),
the Cocktail evaluate() check, and the CSR vector kernel at 1, 2 and 4
lanes per row.

Drivers: every variant of every format is checked inside the driver against
the serial COO SpMV on the host, as the reference's drivers do
(src/spmv_bdia.cpp:118); the report must show no variant above the 1e-5
row-scaled tolerance, the best variant's y must match the FP64 oracle for
x = 1, and for one-variant drivers the report's check figures must equal
the oracle's two_vec_compare of the same two vectors.  evaluate: its report
equals the oracle's two_vec_compare of the serial COO result and the GPU's y
(src/spmv_cocktail.cpp:1134-1153).  CSR lanes: matrices whose mean row
length is at most 1, 2 and 4 make the vector kernel run 1, 2 and 4 lanes
per row (rows longer than the lanes stride through them)."""

import os
import sys

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

pytestmark = pytest.mark.gpu

TOL = 1e-5

MATS = {
    "lap2d_64": lambda: matrices.laplacian_2d(64),
    "rand_long": lambda: matrices.random_sparse(3000, 2500, 0.004, seed=11, long_rows=3, long_len=900),
    "block_small": lambda: matrices.block_stencil(300, 4, 4, seed=6, offsets=(0, 1, -1, 32, -33, 5)),
    "rand_unsorted": lambda: matrices.random_sparse(800, 900, 0.01, seed=12),
}

# Formats each matrix is refused in, by the reference's own rules (DIA and
# BDIA: padded size >= 2 nnz).
REFUSED = {"dia": {"rand_long", "rand_unsorted", "block_small"},
           "bdia": {"rand_long", "rand_unsorted", "block_small"}}


def _x1_reference(oracle, A):
    x = np.ones(A.width, np.float32)
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height)
    coores = np.zeros(A.height, np.float32)
    S = A.sorted()
    oracle.coo_spmv(S.row, S.col, S.val, x, coores)
    return y64, bound, coores


@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("driver", cs.DRIVERS)
def test_driver_checks_every_variant(oracle, mname, driver):
    A = MATS[mname]()
    if mname == "rand_unsorted":  # the driver sorts the caller's matrix in place
        perm = np.random.default_rng(3).permutation(A.nnz)
        A = cs.Coo(A.row[perm], A.col[perm], A.val[perm], A.height, A.width)
    coo = A.copy()
    if mname in REFUSED.get(driver, ()):
        with pytest.raises(cs.ClspmvError):
            cs.run_driver(driver, coo, ntimes=2)
        return
    rep, y = cs.run_driver(driver, coo, ntimes=2)
    assert rep.variants >= 1
    assert rep.variants_over_tol == 0, (rep.max_scaled_error, rep.variants)
    assert rep.max_scaled_error <= TOL
    assert np.all(np.diff(coo.row.astype(np.int64) * coo.width + coo.col) > 0)  # sorted in place
    y64, bound, coores = _x1_reference(oracle, A)
    err, worst = oracle.scaled_error(y, y64, bound)
    assert err <= TOL, f"{driver} best variant: scaled error {err:.3e} at row {worst}"
    if rep.variants == 1:
        ref = oracle.two_vec_compare(coores, y)
        assert rep.check_distance == pytest.approx(ref["distance"], rel=1e-12, abs=0)
        assert rep.check_max_diff == ref["max_diff"]
        assert rep.check_max_ratio == ref["max_ratio"]


def test_csr_scalar_is_one_lane_per_row(oracle):
    """spmv_csr_scalar runs the CSR kernel with one lane per row on a matrix
    whose rows are long (the vector driver would pick 16 or more lanes)."""
    A = matrices.random_sparse(2000, 2000, 0.02, seed=21, long_rows=2, long_len=1500)
    rep, y = cs.run_driver("csr_scalar", A.copy(), ntimes=2)
    rep_v, y_v = cs.run_driver("csr_vector", A.copy(), ntimes=2)
    y64, bound, _ = _x1_reference(oracle, A)
    for yy in (y, y_v):
        err, worst = oracle.scaled_error(yy, y64, bound)
        assert err <= TOL
    assert rep.variants == 1 and rep_v.variants == 2


@pytest.mark.parametrize("mean_len,lanes", [(0.8, 1), (1.7, 2), (3.5, 4)])
@pytest.mark.parametrize("dyadic", [False, True])
def test_csr_vector_few_lanes(ctx, oracle, mean_len, lanes, dyadic):
    """CSR method 0 with the lane count its upload picks for short rows (the
    power of two at or above the mean row length), including a few rows far
    longer than the lanes and empty rows."""
    n = 5000
    A = matrices.random_sparse(n, 4000, mean_len / 4000, seed=int(mean_len * 10), dyadic=dyadic,
                               empty_rows=0.1, long_rows=3, long_len=300)
    assert int(np.ceil(A.nnz / A.height)) <= lanes
    x = (matrices.dyadic_vector if dyadic else matrices.uniform_vector)(A.width, seed=4)
    y0 = (matrices.dyadic_vector if dyadic else matrices.uniform_vector)(A.height, seed=5)
    dA = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()))
    dx, dy = cs.DeviceVector(ctx, A.width, x), cs.DeviceVector(ctx, A.height, y0)
    dA.spmv(dx, dy, 0)
    ctx.sync()
    y = dy.download()
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    err, worst = oracle.scaled_error(y, y64, bound)
    assert err <= TOL, f"lanes {lanes}: scaled error {err:.3e} at row {worst}"
    if dyadic:
        np.testing.assert_array_equal(y, y64.astype(np.float32))


@pytest.mark.parametrize("mname", ["lap2d_64", "rand_long"])
def test_evaluate_reports_reference_check(oracle, mname):
    """clspmv_evaluate (evaluate, src/spmv_cocktail.cpp:1431 with
    check_and_time :1134): one SpMV from res, the reference's comparison
    with the caller's serial result, then timed passes."""
    A = MATS[mname]()
    ck = cs.analyze_matrix(A.copy(), cs.TABLES_DIR)
    x = matrices.uniform_vector(A.width, seed=7)
    coores = np.zeros(A.height, np.float32)
    S = A.sorted()
    oracle.coo_spmv(S.row, S.col, S.val, x, coores)
    res = np.zeros(A.height, np.float32)
    rep = cs.evaluate(ck, x, res, coores)
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height)
    err, _ = oracle.scaled_error(res, y64, bound)
    assert err <= TOL
    ref = oracle.two_vec_compare(coores, res)
    assert rep["l2_distance"] == pytest.approx(ref["distance"], rel=1e-12, abs=0)
    assert rep["max_diff"] == ref["max_diff"] and rep["max_ratio"] == ref["max_ratio"]
    assert rep["ms_per_spmv"] > 0 and rep["gflops"] > 0
