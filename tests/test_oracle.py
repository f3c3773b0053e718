"""Pins the CPU oracle (oracle/spmv_oracle.c) before anything is checked
against it.

* Closed forms: the 2D 5-point and 3D 27-point Laplacians with x = 1 give
  y_i = centre - (number of neighbours of i), for every storage format.
* Brute force: a dense float64 matrix-vector product in numpy on tiny
  matrices; with dyadic values every FP32 sum is exact, so each format's
  serial SpMV must equal it bit for bit, whatever its summation order.
* The reference itself: include/spmv_serial.h compiled into oracle/_ref
  (where the reference tree exists) must agree bit for bit with the
  restatement on non-dyadic inputs, which pins the loop order too.
* The FP64 checker and the comparison report (src/util.cpp:74-131) on
  hand-computed answers.
* A planted mistake (a transposed operand, a dropped term) fails the checks,
  so they are sharp enough to catch one.
"""

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices
import oracle as oracle_mod

BLOCKS = cs.BLOCK_SHAPES  # (bh, bw) pairs


def _dense(A: cs.Coo) -> np.ndarray:
    D = np.zeros((A.height, A.width), np.float64)
    np.add.at(D, (A.row, A.col), A.val.astype(np.float64))
    return D


def _conversions(A: cs.Coo):
    """(name, owned struct) for every storage format the oracle covers."""
    out = [("coo", None), ("csr", cs.coo2csr(A.copy())), ("ell", cs.coo2ell(A.copy())),
           ("sell32", cs.coo2sell(A.copy(), 32)), ("sell128", cs.coo2sell(A.copy(), 128))]
    d = cs.coo2dia(A.copy())
    if d is not None:
        out.append(("dia", d))
    b = cs.coo2bdia(A.copy())
    if b is not None:
        out.append(("bdia", b))
    for bh, bw in BLOCKS:
        out.append((f"bcsr{bh}x{bw}", cs.coo2bcsr(A.copy(), bw, bh)))
        out.append((f"bell{bh}x{bw}", cs.coo2bell(A.copy(), bw, bh)))
        for nm, fn in (("b4csr", cs.coo2b4csr), ("b4ell", cs.coo2b4ell), ("sbell", cs.coo2sbell)):
            m = fn(A.copy(), bw, bh)
            if m is not None:
                out.append((f"{nm}{bh}x{bw}", m))
    return out


def _oracle_spmv(orc, A: cs.Coo, name: str, owned, x: np.ndarray, y0: np.ndarray) -> np.ndarray:
    """y0 + A x through the oracle's routine for `name`; returns height rows."""
    H, W = A.height, A.width
    pad = 64
    xp = np.zeros(W + pad, np.float32)
    xp[:W] = x
    y = np.zeros(H + 4096, np.float32)
    y[:H] = y0
    if name == "coo":
        orc.coo_spmv(A.row, A.col, A.val, xp, y)
        return y[:H]
    a, h = owned.arrays(), owned.header()
    kind = name.rstrip("0123456789x")
    if kind == "csr":
        orc.csr_spmv(a["rp"], a["col"], a["val"], xp, y)
    elif kind == "ell":
        orc.ell_spmv(H, h["ell_height_aligned"], h["ell_num"], a["col"], a["val"], xp, y)
    elif kind == "sell":
        orc.sell_spmv(H, h["sell_slice_height"], a["ptr"], a["col"], a["val"], xp, y)
    elif kind == "dia":
        yy = np.zeros(max(h["dia_length_aligned"], H), np.float32)
        yy[:H] = y0
        orc.dia_spmv(h["dia_length_aligned"], a["off"], a["val"], x, yy)
        return yy[:H]
    elif kind == "bdia":
        yy = np.zeros(max(h["bdia_length_aligned"], H), np.float32)
        yy[:H] = y0
        orc.bdia_spmv(h["bdia_length_aligned"], a["off"], a["bptr"], a["val"], x, yy)
        return yy[:H]
    elif kind == "bcsr":
        orc.bcsr_spmv(H, W, h["bcsr_bheight"], h["bcsr_bwidth"], h["bcsr_aligned_size"], a["rp"],
                      a["col"], a["val"], xp, y)
    elif kind == "bell":
        orc.bell_spmv(h["bell_row_num"], h["bell_bheight"], h["bell_bwidth"], h["bell_height_aligned"],
                      h["bell_block_num"], a["col"], a["val"], xp, y)
    elif kind == "b4csr":
        orc.b4csr_spmv(H, W, h["b4csr_bheight"], h["b4csr_bwidth"], h["b4csr_aligned_size"], a["rp"],
                       a["col"], a["val"], xp, y)
    elif kind == "b4ell":
        orc.b4ell_spmv(H, W, h["b4ell_row_num"], h["b4ell_bheight"], h["b4ell_bwidth"],
                       h["b4ell_height_aligned"], h["b4ell_float4_aligned"], h["b4ell_block_num"],
                       a["col"], a["val"], xp, y)
    elif kind == "sbell":
        orc.sbell_spmv(H, W, h["sbell_bheight"], h["sbell_bwidth"], h["sbell_slice_height"], a["ptr"],
                       a["col"], a["val"], xp, y)
    else:
        raise KeyError(name)
    return y[:H]


def _stencil27_constant(n: int) -> cs.Coo:
    """27-point Laplacian on an n^3 grid: centre 26, every neighbour -1."""
    m = n ** 3
    r = np.arange(m)
    i, j, k = r % n, (r // n) % n, r // (n * n)
    rows, cols, vals = [], [], []
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                ok = ((i + di >= 0) & (i + di < n) & (j + dj >= 0) & (j + dj < n) &
                      (k + dk >= 0) & (k + dk < n))
                rows.append(r[ok])
                cols.append((r + di + dj * n + dk * n * n)[ok])
                v = 26.0 if (di, dj, dk) == (0, 0, 0) else -1.0
                vals.append(np.full(int(ok.sum()), v, np.float32))
    return cs.Coo(np.concatenate(rows), np.concatenate(cols), np.concatenate(vals), m, m).sorted()


# ---- closed forms -------------------------------------------------------------

def test_laplacian_2d_closed_form(oracle):
    n = 24
    A = matrices.laplacian_2d(n)
    x = np.ones(A.width, np.float32)
    i, j = np.arange(n * n) // n, np.arange(n * n) % n
    boundary = (i == 0).astype(int) + (i == n - 1) + (j == 0) + (j == n - 1)
    expect = boundary.astype(np.float32)  # 4 - (4 - boundary sides)
    assert expect.max() == 2 and expect.min() == 0
    for name, owned in _conversions(A):
        y = _oracle_spmv(oracle, A, name, owned, x, np.zeros(A.height, np.float32))
        np.testing.assert_array_equal(y, expect, err_msg=name)


def test_stencil27_closed_form(oracle):
    n = 7
    A = _stencil27_constant(n)
    x = np.ones(A.width, np.float32)
    r = np.arange(n ** 3)
    cnt = np.ones(n ** 3, np.int64)
    for c in (r % n, (r // n) % n, r // (n * n)):
        cnt *= 3 - (c == 0) - (c == n - 1)
    expect = (26 - (cnt - 1)).astype(np.float32)  # interior 0, corner 26 - 7 = 19
    assert expect.min() == 0 and expect.max() == 19
    for name, owned in _conversions(A):
        y = _oracle_spmv(oracle, A, name, owned, x, np.zeros(A.height, np.float32))
        np.testing.assert_array_equal(y, expect, err_msg=name)


# ---- brute force ----------------------------------------------------------------

TINY = {
    "rand": lambda: matrices.random_sparse(37, 41, 0.15, seed=31, dyadic=True),
    "tall": lambda: matrices.random_sparse(70, 9, 0.3, seed=32, dyadic=True),
    "wide": lambda: matrices.random_sparse(9, 70, 0.3, seed=33, dyadic=True),
    "empty_rows": lambda: matrices.random_sparse(50, 50, 0.08, seed=34, dyadic=True, empty_rows=0.4),
    "banded": lambda: matrices.banded(45, (-3, -1, 0, 1, 2, 9), seed=35, dyadic=True),
    "long_rows": lambda: matrices.random_sparse(40, 300, 0.01, seed=36, dyadic=True, long_rows=2,
                                                long_len=200),
}


@pytest.mark.parametrize("mname", list(TINY))
def test_dyadic_equals_dense(oracle, mname):
    A = TINY[mname]()
    x = matrices.dyadic_vector(A.width, seed=41)
    y0 = matrices.dyadic_vector(A.height, seed=42)
    want = y0.astype(np.float64) + _dense(A) @ x.astype(np.float64)
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    np.testing.assert_array_equal(y64, want)
    np.testing.assert_array_equal(bound, np.abs(y0) + np.abs(_dense(A)) @ np.abs(x.astype(np.float64)))
    for name, owned in _conversions(A):
        y = _oracle_spmv(oracle, A, name, owned, x, y0)
        np.testing.assert_array_equal(y, want.astype(np.float32), err_msg=name)


@pytest.mark.parametrize("mname", ["rand", "long_rows", "banded"])
def test_float_within_tolerance_of_dense(oracle, mname):
    A0 = TINY[mname]()
    rng = np.random.default_rng(7)
    A = cs.Coo(A0.row, A0.col, rng.standard_normal(A0.nnz).astype(np.float32), A0.height, A0.width)
    x = matrices.uniform_vector(A.width, seed=43)
    y0 = matrices.uniform_vector(A.height, seed=44)
    D = _dense(A)
    want = y0.astype(np.float64) + D @ x.astype(np.float64)
    bound = np.abs(y0) + np.abs(D) @ np.abs(x.astype(np.float64))
    for name, owned in _conversions(A):
        y = _oracle_spmv(oracle, A, name, owned, x, y0)
        err = np.max(np.abs(y - want) / np.maximum(bound, 1e-300))
        assert err <= 1e-5, f"{name}: {err}"


# ---- the reference's own serial SpMV -------------------------------------------

needs_ref = pytest.mark.skipif(not oracle_mod.reference_available(),
                               reason="oracle/_ref not built (reference tree absent)")

REF_MATS = {
    "rand": lambda: matrices.random_sparse(300, 257, 0.03, seed=51),
    "banded": lambda: matrices.banded(400, (-9, -1, 0, 1, 3, 40), seed=52),
    "block": lambda: matrices.block_stencil(96, 4, 4, seed=53, offsets=(0, 1, -1, 7)),
    "powerlaw": lambda: matrices.power_law(3000, seed=54, max_len=200, band=(0, 1, -1)),
}


@needs_ref
@pytest.mark.parametrize("mname", list(REF_MATS))
def test_restatement_bit_exact_with_reference(oracle, mname):
    ref = oracle_mod.Reference()
    A = REF_MATS[mname]()
    x = matrices.uniform_vector(A.width, seed=55)
    y0 = matrices.uniform_vector(A.height, seed=56)
    pad = 64
    xp = np.zeros(A.width + pad, np.float32)
    xp[:A.width] = x
    checked = 0
    for name, owned in _conversions(A):
        kind = name.rstrip("0123456789x")
        if kind in ("b4csr", "b4ell"):
            continue  # no serial routine in the reference
        mine = _oracle_spmv(oracle, A, name, owned, x, y0)
        n_out = A.height + 4096
        if kind == "dia":
            n_out = max(owned.header()["dia_length_aligned"], A.height)
        if kind == "bdia":
            n_out = max(owned.header()["bdia_length_aligned"], A.height)
        yr = np.zeros(n_out, np.float32)
        yr[:A.height] = y0
        struct = A.c if name == "coo" else owned.s
        ref.spmv(kind, struct, xp if kind not in ("dia", "bdia") else x, yr)
        np.testing.assert_array_equal(mine, yr[:A.height], err_msg=name)
        checked += 1
    assert checked >= 12


# ---- FP64 checker and comparison report -----------------------------------------

def test_scaled_error_known_answer(oracle):
    y64 = np.array([1.0, -2.0, 0.0, 3.0])
    bound = np.array([2.0, 4.0, 0.0, 1.0])
    y = np.array([1.0, -2.0 + 4 * 3e-6, 0.0, 3.0], np.float32)
    err, worst = oracle.scaled_error(y, y64, bound)
    assert worst == 1 and abs(err - abs(float(y[1]) + 2.0) / 4.0) < 1e-15
    y[2] = 1e-30  # a row with bound 0 must match exactly
    err, worst = oracle.scaled_error(y, y64, bound)
    assert worst == 2 and np.isinf(err)


def test_two_vec_compare_known_answer(oracle):
    a = np.array([1.0, 2.0, -3.0, 0.0, 1.0], np.float32)
    b = np.array([1.0, 2.5, 3.0, 0.0, 1e-8], np.float32)
    r = oracle.two_vec_compare(a, b)
    # Differences are taken in float (src/util.cpp:79,108): 1 - 1e-8 rounds to 1.
    assert r["distance"] == np.sqrt(0.25 + 36.0 + 1.0)
    assert r["max_diff"] == 1.0 and r["max_id"] == 4
    assert r["max_ratio"] == 1.0
    assert r["differ"] == 3


# ---- planted mistakes ------------------------------------------------------------

def test_checks_catch_planted_mistakes(oracle):
    """A transposed operand and a dropped term, planted in the oracle's
    inputs, must both fail the brute-force comparison."""
    A = TINY["rand"]()
    x = matrices.dyadic_vector(max(A.width, A.height), seed=61)
    want = _dense(A) @ x[:A.width].astype(np.float64)
    # transposed: rows and columns swapped (a square slice to keep shapes)
    sq = (A.row < 37) & (A.col < 37)
    At = cs.Coo(A.col[sq], A.row[sq], A.val[sq], 37, 37).sorted()
    As = cs.Coo(A.row[sq], A.col[sq], A.val[sq], 37, 37)
    y_t = _oracle_spmv(oracle, At, "csr", cs.coo2csr(At.copy()), x[:37], np.zeros(37, np.float32))
    assert not np.array_equal(y_t, (_dense(As) @ x[:37].astype(np.float64)).astype(np.float32))
    # dropped term: one product with a nonzero value removed
    keep = np.ones(A.nnz, bool)
    keep[np.flatnonzero((A.val != 0) & (x[A.col] != 0))[0]] = False
    Ad = cs.Coo(A.row[keep], A.col[keep], A.val[keep], A.height, A.width)
    y_d = _oracle_spmv(oracle, Ad, "csr", cs.coo2csr(Ad.copy()), x[:A.width], np.zeros(A.height, np.float32))
    assert not np.array_equal(y_d, want.astype(np.float32))


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_csr_matches_serial(oracle, threads):
    """The all-core baseline's CSR SpMV sums each row in the serial order:
    bit-identical to oracle_csr_spmv (and so to the reference's csr_spmv)."""
    A = matrices.random_sparse(5000, 4000, 0.003, seed=71, long_rows=4, long_len=3000, empty_rows=0.1)
    csr = cs.coo2csr(A.copy())
    arr = csr.arrays()
    x = matrices.uniform_vector(A.width, seed=72)
    y1 = matrices.uniform_vector(A.height, seed=73)
    y2 = y1.copy()
    oracle.csr_spmv(arr["rp"], arr["col"], arr["val"], x, y1)
    oracle.csr_spmv_threads(threads, arr["rp"], arr["col"], arr["val"], x, y2)
    np.testing.assert_array_equal(y1, y2)
    assert oracle.max_threads() >= 1
