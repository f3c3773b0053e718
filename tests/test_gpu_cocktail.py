"""GPU parity of the Cocktail runtime: partitions from analyze_matrix on the
synthetic table sets, run as one HIP graph per SpMV, against the FP64
oracle of the unpartitioned matrix (tolerance 1e-5 of |y0| + |A||x|, and
bit-exact on dyadic inputs)."""

import os
import sys

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(HERE, "golden")
TOL = 1e-5


def _dyadic_copy(A):
    v = (np.round(A.val * 4) / 64.0).astype(np.float32)
    v[v == 0] = 1.0 / 64
    return cs.Coo(A.row, A.col, v, A.height, A.width)


@pytest.mark.parametrize("tset", cases.TABLE_SETS)
@pytest.mark.parametrize("mname", cases.PARTITION_MATRICES)
def test_cocktail_matches_oracle(oracle, mname, tset):
    A = cases.matrix(mname)
    ck = cs.analyze_matrix(A.copy(), os.path.join(GOLD, "ben_" + tset))
    assert sum(ck.summary().values()) == A.nnz
    x = matrices.uniform_vector(A.width, seed=21)
    y0 = matrices.uniform_vector(A.height, seed=22)
    rt = cs.CocktailRuntime(ck, x, y0)
    rt.do_spmv(1)
    y = rt.get_y()
    rt.close()
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    err, worst = oracle.scaled_error(y, y64, bound)
    assert err <= TOL, f"{ck.summary()}: scaled error {err:.3e} at row {worst}"


@pytest.mark.parametrize("tset", ["dia_first", "bdia_first", "block_first", "bell_first", "sell_first"])
def test_cocktail_dyadic_exact(oracle, tset):
    for mname in ["stencil27_12", "powerlaw_band", "block_thinned"]:
        A = _dyadic_copy(cases.matrix(mname))
        ck = cs.analyze_matrix(A.copy(), os.path.join(GOLD, "ben_" + tset))
        x = matrices.dyadic_vector(A.width, seed=3)
        y0 = matrices.dyadic_vector(A.height, seed=4)
        rt = cs.CocktailRuntime(ck, x, y0)
        rt.do_spmv(1)
        y = rt.get_y()
        rt.close()
        y64, _ = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
        np.testing.assert_array_equal(y, y64.astype(np.float32), err_msg=f"{mname} {ck.summary()}")


def test_iterated_spmv_example(oracle):
    """The reference's example(): y = A^k x with cpy_vector_from_gpu between
    passes (src/eval.cpp:19-72), against k FP64 SpMVs."""
    A = matrices.laplacian_2d(48)
    A = cs.Coo(A.row, A.col, (A.val / 8.0).astype(np.float32), A.height, A.width)
    ck = cs.analyze_matrix(A.copy(), os.path.join(GOLD, "ben_dia_first"))
    x = matrices.uniform_vector(A.width, seed=5)
    rt = cs.CocktailRuntime(ck, x)
    ref = x.astype(np.float64)
    for _ in range(10):
        rt.set_y(np.zeros(A.height, np.float32))
        rt.do_spmv(1)
        rt.y_to_x()
        y64, _ = oracle.spmv_f64(A.row, A.col, A.val, ref.astype(np.float32), A.height)
        ref = y64
    y = rt.get_y()
    rt.close()
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-6)


def test_do_spmv_accumulates(oracle):
    """do_spmv(k) runs k passes of y += A x."""
    A = cases.matrix("powerlaw_20k")
    ck = cs.analyze_matrix(A.copy(), os.path.join(GOLD, "ben_csr_first"))
    x = matrices.uniform_vector(A.width, seed=7)
    rt = cs.CocktailRuntime(ck, x)
    rt.do_spmv(3)
    y = rt.get_y()
    rt.close()
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height)
    err, _ = oracle.scaled_error(y, 3 * y64, 3 * bound)
    assert err <= 3 * TOL


def test_profile_restores_y(oracle):
    A = cases.matrix("stencil27_12")
    ck = cs.analyze_matrix(A.copy(), os.path.join(GOLD, "ben_dia_first"))
    x = matrices.uniform_vector(A.width, seed=8)
    rt = cs.CocktailRuntime(ck, x)
    rt.do_spmv(1)
    before = rt.get_y()
    parts, gms = rt.profile(5)
    assert parts and gms > 0
    np.testing.assert_array_equal(rt.get_y(), before)
    for graph in (True, False):  # cold-cache timing keeps y too, in both launch modes
        rt.set_graph(graph)
        assert rt.time_cold(3, 64 << 20) > 0
        np.testing.assert_array_equal(rt.get_y(), before)
    rt.close()


@pytest.mark.parametrize("mname", ["stencil27_12", "powerlaw_band", "block_thinned"])
def test_tuned_runtime_dyadic_exact(oracle, mname):
    """Measured tuning may pick any allowed method or lane count per part and
    either launch mode; y is kept and every choice stays exact on dyadic
    inputs."""
    A = _dyadic_copy(cases.matrix(mname))
    ck = cs.analyze_matrix(A.copy(), cs.TABLES_DIR)
    x = matrices.dyadic_vector(A.width, seed=5)
    y0 = matrices.dyadic_vector(A.height, seed=6)
    rt = cs.CocktailRuntime(ck, x, y0)
    rt.tune(3)
    np.testing.assert_array_equal(rt.get_y(), y0)
    y64, _ = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    for graph in (True, False):
        rt.set_y(y0)
        rt.set_graph(graph)
        rt.do_spmv(1)
        np.testing.assert_array_equal(rt.get_y(), y64.astype(np.float32), err_msg=f"graph={graph}")
    rt.close()


def test_overlapped_first_pass_identical(oracle):
    """A power-law matrix with a band partitions into DIA + an x-panel COO
    part; the COO first pass runs on a second stream beside the DIA kernel.
    The pass must give the same y, bit for bit, as the one-stream pass, in
    both launch modes and over repeated passes, and stay within the
    tolerance of the FP64 oracle."""
    A = matrices.power_law(500000, seed=7)  # DIA (the band) + COO under the shipped tables
    ck = cs.analyze_matrix(A.copy(), cs.TABLES_DIR)
    summ = ck.summary()
    assert summ.get("dia", 0) > 0 and summ.get("coo", 0) + summ.get("csr", 0) >= 65536, summ
    x = matrices.uniform_vector(A.width, seed=31)
    y0 = matrices.uniform_vector(A.height, seed=32)
    rt = cs.CocktailRuntime(ck, x, y0)
    assert rt.launch_modes()["overlap"], "an x-panel part after the DIA part should overlap"
    outs = {}
    for overlap in (True, False):
        rt.set_overlap(overlap)
        assert rt.launch_modes()["overlap"] == overlap
        for graph in (True, False):
            rt.set_graph(graph)
            rt.set_y(y0)
            rt.do_spmv(3)
            outs[(overlap, graph)] = rt.get_y()
    rt.close()
    ref = outs[(False, False)]
    for key, y in outs.items():
        np.testing.assert_array_equal(y, ref, err_msg=f"overlap, graph = {key}")
    # Three passes from y0: y0 + 3 A x.
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    y64_3 = y64 + 2 * (y64 - y0.astype(np.float64))
    err, worst = oracle.scaled_error(ref, y64_3, 3 * bound)
    assert err <= 3 * TOL, f"scaled error {err:.3e} at row {worst}"


def test_dia_in_first_pass_launch(oracle):
    """The DIA band summed by extra work-groups of the COO part's first-pass
    launch: within the tolerance of the FP64 oracle, the same in graph and
    plain launches, bit-identical to separate launches when the DIA part
    runs method 1 (whose arithmetic the tail uses), and exact on dyadic
    values."""
    A = matrices.power_law(500000, seed=7)  # DIA (the band) + COO under the shipped tables
    ck = cs.analyze_matrix(A.copy(), cs.TABLES_DIR)
    x = matrices.uniform_vector(A.width, seed=41)
    y0 = matrices.uniform_vector(A.height, seed=42)
    rt = cs.CocktailRuntime(ck, x, y0)
    parts, _ = rt.profile(2)
    dia_method = [p["method"] for p in parts if p["format"] == "dia"][0]
    assert not rt.launch_modes()["fuse"]  # off after init
    outs = {}
    for fuse in (False, True):
        rt.set_fuse(fuse)
        assert rt.launch_modes()["fuse"] == fuse
        for graph in (True, False):
            rt.set_graph(graph)
            rt.set_y(y0)
            rt.do_spmv(2)
            outs[(fuse, graph)] = rt.get_y()
    rt.close()
    np.testing.assert_array_equal(outs[(True, True)], outs[(True, False)])
    np.testing.assert_array_equal(outs[(False, True)], outs[(False, False)])
    if dia_method == 1:
        np.testing.assert_array_equal(outs[(True, True)], outs[(False, True)])
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    y64_2 = 2 * y64 - y0.astype(np.float64)
    for key, y in outs.items():
        err, worst = oracle.scaled_error(y, y64_2, 2 * bound)
        assert err <= 2 * TOL, f"fuse, graph = {key}: scaled error {err:.3e} at row {worst}"
    # Dyadic values: every partial sum exact, whatever the order.
    D = _dyadic_copy(A)
    ckd = cs.analyze_matrix(D.copy(), cs.TABLES_DIR)
    xd = matrices.dyadic_vector(D.width, seed=43)
    yd = matrices.dyadic_vector(D.height, seed=44)
    rt = cs.CocktailRuntime(ckd, xd, yd)
    rt.set_fuse(True)
    rt.do_spmv(1)
    got = rt.get_y()
    modes = rt.launch_modes()
    rt.close()
    y64, _ = oracle.spmv_f64(D.row, D.col, D.val, xd, D.height, yd)
    np.testing.assert_array_equal(got, y64.astype(np.float32), err_msg=str(modes))


def test_many_passes_graph_equals_direct(oracle):
    """do_spmv(ntimes) in graph mode launches a graph of 16 passes while
    ntimes allows, then single passes: 37 passes give bit for bit the y of
    37 plain launches, in every launch mode (x-panel first pass on a second
    stream or not, the DIA part fused into the first pass or not), and stay
    within the tolerance of y0 + 37 A x."""
    A = matrices.power_law(500000, seed=7)  # DIA (the band) + COO under the shipped tables
    ck = cs.analyze_matrix(A.copy(), cs.TABLES_DIR)
    x = matrices.uniform_vector(A.width, seed=51)
    y0 = matrices.uniform_vector(A.height, seed=52)
    rt = cs.CocktailRuntime(ck, x, y0)
    n = 37
    outs = {}
    for overlap, fuse in ((True, False), (False, False), (False, True)):
        rt.set_overlap(overlap)
        rt.set_fuse(fuse)
        for graph in (True, False):
            rt.set_graph(graph)
            rt.set_y(y0)
            rt.do_spmv(n)
            outs[(overlap, fuse, graph)] = rt.get_y()
    rt.close()
    for overlap, fuse in ((True, False), (False, False), (False, True)):
        np.testing.assert_array_equal(outs[(overlap, fuse, True)], outs[(overlap, fuse, False)],
                                      err_msg=f"overlap={overlap} fuse={fuse}")
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    want = y0.astype(np.float64) + n * (y64 - y0.astype(np.float64))
    for key, y in outs.items():
        err, worst = oracle.scaled_error(y, want, n * bound)
        assert err <= n * TOL, f"{key}: scaled error {err:.3e} at row {worst}"


def test_failed_tune_leaves_runtime_usable(oracle):
    """A tuning call that fails part-way (a host-side error injected at trial
    n, for every n up to the last trial) returns the error, leaves y as it
    was, and the runtime still computes
    the exact product (dyadic inputs) in both launch modes: the graph is
    captured again from the parts' current encodings."""
    A = _dyadic_copy(cases.matrix("powerlaw_band"))
    ck = cs.analyze_matrix(A.copy(), os.path.join(GOLD, "ben_dia_first"))
    assert len(ck.summary()) >= 3, ck.summary()
    x = matrices.dyadic_vector(A.width, seed=5)
    y0 = matrices.dyadic_vector(A.height, seed=6)
    want = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)[0].astype(np.float32)
    failed = 0
    for n in range(1, 64):
        rt = cs.CocktailRuntime(ck, x, y0)
        os.environ["CLSPMV_TUNE_FAIL_TRIAL"] = str(n)
        try:
            rt.tune(1)
            ok = True
        except cs.ClspmvError as e:
            ok = False
            assert "injected failure" in str(e)
        finally:
            del os.environ["CLSPMV_TUNE_FAIL_TRIAL"]
        np.testing.assert_array_equal(rt.get_y(), y0)
        failed += not ok
        for graph in (True, False):
            rt.set_y(y0)
            rt.set_graph(graph)
            rt.do_spmv(1)
            np.testing.assert_array_equal(rt.get_y(), want, err_msg=f"fail at trial {n}, graph={graph}")
        rt.close()
        if ok:
            break
    assert failed >= 3
