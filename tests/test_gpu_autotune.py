"""GPU test of the MI355X autotuner (clspmv_autotune, the counterpart of the
reference's bench_* table generators, src/bench_*.cpp) on its reduced probe
grid (level -1).

The ten tables it writes must follow the reference's grammar (readers at
src/analyze.cpp:16-175): overhead rows `%d %d %lf`; bdia / dia / sell / ell
rows `%d %d` and one GFLOP/s per method; sbell / bell / bcsr rows
`%d %d %d %d` and one per method, shapes in the order (1,4) (2,4) (4,4)
(8,4) (1,8) (2,8) (4,8) (8,8) (:545-560); csr / coo rows `%d %d %d` and
one per method; dia.ben and bdia.ben on the same row grid (:893).  The
partitions the engine's analyze_matrix computes from the fresh tables must
equal, array for array, those of the reference's own analyze_matrix
(oracle/_ref) on the same tables, and the Cocktail SpMV of each partition
must match the FP64 oracle."""

import os
import sys

import numpy as np
import pytest

import clspmv_mic_amd as cs
import oracle as oracle_mod
from clspmv_mic_amd import matrices

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases  # noqa: E402
from cocktail_digest import digest  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-5
METHODS = {"bdia": 5, "dia": 4, "sell": 2, "ell": 2, "sbell": 1, "bell": 2, "bcsr": 1, "csr": 2, "coo": 1}
KEYS = {"bdia": 2, "dia": 2, "sell": 2, "ell": 2, "sbell": 4, "bell": 4, "bcsr": 4, "csr": 3, "coo": 3}
SHAPES = [(1, 4), (2, 4), (4, 4), (8, 4), (1, 8), (2, 8), (4, 8), (8, 8)]


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("tables"))
    cs.autotune(d, 0, -1, False)
    return d


def _rows(path):
    return [line.split() for line in open(path) if line.strip()]


def test_tables_follow_reference_grammar(tables):
    bench = os.path.join(tables, "benchmark")
    rows = _rows(os.path.join(bench, "overhead.ben"))
    assert rows and all(len(r) == 3 for r in rows)
    for r in rows:
        int(r[0]), int(r[1]), float(r[2])
    grids = {}
    for name, nm in METHODS.items():
        rows = _rows(os.path.join(bench, name + ".ben"))
        assert rows, name
        nk = KEYS[name]
        best = 0.0
        for r in rows:
            assert len(r) == nk + nm, (name, r)
            [int(v) for v in r[:nk]]
            gf = [float(v) for v in r[nk:]]
            # 0 marks a method that cannot run, or a probe over the size cap
            assert all(g >= 0 for g in gf), (name, r)
            best = max(best, max(gf))
        assert best > 0, name
        if nk == 4:  # block tables: shapes in the partitioner's order, then dimension, blocks
            shapes = []
            for r in rows:
                s = (int(r[0]), int(r[1]))
                if not shapes or shapes[-1] != s:
                    shapes.append(s)
            assert shapes == SHAPES, (name, shapes)
        else:
            dims = [int(r[0]) for r in rows]
            assert dims == sorted(dims), name
            grids[name] = [(int(r[0]), int(r[1])) for r in rows]
    assert grids["dia"] == grids["bdia"]


@pytest.mark.skipif(not oracle_mod.reference_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("mname", ["stencil27_12", "powerlaw_band", "block_4x4", "rand_long", "lap2d_64"])
def test_fresh_tables_partition_like_reference(oracle, tables, mname):
    A = cases.matrix(mname)
    ck = cs.analyze_matrix(A.copy(), tables)
    ref = oracle_mod.Reference()
    rc = ref.analyze(A.copy(), tables)
    try:
        assert digest(ck.c) == digest(rc)
    finally:
        ref.free_cocktail(rc)
    assert sum(ck.summary().values()) == A.nnz
    x = matrices.uniform_vector(A.width, seed=3)
    y0 = matrices.uniform_vector(A.height, seed=4)
    rt = cs.CocktailRuntime(ck, x, y0)
    rt.tune(2)
    rt.set_y(y0)
    rt.do_spmv(1)
    y = rt.get_y()
    rt.close()
    y64, bound = oracle.spmv_f64(A.row, A.col, A.val, x, A.height, y0)
    err, worst = oracle.scaled_error(y, y64, bound)
    assert err <= TOL, f"{ck.summary()}: scaled error {err:.3e} at row {worst}"
