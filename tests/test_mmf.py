"""Matrix Market ingest (clspmv_read_mmf, the reference's ReadMMF,
src/fileio.cpp:10-111): hand-computed answers, bit-exact agreement with the
reference's own reader (oracle/_ref) where the reference is well defined,
round trips, and refusals."""

import ctypes as C

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices
import oracle as oracle_mod


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


GENERAL = """%%MatrixMarket matrix coordinate real general
% a comment line
%another one

4 5 6
3 2 -1.5
1 1 2.0
4 5 0.25
1 3 1e-3
2 4 3
3 1 7.125
"""

PATTERN = """%%MatrixMarket matrix coordinate pattern general
3 3 4
2 1
1 2
3 3
1 1
"""

SYMMETRIC = """%%MatrixMarket matrix coordinate real symmetric
3 3 4
1 1 4.0
2 1 -1.0
3 2 -2.5
3 3 5.0
"""

SKEW = """%%MatrixMarket matrix coordinate real skew-symmetric
3 3 2
2 1 1.5
3 1 -0.5
"""

INTEGER = """%%MatrixMarket matrix coordinate integer general
2 2 3
1 1 7
2 1 -3
2 2 12
"""


def _arrays(A):
    return A.row.tolist(), A.col.tolist(), A.val.tolist(), A.height, A.width


def test_general_sorted_and_zero_based(tmp_path):
    A = cs.read_mmf(_write(tmp_path, "g.mtx", GENERAL))
    assert _arrays(A) == ([0, 0, 1, 2, 2, 3], [0, 2, 3, 0, 1, 4],
                          [2.0, np.float32(1e-3), 3.0, 7.125, -1.5, 0.25], 4, 5)


def test_pattern_values_are_one(tmp_path):
    A = cs.read_mmf(_write(tmp_path, "p.mtx", PATTERN))
    assert _arrays(A) == ([0, 0, 1, 2], [0, 1, 0, 2], [1.0, 1.0, 1.0, 1.0], 3, 3)


def test_symmetric_expansion_keeps_one_diagonal(tmp_path):
    A = cs.read_mmf(_write(tmp_path, "s.mtx", SYMMETRIC))
    assert _arrays(A) == ([0, 0, 1, 1, 2, 2], [0, 1, 0, 2, 1, 2],
                          [4.0, -1.0, -1.0, -2.5, -2.5, 5.0], 3, 3)


def test_skew_symmetric_mirror_is_negated(tmp_path):
    A = cs.read_mmf(_write(tmp_path, "k.mtx", SKEW))
    assert _arrays(A) == ([0, 0, 1, 2], [1, 2, 0, 0], [-1.5, 0.5, 1.5, -0.5], 3, 3)


def test_integer_values_are_read(tmp_path):
    A = cs.read_mmf(_write(tmp_path, "i.mtx", INTEGER))
    assert _arrays(A) == ([0, 1, 1], [0, 0, 1], [7.0, -3.0, 12.0], 2, 2)


@pytest.mark.parametrize("text, code", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1.0 2.0\n", cs.ERR_UNSUPPORTED),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", cs.ERR_UNSUPPORTED),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", cs.ERR_INVALID),
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n", cs.ERR_IO),
])
def test_refusals(tmp_path, text, code):
    with pytest.raises(cs.ClspmvError) as e:
        cs.read_mmf(_write(tmp_path, "bad.mtx", text))
    assert e.value.code == code


def test_missing_file(tmp_path):
    with pytest.raises(cs.ClspmvError) as e:
        cs.read_mmf(str(tmp_path / "nope.mtx"))
    assert e.value.code == cs.ERR_IO


def test_round_trip_random(tmp_path):
    A = matrices.random_sparse(200, 150, 0.05, seed=71)
    lines = ["%%MatrixMarket matrix coordinate real general", f"{A.height} {A.width} {A.nnz}"]
    order = np.random.default_rng(3).permutation(A.nnz)  # unsorted on disk
    lines += [f"{A.row[i] + 1} {A.col[i] + 1} {float(A.val[i])!r}" for i in order]
    B = cs.read_mmf(_write(tmp_path, "r.mtx", "\n".join(lines) + "\n"))
    S = A.sorted()
    np.testing.assert_array_equal(B.row, S.row)
    np.testing.assert_array_equal(B.col, S.col)
    np.testing.assert_array_equal(B.val, S.val)


def test_symmetric_round_trip_matches_dense(tmp_path):
    rng = np.random.default_rng(5)
    n = 60
    D = np.where(rng.random((n, n)) < 0.1, rng.standard_normal((n, n)), 0.0).astype(np.float32)
    D = np.tril(D) + np.tril(D, -1).T
    r, c = np.nonzero(np.tril(D))
    lines = ["%%MatrixMarket matrix coordinate real symmetric", f"{n} {n} {r.size}"]
    lines += [f"{i + 1} {j + 1} {float(D[i, j])!r}" for i, j in zip(r, c)]
    B = cs.read_mmf(_write(tmp_path, "sym.mtx", "\n".join(lines) + "\n"))
    got = np.zeros((n, n), np.float32)
    got[B.row, B.col] = B.val
    np.testing.assert_array_equal(got, D)


needs_ref = pytest.mark.skipif(not oracle_mod.reference_available(),
                               reason="oracle/_ref not built (reference tree absent)")


@needs_ref
@pytest.mark.parametrize("text", [GENERAL, PATTERN, SYMMETRIC], ids=["general", "pattern", "symmetric"])
def test_matches_reference_reader(tmp_path, text):
    path = _write(tmp_path, "ref.mtx", text)
    lib = C.CDLL(oracle_mod.REF_LIB)
    m = cs.CooMatrix()
    lib.ref_read_mmf(path.encode(), C.byref(m))
    ref = cs.struct_arrays(m)
    libc = C.CDLL(None)
    A = cs.read_mmf(path)
    try:
        assert (m.matinfo.height, m.matinfo.width, m.matinfo.nnz) == (A.height, A.width, A.nnz)
        np.testing.assert_array_equal(ref["row"], A.row)
        np.testing.assert_array_equal(ref["col"], A.col)
        np.testing.assert_array_equal(ref["val"], A.val)
    finally:
        for ptr in (m.coo_row_id, m.coo_col_id, m.coo_data):
            libc.free(C.cast(ptr, C.c_void_p))
