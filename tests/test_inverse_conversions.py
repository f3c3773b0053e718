"""The X -> COO inverse conversions (include/matrix_storage.h:826-1901) and
coo2diaext (:1054).

Pinned two ways:
  - against the reference's own routines compiled from its sources
    (oracle/_ref), array for array, on the forward conversions' output
    (skipped where the reference tree was not available to build it);
    csr2coo has no such check: the reference's template calls
    if_sorted_coo on the CSR struct (:1241-1244) and does not compile;
  - by the round trip the formats define: X2coo(coo2X(A)) is A sorted by
    (row, column), checked against a numpy restatement.  csr2coo copies
    every stored entry (:1255-1265); the other inverses skip stored zeros
    (which also removes the formats' padding), so explicit zeros of A are
    dropped; DIA reads columns only below dia_length (:845), so for
    matrices with more columns than rows the inverse keeps the entries left
    of that bound.  The reference compares only on matrices without
    explicit zeros: for those its inverses (other than csr2coo) keep the
    input's nnz in the header while writing fewer entries (ell2coo asserts
    they match, :1334), leaving the arrays' tail unset.
"""

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

NAMES = ["lap2d_64", "stencil27_12", "powerlaw_band", "block_thinned", "rand_nonsquare"]


def _with_zeros(A):
    """A with a few stored zeros, which the inverses drop."""
    v = A.val.copy()
    v[::17] = 0.0
    return cs.Coo(A.row, A.col, v, A.height, A.width)


def _mats():
    out = {n: cases.matrix(n) for n in NAMES}
    out["lap2d_zeros"] = _with_zeros(cases.matrix("lap2d_64"))
    out["tall_zeros"] = _with_zeros(matrices.random_sparse(700, 300, 0.02, seed=5))
    return out


MATS = _mats()


def _expected(A, col_bound=None):
    """A sorted by (row, col) without its zeros (and without columns at or
    past col_bound)."""
    S = A.sorted()
    keep = S.val != 0
    if col_bound is not None:
        keep &= S.col < col_bound
    return S.row[keep], S.col[keep], S.val[keep]


def _forward(name, A):
    if name == "csr":
        return cs.coo2csr(A.copy())
    if name == "ell":
        return cs.coo2ell(A.copy())
    if name == "dia":
        return cs.coo2dia(A.copy())
    if name == "diaext":
        return cs.coo2diaext(A.copy())
    if name == "bdia":
        return cs.coo2bdia(A.copy())
    if name.startswith("bcsr"):
        bh, bw = map(int, name[4:].split("x"))
        return cs.coo2bcsr(A.copy(), bw, bh)
    if name.startswith("bell"):
        bh, bw = map(int, name[4:].split("x"))
        return cs.coo2bell(A.copy(), bw, bh)
    raise KeyError(name)


FORMS = ["csr", "ell", "dia", "diaext", "bdia", "bcsr1x4", "bcsr4x4", "bcsr8x8", "bell2x4", "bell4x8"]


@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("form", FORMS)
def test_round_trip(mname, form):
    A = MATS[mname]
    owned = _forward(form, A)
    if owned is None:  # DIA / BDIA refused by the padding rule
        assert form in ("dia", "diaext", "bdia")
        return
    B = cs.to_coo(owned)
    assert (B.height, B.width) == (A.height, A.width)
    bound = None
    if form in ("dia", "bdia") and A.width > A.height:
        bound = owned.s.dia_length if form == "dia" else owned.s.bdia_length
    if form == "csr":
        S = A.sorted()
        r, c, v = S.row, S.col, S.val
    else:
        r, c, v = _expected(A, bound)
    np.testing.assert_array_equal(B.row, r)
    np.testing.assert_array_equal(B.col, c)
    np.testing.assert_array_equal(B.val, v)


REF = oracle_mod.reference_available()


@pytest.mark.skipif(not REF, reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("form", [f for f in FORMS if f != "csr"])
def test_inverse_matches_reference(mname, form):
    ref = oracle_mod.Reference()
    A = MATS[mname]
    if np.any(A.val == 0):
        pytest.skip("explicit zeros: the reference's inverse leaves part of its output unset")
    owned = _forward(form, A)
    if owned is None:
        return
    B = cs.to_coo(owned)
    kind = "bcsr" if form.startswith("bcsr") else "bell" if form.startswith("bell") else form
    row, col, val, h, w = ref.to_coo(kind, owned.s)
    assert (B.height, B.width) == (h, w)
    np.testing.assert_array_equal(B.row, row)
    np.testing.assert_array_equal(B.col, col)
    np.testing.assert_array_equal(B.val, val)


@pytest.mark.skipif(not REF, reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("mname", list(MATS))
def test_coo2diaext_matches_reference(mname):
    ref = oracle_mod.Reference()
    A = MATS[mname]
    ours = cs.coo2diaext(A.copy())
    rc, theirs = ref.convert("diaext", A.copy(), cs.GPU_ALIGNMENT)
    assert (ours is not None) == bool(rc)
    if ours is None:
        return
    assert cs.struct_header(ours.s) == cs.struct_header(theirs)
    a, b = ours.arrays(), cs.struct_arrays(theirs)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
