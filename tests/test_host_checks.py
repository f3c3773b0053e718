"""Host-side argument checks of the Python mirror that run before any call
into the library: the SpMM entry points take raw device pointers, so the
mirror compares buffer sizes with k, ldx / ldy and the tile count and
refuses short buffers (no GPU needed: the check fires before the call)."""

import types

import pytest

import clspmv_mic_amd as cs


def _spmm(rows, cols):
    M = cs.BlockSpmm.__new__(cs.BlockSpmm)
    M.ctx, M.h, M.dtype = None, None, cs.DT_BF16
    M.info = cs.DmatInfo()
    M.info.rows, M.info.cols = rows, cols
    return M


def _buf(nbytes):
    return types.SimpleNamespace(nbytes=nbytes, p=None)


def _vec(n):
    return types.SimpleNamespace(n=n, p=None)


@pytest.mark.parametrize("k", [1, 3, 16])
def test_spmm_short_x_or_y(k):
    M = _spmm(40, 36)
    ldx, ldy = 40, 44
    x_need = 2 * (ldx * (k - 1) + 36)
    y_need = ldy * (k - 1) + 40
    with pytest.raises(cs.ClspmvError) as e:
        M.spmm(k, _buf(x_need - 2), ldx, _vec(y_need), ldy)
    assert e.value.code == cs.ERR_INVALID and "X holds" in str(e.value)
    with pytest.raises(cs.ClspmvError) as e:
        M.spmm(k, _buf(x_need), ldx, _vec(y_need - 1), ldy)
    assert "Y holds" in str(e.value)


def test_spmm_tiled_short_x4():
    M = _spmm(40, 37)  # 10 tiles of 4 rows
    with pytest.raises(cs.ClspmvError) as e:
        M.spmm_tiled(8, _buf(8 * 10 * 8 - 8), _vec(8 * 40), 40)
    assert "X4 holds" in str(e.value)
    with pytest.raises(cs.ClspmvError):
        M.spmm_tiled(8, _buf(8 * 10 * 8), _vec(7 * 40 + 39), 40)


def test_pack_x_short_x():
    with pytest.raises(cs.ClspmvError) as e:
        cs.spmm_pack_x(None, 4, 100, _buf(2 * (100 * 3 + 100) - 2), 100)
    assert "X holds" in str(e.value)
