"""Host conversions to BF16 / FP16 (round to nearest, ties to even) used by
the mixed-precision SpMM, checked against numpy's float16 and, for BF16,
against the two neighbouring representable values."""

import numpy as np
import pytest

import clspmv_mic_amd as cs


def _samples():
    rng = np.random.default_rng(9)
    v = np.concatenate([
        rng.standard_normal(20000).astype(np.float32),
        (rng.standard_normal(5000) * 1e-6).astype(np.float32),  # FP16 subnormals
        (rng.standard_normal(5000) * 1e5).astype(np.float32),   # FP16 overflow
        np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65520.0, 6.1e-5, np.inf, -np.inf], np.float32),
    ])
    return v


def test_f16_matches_numpy():
    v = _samples()
    got = cs.convert_from_f32(v, cs.DT_F16)
    with np.errstate(over="ignore"):  # overflow to inf is the point of some samples
        want = v.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(cs.convert_to_f32(got, cs.DT_F16), want.view(np.float16).astype(np.float32))


def test_bf16_is_nearest_even():
    v = _samples()
    v = v[np.isfinite(v)]
    got = cs.convert_to_f32(cs.convert_from_f32(v, cs.DT_BF16), cs.DT_BF16).astype(np.float64)
    bits = v.view(np.uint32)
    lo = (bits & 0xFFFF0000).view(np.float32).astype(np.float64)         # toward zero
    hi = ((bits & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32).astype(np.float64)
    x = v.astype(np.float64)
    exact = (bits & 0xFFFF) == 0
    assert np.all(got[exact] == x[exact])
    d_lo, d_hi, d_got = np.abs(x - lo), np.abs(x - hi), np.abs(x - got)
    assert np.all((got == lo) | (got == hi))
    assert np.all(d_got <= np.minimum(d_lo, d_hi))
    tie = (d_lo == d_hi) & ~exact
    # ties go to the value whose last kept bit is 0
    kept = cs.convert_from_f32(v[tie], cs.DT_BF16)
    assert np.all((kept & 1) == 0)


def test_bf16_nan_stays_nan():
    v = np.array([np.nan], np.float32)
    assert np.isnan(cs.convert_to_f32(cs.convert_from_f32(v, cs.DT_BF16), cs.DT_BF16)[0])


def test_bad_dtype():
    with pytest.raises(cs.ClspmvError):
        cs.convert_from_f32(np.ones(3, np.float32), cs.DT_F32)
