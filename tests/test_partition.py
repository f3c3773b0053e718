"""The Cocktail partitioner (analyze_matrix) reproduces the reference's
partition index for index.

Expected values come from the reference's own analyze_matrix compiled from
its sources (oracle/_ref, written by tests/golden/make_golden.py) on the same
matrices, tables and knobs; every output array of every part is compared
through a digest of the regions the reference defines.  When oracle/_ref is
present a second test also compares live against it on fresh cases.
"""

import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import cases  # noqa: E402
import clspmv_mic_amd as cs  # noqa: E402
from cocktail_digest import digest, part_digest  # noqa: E402

GOLD = os.path.join(HERE, "golden")
PARTS = json.load(open(os.path.join(GOLD, "partitions.json")))
CONVS = json.load(open(os.path.join(GOLD, "conversions.json")))

_MATS = {}


def mat(name):
    if name not in _MATS:
        _MATS[name] = cases.matrix(name)
    return _MATS[name]


def _diff(a, b, path=""):
    out = []
    for k in sorted(set(a) | set(b)):
        if k not in a or k not in b:
            out.append(f"{path}{k}: {'missing' if k not in a else 'extra'}")
        elif isinstance(a[k], dict) and isinstance(b[k], dict):
            out += _diff(a[k], b[k], path + k + ".")
        elif a[k] != b[k]:
            out.append(f"{path}{k}: {a[k]} != {b[k]}")
    return out


@pytest.mark.parametrize("key", sorted(PARTS))
def test_partition_matches_reference(key):
    mname, tset, mult, partial, full = key.split("|")
    expected = PARTS[key]
    assert "reference_failed" not in expected
    ck = cs.analyze_matrix(mat(mname).copy(), os.path.join(GOLD, "ben_" + tset), float(mult),
                           bool(int(partial)), bool(int(full)))
    got = digest(ck.c)
    d = _diff(expected, got)
    assert not d, "\n".join(d[:20])


@pytest.mark.parametrize("key", sorted(CONVS))
def test_conversion_matches_reference(key):
    mname, conv, params = key.split("|")
    params = [int(p) for p in params.split(",")] if params else []
    expected = CONVS[key]
    assert "reference_failed" not in expected
    A = mat(mname).copy()
    fn = {"csr": cs.coo2csr, "ell": cs.coo2ell, "sell": cs.coo2sell, "dia": cs.coo2dia,
          "bdia": cs.coo2bdia, "bcsr": cs.coo2bcsr, "b4csr": cs.coo2b4csr, "bell": cs.coo2bell,
          "b4ell": cs.coo2b4ell, "sbell": cs.coo2sbell}[conv]
    out = fn(A, *params)
    if expected.get("refused"):
        assert out is None
        return
    assert out is not None
    got = part_digest(conv, out.s)
    d = _diff(expected, got)
    assert not d, "\n".join(d[:20])
