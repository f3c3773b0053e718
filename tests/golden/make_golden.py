"""Writes the golden fixtures of the partitioner and converter parity tests.

Runs the reference's own CPU code -- analyze_matrix and the coo2* converters
compiled from /root/reference into oracle/_ref by oracle/Makefile -- on the
synthetic matrices of tests/golden/cases.py and the table sets of
tests/golden/tables.py, and stores digests of every output array
(tests/golden/partitions.json, tests/golden/conversions.json).  Each case
runs in a child process, so an abort inside the reference (it asserts on
internal invariants) is recorded instead of ending the run.

Usage (in the container that holds the reference tree):
    make -C oracle && python tests/golden/make_golden.py
"""

import json
import multiprocessing as mp
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import cases  # noqa: E402
import tables  # noqa: E402


def _partition(args):
    mname, tset, knobs = args
    from oracle import Reference
    from cocktail_digest import digest
    ref = Reference()
    A = cases.matrix(mname)
    c = ref.analyze(A, os.path.join(HERE, "ben_" + tset), *knobs)
    d = digest(c)
    ref.free_cocktail(c)
    return d


def _conversion(args):
    mname, conv, params = args
    from oracle import Reference
    from cocktail_digest import part_digest
    ref = Reference()
    A = cases.matrix(mname)
    rc, s = ref.convert(conv, A, *params)
    if conv in ("dia", "bdia", "b4csr", "b4ell", "sbell") and rc == 0:
        return {"refused": True}
    return part_digest(conv, s)


def run_isolated(fn, arg):
    with mp.get_context("fork").Pool(1) as pool:
        try:
            return pool.apply(fn, (arg,))
        except Exception as e:  # child died or raised
            return {"reference_failed": repr(e)}


def main():
    tables.make_all()
    tables.make_knc("/root/reference/benchmark/bdia.ben")
    parts = {}
    for mname in cases.PARTITION_MATRICES:
        for tset in cases.TABLE_SETS:
            for knobs in cases.KNOBS:
                key = f"{mname}|{tset}|{knobs[0]}|{int(knobs[1])}|{int(knobs[2])}"
                parts[key] = run_isolated(_partition, (mname, tset, knobs))
                print(key, "ok" if "reference_failed" not in parts[key] else parts[key], flush=True)
    with open(os.path.join(HERE, "partitions.json"), "w") as f:
        json.dump(parts, f, indent=1, sort_keys=True)
    convs = {}
    for mname in cases.CONVERSION_MATRICES:
        for conv, params in cases.CONVERSIONS:
            key = f"{mname}|{conv}|{','.join(map(str, params))}"
            convs[key] = run_isolated(_conversion, (mname, conv, params))
            print(key, "ok" if "reference_failed" not in convs[key] else convs[key], flush=True)
    with open(os.path.join(HERE, "conversions.json"), "w") as f:
        json.dump(convs, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
