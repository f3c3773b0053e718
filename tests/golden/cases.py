"""Matrices, table sets and knobs of the golden partition / conversion
fixtures (regenerated bit-identically from their seeds by the tests)."""

import numpy as np

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices


def _thinned_blocks():
    # 2x8 block pattern with 40% of the entries dropped: blocks of mixed
    # density for the dense-block threshold.
    A = matrices.block_stencil(512, 2, 8, seed=7, offsets=(0, 1, -1, 3, 40))
    keep = np.random.default_rng(8).random(A.nnz) >= 0.4
    return cs.Coo(A.row[keep], A.col[keep], A.val[keep], A.height, A.width)


MATRICES = {
    "lap2d_64": lambda: matrices.laplacian_2d(64),
    "stencil27_12": lambda: matrices.stencil_27pt(12, seed=3),
    "powerlaw_20k": lambda: matrices.power_law(20000, seed=2),
    "powerlaw_band": lambda: matrices.power_law(8192, seed=5, max_len=40,
                                                band=(-2, -1, 0, 1, 2, 7, 8, 100, 101, 102, 103)),
    "block_4x4": lambda: matrices.block_stencil(2048, 4, 4, seed=4, offsets=(0, 1, -1, 32, -32, 5)),
    "block_thinned": _thinned_blocks,
    "rand_nonsquare": lambda: matrices.random_sparse(3000, 2000, 0.005, seed=6),
    "rand_long": lambda: matrices.random_sparse(4000, 4000, 0.002, seed=9, long_rows=4, long_len=1500),
}


def matrix(name: str) -> "cs.Coo":
    return MATRICES[name]()


PARTITION_MATRICES = list(MATRICES)
TABLE_SETS = ["dia_first", "bdia_first", "block_first", "sbell_first", "bell_first", "sell_first",
              "csr_first", "coo_first", "knc"]
# (dia_priority_multiplier, estimate_flat_partial, estimate_flat_full):
# the reference's default and the three settings its driver documents
# (src/eval.cpp:85-95).
KNOBS = [(1.0, True, False), (0.5, True, False), (1.0, False, False), (1.0, True, True)]

CONVERSION_MATRICES = ["lap2d_64", "stencil27_12", "powerlaw_band", "block_thinned", "rand_nonsquare"]
CONVERSIONS = ([("csr", ()), ("ell", (32, 0)), ("ell", (7, 0)), ("sell", (32,)), ("sell", (128,)),
                ("dia", (32,)), ("bdia", (32,))] +
               [(c, (bw, bh, 32)) for c in ("bcsr", "b4csr") for bh, bw in cs.BLOCK_SHAPES] +
               [(c, (bw, bh, 32, 0)) for c in ("bell", "b4ell") for bh, bw in cs.BLOCK_SHAPES] +
               [("sbell", (bw, bh, 32)) for bh, bw in cs.BLOCK_SHAPES])
