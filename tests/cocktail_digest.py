"""Digest of a partitioned matrix (clspmv_cocktail / the reference's
cocktail<int,int,float>, same memory layout) for comparing partitions.

Only array regions the reference defines are hashed: b4ell column ids of
padding block rows and sbell column ids of the padding rows of the last
slice are left uninitialised by the reference (src/analyze.cpp:2064-2101,
1833-1879) and are skipped.
"""

from __future__ import annotations

import hashlib

import numpy as np

import clspmv_mic_amd as cs

FLAGS = ["bdia", "dia", "sbell", "bell", "bcsr", "sell", "ell", "csr", "coo"]


def _h(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:24]


def part_digest(name: str, m) -> dict:
    d = {k: v for k, v in cs.struct_header(m).items()}
    a = cs.struct_arrays(m)
    if isinstance(m, cs.B4ellMatrix):
        nbr, hal, ell = m.b4ell_row_num, m.b4ell_height_aligned, m.b4ell_block_num
        col = a["col"].reshape(ell, hal)[:, :nbr] if ell else a["col"]
        a["col"] = col
    if isinstance(m, cs.SbellMatrix):
        sh, nbr = m.sbell_slice_height, m.sbell_row_num
        ptr = a["ptr"]
        keep = []
        for s in range(len(ptr) - 1):
            w = (ptr[s + 1] - ptr[s]) // sh if sh else 0
            rows = min(sh, nbr - s * sh)
            blk = a["col"][ptr[s]:ptr[s + 1]].reshape(w, sh)[:, :rows] if w else np.zeros(0, np.int32)
            keep.append(blk.ravel())
        a["col"] = np.concatenate(keep) if keep else np.zeros(0, np.int32)
    for k, v in a.items():
        d[k] = _h(v)
        d[k + "_n"] = int(v.size)
    return d


def digest(c) -> dict:
    """c: a CocktailMatrix struct."""
    out = {}
    for n in FLAGS:
        out["ifuse" + n] = bool(getattr(c, "ifuse" + n))
        out[n + "_meth_num"] = int(getattr(c, n + "_meth_num"))
    for k in ["csr_group_num", "coo_group_num", "mat_width", "mat_height", "mat_nnz"]:
        out[k] = int(getattr(c, k))
    parts = {}
    for n in FLAGS:
        if getattr(c, "ifuse" + n):
            parts[n] = part_digest(n, getattr(c, n))
    out["parts"] = parts
    if not out["ifusecsr"]:
        out.pop("csr_group_num")   # unset by the reference unless CSR is used
    if not out["ifusecoo"]:
        out.pop("coo_group_num")
    return out
