"""Synthetic matrices of the engine's benchmark and test workloads.

Every generator is deterministic in its seed (numpy PCG64) and returns a
row-sorted COO without duplicate entries (``clspmv_mic_amd.Coo``).

* ``laplacian_2d(n)``       -- 5-point Laplacian on an n x n grid, diagonal
                               4.0, off-diagonals -1.0.
* ``stencil_27pt(n)``       -- 27-point stencil on an n^3 grid, N(0,1)
                               off-diagonals, diagonally dominant centre
                               (|centre| = sum |off-diagonals| + 1).
* ``power_law(m)``          -- row lengths Zipf(1.8) clipped to [1, 5000],
                               uniform random columns, N(0,1) values, united
                               with a band at offsets {0, +-1, +-1000}.
* ``block_stencil(nbr)``    -- 4x4 dense blocks, 12 blocks per block row at
                               block-column offsets {0, +-1, +-32, +-33,
                               +-4096, +-4097, +31} (mod nbr), N(0,1).
* ``random_sparse(...)``    -- small random matrices for parity tests,
                               optionally with dyadic values (exact FP32
                               sums) and empty / long rows.
"""

from __future__ import annotations

import numpy as np

from . import Coo


def _finish(rows, cols, vals, height, width) -> Coo:
    key = rows.astype(np.int64) * width + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    key = key[order]
    keep = np.ones(key.size, dtype=bool)
    keep[1:] = key[1:] != key[:-1]          # first occurrence wins
    order = order[keep]
    return Coo(rows[order], cols[order], vals[order], height, width)


def laplacian_2d(n: int = 256) -> Coo:
    m = n * n
    r = np.arange(m, dtype=np.int64)
    i, j = r // n, r % n
    parts = [(r, r, np.full(m, 4.0, np.float32))]
    for di, dj in [(-1, 0), (1, 0), (0, -1), (0, 1)]:
        ok = (i + di >= 0) & (i + di < n) & (j + dj >= 0) & (j + dj < n)
        parts.append((r[ok], (r + di * n + dj)[ok], np.full(int(ok.sum()), -1.0, np.float32)))
    rows = np.concatenate([p[0] for p in parts])
    cols = np.concatenate([p[1] for p in parts])
    vals = np.concatenate([p[2] for p in parts])
    return _finish(rows.astype(np.int32), cols.astype(np.int32), vals, m, m)


def stencil_27pt(n: int = 128, seed: int = 3, dyadic: bool = False) -> Coo:
    """27-point stencil; rows in natural (x fastest) order.

    Values are drawn neighbour by neighbour in (dz, dy, dx) order, each
    draw covering the rows that have that neighbour; since the column
    offsets dz*n^2 + dy*n + dx increase in that same order, filling a
    rows x 27 table and dropping the missing neighbours yields the entries
    already sorted by (row, column)."""
    m = n ** 3
    rng = np.random.default_rng(seed)
    r = np.arange(m, dtype=np.int64)
    x, y, z = r % n, (r // n) % n, r // (n * n)
    val = np.zeros((m, 27), np.float32)
    ok_all = np.zeros((m, 27), bool)
    offs = np.zeros(27, np.int64)
    absum = np.zeros(m, np.float64)
    j = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                offs[j] = dx + dy * n + dz * n * n
                if dx == dy == dz == 0:
                    j += 1
                    continue
                ok = ((x + dx >= 0) & (x + dx < n) & (y + dy >= 0) & (y + dy < n) &
                      (z + dz >= 0) & (z + dz < n))
                cnt = int(ok.sum())
                if dyadic:
                    v = (rng.integers(-64, 65, size=cnt) / 64.0).astype(np.float32)
                else:
                    v = rng.standard_normal(cnt).astype(np.float32)
                absum[ok] += np.abs(v)
                val[ok, j] = v
                ok_all[:, j] = ok
                j += 1
    centre = (absum + 1.0).astype(np.float32)
    if dyadic:
        centre = np.ceil(absum + 1.0).astype(np.float32)
    val[:, 13] = centre
    ok_all[:, 13] = True
    del x, y, z, absum
    rows = np.broadcast_to(r.astype(np.int32)[:, None], (m, 27))[ok_all]
    cols = (r[:, None] + offs[None, :]).astype(np.int32)[ok_all]
    return Coo(rows, cols, val[ok_all], m, m)


def power_law(m: int = 1_000_000, seed: int = 2, s: float = 1.8, max_len: int = 5000,
              band=(0, 1, -1, 1000, -1000)) -> Coo:
    rng = np.random.default_rng(seed)
    lengths = np.minimum(rng.zipf(s, size=m), max_len).astype(np.int64)
    lengths = np.minimum(lengths, m)
    rows = np.repeat(np.arange(m, dtype=np.int64), lengths)
    cols = rng.integers(0, m, size=rows.size, dtype=np.int64)
    vals = rng.standard_normal(rows.size).astype(np.float32)
    brow, bcol = [], []
    for off in band:
        r = np.arange(max(0, -off), min(m, m - off), dtype=np.int64)
        brow.append(r)
        bcol.append(r + off)
    brow = np.concatenate(brow) if brow else np.zeros(0, np.int64)
    bcol = np.concatenate(bcol) if bcol else np.zeros(0, np.int64)
    bval = rng.standard_normal(brow.size).astype(np.float32)
    # Band entries first so that they win over random duplicates.
    return _finish(np.concatenate([brow, rows]).astype(np.int32),
                   np.concatenate([bcol, cols]).astype(np.int32),
                   np.concatenate([bval, vals]), m, m)


BLOCK_OFFSETS = (0, 1, -1, 32, -32, 33, -33, 4096, -4096, 4097, -4097, 31)


def block_stencil(nbr: int = 524_288, bh: int = 4, bw: int = 4, seed: int = 4,
                  offsets=BLOCK_OFFSETS, dyadic: bool = False) -> Coo:
    """Block-row pattern with len(offsets) dense bh x bw blocks per block row.

    Values are drawn block by block (block row, then offset, then the
    block's entries row-major); the entries are then laid out by (row,
    column): each block row's blocks sorted by block column.  Where two
    offsets meet the same block column (small nbr), the first block drawn
    is kept."""
    rng = np.random.default_rng(seed)
    nbc = nbr  # square in blocks
    nb = len(offsets)
    bc = (np.arange(nbr, dtype=np.int64)[:, None] + np.array(offsets, dtype=np.int64)[None, :]) % nbc
    if dyadic:
        vals = (rng.integers(-64, 65, size=nbr * nb * bh * bw) / 64.0).astype(np.float32)
    else:
        vals = rng.standard_normal(nbr * nb * bh * bw).astype(np.float32)
    perm = np.argsort(bc, axis=1, kind="stable")
    bcs = np.take_along_axis(bc, perm, axis=1)
    if np.any(bcs[:, 1:] == bcs[:, :-1]):
        # Colliding offsets: general path (sort + first occurrence wins).
        br = np.repeat(np.arange(nbr, dtype=np.int64), nb)
        hh, ww = np.meshgrid(np.arange(bh), np.arange(bw), indexing="ij")
        rows = (br[:, None] * bh + hh.ravel()[None, :]).ravel()
        cols = (bc.ravel()[:, None] * bw + ww.ravel()[None, :]).ravel()
        return _finish(rows.astype(np.int32), cols.astype(np.int32), vals, nbr * bh, nbc * bw)
    V = vals.reshape(nbr, nb, bh, bw)
    V = np.take_along_axis(V, perm[:, :, None, None], axis=1)      # blocks by column
    V = V.transpose(0, 2, 1, 3)                                       # (block row, h, block, w)
    rows = np.repeat(np.arange(nbr * bh, dtype=np.int32), nb * bw)
    cols = (bcs[:, None, :, None] * bw + np.arange(bw)[None, None, None, :])
    cols = np.broadcast_to(cols, (nbr, bh, nb, bw)).astype(np.int32)
    return Coo(rows, cols.ravel(), np.ascontiguousarray(V).ravel(), nbr * bh, nbc * bw)


def random_sparse(height: int, width: int, density: float = 0.05, seed: int = 0,
                  dyadic: bool = False, empty_rows: float = 0.0, long_rows: int = 0,
                  long_len: int = 0) -> Coo:
    rng = np.random.default_rng(seed)
    nnz = int(height * width * density)
    rows = rng.integers(0, height, size=nnz)
    cols = rng.integers(0, width, size=nnz)
    if empty_rows > 0:
        dead = rng.random(height) < empty_rows
        keep = ~dead[rows]
        rows, cols = rows[keep], cols[keep]
    if long_rows > 0 and long_len > 0:
        lr = rng.choice(height, size=min(long_rows, height), replace=False)
        extra_r = np.repeat(lr, long_len)
        extra_c = rng.integers(0, width, size=extra_r.size)
        rows = np.concatenate([rows, extra_r])
        cols = np.concatenate([cols, extra_c])
    if dyadic:
        vals = (rng.integers(-64, 65, size=rows.size) / 64.0).astype(np.float32)
    else:
        vals = rng.standard_normal(rows.size).astype(np.float32)
    return _finish(rows.astype(np.int32), cols.astype(np.int32), vals, height, width)


def banded(n: int, offsets, seed: int = 0, dyadic: bool = False) -> Coo:
    """Square n x n matrix holding full diagonals at the given offsets."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for off in offsets:
        r = np.arange(max(0, -off), min(n, n - off), dtype=np.int64)
        rows.append(r)
        cols.append(r + off)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    if dyadic:
        vals = (rng.integers(-64, 65, size=rows.size) / 64.0).astype(np.float32)
    else:
        vals = rng.standard_normal(rows.size).astype(np.float32)
    return _finish(rows.astype(np.int32), cols.astype(np.int32), vals, n, n)


def serving_mix(nmat: int = 64, seed: int = 6):
    """The serving workload's resident matrices: sizes log-uniform in
    [10k, 2M] rows; even entries stencils (2D 5-point and 3D 27-point in
    turn), odd entries power-law matrices without the band; value types
    FP32 and BF16 alternating in pairs.  Returns (sizes rng, [(kind, bf16,
    COO)]); the rng continues into the request stream."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(2e6), size=nmat)).astype(np.int64)
    out = []
    for i in range(nmat):
        m = int(sizes[i])
        bf16 = (i // 2) % 2 == 1
        if i % 2 == 0:
            if (i // 2) % 4 < 2:
                A = laplacian_2d(max(8, int(round(m ** 0.5))))
            else:
                A = stencil_27pt(max(4, int(round(m ** (1 / 3)))), seed=100 + i)
            out.append(("stencil", bf16, A))
        else:
            out.append(("powerlaw", bf16, power_law(m, seed=100 + i, band=())))
    return rng, out


def dyadic_vector(n: int, seed: int = 1) -> np.ndarray:
    """x with entries k/64, |k| <= 64: products with dyadic values k/64 and
    their sums stay exact in FP32 while partial sums are below 2^24/4096."""
    rng = np.random.default_rng(seed)
    return (rng.integers(-64, 65, size=n) / 64.0).astype(np.float32)


def uniform_vector(n: int, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, size=n).astype(np.float32)
