"""Synthetic .ben table sets for the partitioner's parity tests.

Each set is a directory <name>/benchmark/{overhead,bdia,dia,sbell,bell,bcsr,
sell,ell,csr,coo}.ben in the grammar of src/analyze.cpp:16-175 and the row
order the reference's readers assume.  The sets are shaped to drive the
partitioner down different branches (diagonals first, blocks first, SELL/ELL,
CSR/COO).  The "knc" set reuses the only measured data the reference ships:
its launch overheads (runlog.txt:65-123, the values bench_overhead writes,
src/bench_overhead.cpp:165) and benchmark/bdia.ben, stored here as data.
"""

from __future__ import annotations

import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))

DIMS = [1024 * 2 ** k for k in range(12)]             # 1024 .. 2M rows
BDIA_WIDTHS = [1, 3, 5, 7, 9, 11, 13, 15, 19, 23, 27, 31, 35, 39, 43, 47, 51, 55, 59, 63]
FLAT_WIDTHS = [1, 2, 4, 8, 16, 32, 64, 128]
BLOCKNUMS = [1, 2, 4, 8, 16, 32]
GROUP_WIDTHS = [2, 8, 32, 128, 512, 2048]
GROUPS = [256, 512, 768, 1024]
SHAPES = [(1, 4), (2, 4), (4, 4), (8, 4), (1, 8), (2, 8), (4, 8), (8, 8)]

# Launch overheads of the Xeon Phi run (groups, threads, microseconds),
# runlog.txt:65-123.
KNC_OVERHEAD = [
    (236, 64, 264.667034), (236, 128, 264.492989), (236, 256, 260.092974),
    (472, 64, 291.541100), (472, 128, 289.776087), (472, 256, 293.596983),
    (944, 64, 305.261850), (944, 128, 307.219028), (944, 256, 309.284925),
    (1888, 64, 342.428923), (1888, 128, 339.557886), (1888, 256, 342.459202),
    (3776, 64, 379.523039), (3776, 128, 390.759945), (3776, 256, 389.564037),
    (7552, 64, 468.387842), (7552, 128, 466.649055), (7552, 256, 457.923174),
    (15104, 64, 612.858057), (15104, 128, 614.279032), (15104, 256, 614.061117),
    (30208, 64, 815.468073), (30208, 128, 818.267107), (30208, 256, 819.042921),
    (60416, 64, 1057.935953), (60416, 128, 1059.369087), (60416, 256, 1053.516865),
    (120832, 64, 1399.020910), (120832, 128, 1394.323826), (120832, 256, 1402.035952),
]


def _sat(dim: int, half: float) -> float:
    """Throughput saturating with the row count."""
    return dim / (dim + half)


def _grow(width: float, half: float) -> float:
    return width / (width + half)


def make_set(name: str, peaks: dict, overhead=None, bdia_rows=None, halves=None) -> str:
    """Writes one table set; `peaks` maps format -> list of per-method peak
    GFLOP/s (dia family also scales with band width, flat formats with row
    length).  Returns the set's directory."""
    halves = {**dict(bdia=6.0, dia=2.0, sell=3.0, ell=2.0), **(halves or {})}
    root = os.path.join(HERE, "ben_" + name)
    bdir = os.path.join(root, "benchmark")
    os.makedirs(bdir, exist_ok=True)

    def write(fname, rows):
        with open(os.path.join(bdir, fname), "w") as f:
            for r in rows:
                ints = " ".join(str(int(v)) for v in r[0])
                dbls = " ".join("%f" % v for v in r[1])
                f.write(f"{ints} {dbls}\n")

    if overhead is None:
        overhead = [(g, t, 4.0 + 0.0004 * g + 0.002 * t)
                    for g in [256 * 2 ** k for k in range(10)] for t in (64, 128, 256)]
    with open(os.path.join(bdir, "overhead.ben"), "w") as f:
        for g, t, us in overhead:
            f.write("%d %d %f\n" % (g, t, us))

    def width_rows(fmt, widths, half_w):
        out = []
        for d in DIMS:
            for w in widths:
                out.append(((d, w), [p * _sat(d, 20000) * _grow(w, half_w) * (1 + 0.01 * m)
                                     for m, p in enumerate(peaks[fmt])]))
        return out

    if bdia_rows is None:
        write("bdia.ben", width_rows("bdia", BDIA_WIDTHS, halves["bdia"]))
    else:
        with open(os.path.join(bdir, "bdia.ben"), "w") as f:
            f.write(bdia_rows)
    # dia.ben on bdia.ben's grid (find_min_bdia_width pairs them by row).
    write("dia.ben", width_rows("dia", BDIA_WIDTHS, halves["dia"]))
    write("sell.ben", width_rows("sell", FLAT_WIDTHS, halves["sell"]))
    write("ell.ben", width_rows("ell", FLAT_WIDTHS, halves["ell"]))
    for fmt in ("sbell", "bell", "bcsr"):
        rows = []
        for si, (bh, bw) in enumerate(SHAPES):
            for d in DIMS:
                for b in BLOCKNUMS:
                    scale = (1.0 + 0.05 * si) * _sat(d, 30000) * _grow(b, 3.0)
                    rows.append(((bh, bw, d, b), [p * scale * (1 - 0.02 * m)
                                                 for m, p in enumerate(peaks[fmt])]))
        write(fmt + ".ben", rows)
    for fmt in ("csr", "coo"):
        rows = []
        for d in DIMS:
            for w in GROUP_WIDTHS:
                for gi, g in enumerate(GROUPS):
                    scale = _sat(d, 15000) * _grow(w, 8.0) * (1 + 0.03 * math.sin(gi + d / 1e5))
                    rows.append(((d, w, g), [p * scale * (1 + 0.05 * m * (w > 16))
                                             for m, p in enumerate(peaks[fmt])]))
        write(fmt + ".ben", rows)
    return root


SETS = {
    # Diagonal formats fastest: DIA/BDIA extraction first, flat for the rest.
    "dia_first": dict(bdia=[9, 10, 11, 10.5, 12], dia=[8, 8.5, 9, 9.2], sbell=[5], bell=[4.5, 4.8],
                      bcsr=[4.2], sell=[3.0, 3.2], ell=[2.8, 3.1], csr=[2.5, 2.0], coo=[1.5]),
    # Block formats fastest; flat remainder picks SELL/ELL.
    "block_first": dict(bdia=[3, 3.1, 3.2, 3.3, 3.4], dia=[2.9, 3.0, 3.05, 3.1], sbell=[9],
                        bell=[8.5, 8.8], bcsr=[9.5], sell=[4.0, 4.2], ell=[3.9, 4.1], csr=[3.0, 2.8],
                        coo=[2.0]),
    # SBELL/BELL beat BCSR; diagonals still above flat.
    "sbell_first": dict(bdia=[6, 6.2, 6.4, 6.1, 6.3], dia=[5.5, 5.6, 5.7, 5.8], sbell=[9.8],
                        bell=[9.0, 9.6], bcsr=[7.0], sell=[4.0, 4.3], ell=[3.8, 3.9], csr=[3.5, 3.2],
                        coo=[2.5]),
    # Flat only, CSR/COO preferred over SELL/ELL.
    "csr_first": dict(bdia=[1, 1, 1, 1, 1], dia=[0.9, 0.9, 0.9, 0.9], sbell=[1.5], bell=[1.4, 1.4],
                      bcsr=[1.3], sell=[2.0, 2.1], ell=[1.9, 2.0], csr=[4.0, 4.5], coo=[3.0]),
    # Flat only, COO wins outright.
    "coo_first": dict(bdia=[1, 1, 1, 1, 1], dia=[0.9, 0.9, 0.9, 0.9], sbell=[1.5], bell=[1.4, 1.4],
                      bcsr=[1.3], sell=[2.0, 2.1], ell=[1.9, 2.0], csr=[3.0, 2.9], coo=[5.0]),
}


# Per-set saturation widths (defaults in make_set).
HALVES = {
    "bdia_first": dict(bdia=0.5, dia=4.0),
    "sell_first": dict(sell=1.0, ell=6.0),
}

SETS.update({
    # BDIA beats DIA from band width 1: bands become BDIA.
    "bdia_first": dict(bdia=[12, 13, 14, 13.5, 14.5], dia=[8, 8.5, 9, 9.2], sbell=[5], bell=[4.5, 4.8],
                       bcsr=[4.2], sell=[3.0, 3.2], ell=[2.8, 3.1], csr=[2.5, 2.0], coo=[1.5]),
    # BELL beats SBELL and BCSR in the block family.
    "bell_first": dict(bdia=[2, 2, 2, 2, 2], dia=[1.9, 1.9, 1.9, 1.9], sbell=[6], bell=[10, 10.5],
                       bcsr=[5], sell=[3.0, 3.2], ell=[2.8, 3.1], csr=[2.5, 2.0], coo=[1.5]),
    # Flat only with SELL ahead of ELL, CSR and COO.
    "sell_first": dict(bdia=[1, 1, 1, 1, 1], dia=[0.9, 0.9, 0.9, 0.9], sbell=[1.5], bell=[1.4, 1.4],
                       bcsr=[1.3], sell=[5.0, 5.5], ell=[3.0, 3.5], csr=[2.5, 2.2], coo=[1.5]),
})


def make_all() -> list:
    dirs = [make_set(name, peaks, halves=HALVES.get(name)) for name, peaks in SETS.items()]
    return dirs


def make_knc(bdia_ben_path: str) -> str:
    """The set built from the reference's measured Xeon Phi data."""
    with open(bdia_ben_path) as f:
        bdia_rows = f.read()
    peaks = SETS["dia_first"]
    return make_set("knc", peaks, overhead=KNC_OVERHEAD, bdia_rows=bdia_rows)


if __name__ == "__main__":
    for d in make_all():
        print(d)
