"""Command-line front of the MI355X autotuner (clspmv_autotune): measures
every kernel variant on this GPU and writes the ten .ben tables the
Cocktail partitioner reads, the counterpart of the reference's
`benchmarking` script (bench_* programs, src/bench_*.cpp).

    python -m clspmv_mic_amd.autotune OUT_DIR [--level 0|1|-1] [--device N] [--note TEXT]

Tables land in OUT_DIR/benchmark/*.ben; OUT_DIR/benchmark/PROVENANCE holds
the device, level, wall time and the note (e.g. the source revision the
library was built from).
"""

import argparse
import os
import sys
import time

from . import Context, autotune


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m clspmv_mic_amd.autotune")
    ap.add_argument("out_dir")
    ap.add_argument("--level", type=int, default=0,
                    help="0: up to 64M stored entries per probe; 1: 256M; -1: reduced grid")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--note", default="")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args(argv)
    os.makedirs(args.out_dir, exist_ok=True)
    name = Context(args.device).info()[0]
    t0 = time.time()
    autotune(args.out_dir, args.device, args.level, args.verbose)
    dt = time.time() - t0
    with open(os.path.join(args.out_dir, "benchmark", "PROVENANCE"), "w") as f:
        f.write(f"device: {name}\nlevel: {args.level}\nseconds: {dt:.1f}\nnote: {args.note}\n")
    print(f"tables written to {args.out_dir}/benchmark in {dt:.1f} s", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
