"""Benchmark of the clspmv MI355X engine: autotuned Cocktail SpMV.

One step = one SpMV y += A x over all parts of the partitioned matrix (one
HIP-graph launch).  The default workload is the 1,000,000-row power-law
matrix (row lengths Zipf(1.8) clipped to [1, 5000], uniform random columns,
N(0,1) values) united with a band at offsets {0, +-1, +-1000}; --workload
selects the other configurations.  Matrix, x (uniform[-1,1]) and y are
resident in HBM before timing starts.

Prints one JSON line: GFLOP/s of the whole SpMV (2 * nnz / t), the roofline
of the dominant kernel (algorithmic bytes / its event-timed duration against
8 TB/s), and the reference's serial CSR SpMV timed on the host as the CPU
baseline.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0


def build_matrix(name):
    from clspmv_mic_amd import matrices
    if name == "powerlaw":
        return matrices.power_law(1_000_000, seed=2), "power-law 1M rows, Zipf(1.8) in [1,5000] + band {0,+-1,+-1000}"
    if name == "stencil27":
        return matrices.stencil_27pt(128, seed=3), "27-point stencil on a 128^3 grid (2,097,152 rows)"
    if name == "lap2d":
        return matrices.laplacian_2d(256), "2D 5-point Laplacian on a 256x256 grid (65,536 rows)"
    if name == "block4x4":
        return matrices.block_stencil(524_288, 4, 4, seed=4), "4x4-block matrix, 524,288 block rows x 12 blocks"
    raise SystemExit(f"unknown workload {name}")


def part_bytes(part, rows, cols):
    """Algorithmic HBM bytes of one part per SpMV: true nonzeros only (no
    padding), x read once, y read and written once."""
    nnz = part["nnz"]
    fmt = part["format"]
    vec = 4 * cols + 8 * rows
    if fmt in ("dia", "bdia"):
        return 4 * nnz + vec
    if fmt == "coo":
        return 12 * nnz + vec
    if fmt in ("bcsr", "bell", "sbell"):
        return 4 * nnz + part.get("block_index_bytes", 0) + vec
    return 8 * nnz + (4 * (rows + 1) if fmt == "csr" else 0) + vec


def cpu_baseline(A, x, y_gpu, budget_s=10.0):
    """The reference's serial csr_spmv (include/spmv_serial.h:199) compiled
    from its sources (oracle/_ref), else the oracle's restatement, on the
    whole matrix for as many iterations as fit in budget_s."""
    import ctypes as C
    import numpy as np
    import clspmv_mic_amd as cs
    import oracle
    csr = cs.coo2csr(A.copy())
    arrs = csr.arrays()
    y = np.zeros(A.height, np.float32)
    kind = "port"
    if oracle.reference_available():
        ref = oracle.Reference()
        run = lambda: ref.spmv("csr", csr.s, x, y)  # noqa: E731
        kind = "reference"
    else:
        orc = oracle.Oracle()
        run = lambda: orc.csr_spmv(arrs["rp"], arrs["col"], arrs["val"], x, y)  # noqa: E731
    t0 = time.perf_counter()
    run()
    t1 = time.perf_counter()
    single = max(t1 - t0, 1e-6)
    y_first = y.copy()
    iters = max(2, min(200, int(budget_s / single)))
    t0 = time.perf_counter()
    for _ in range(iters):
        run()
    t1 = time.perf_counter()
    per = (t1 - t0) / iters
    # GPU result against the FP64 oracle, scaled by |A||x| per row.
    orc = oracle.Oracle()
    y64, bound = orc.spmv_f64(A.row, A.col, A.val, x, A.height)
    err, _ = orc.scaled_error(y_gpu, y64, bound)
    return {
        "value": round(2.0 * A.nnz / per / 1e9, 4),
        "unit": "GFLOP/s",
        "cores": 1,
        "kind": kind,
        "sample": f"whole matrix ({A.nnz} nnz), {iters} serial csr_spmv calls, {per * 1e3:.2f} ms each",
        "gpu_scaled_error_vs_fp64": err,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="powerlaw", choices=["powerlaw", "stencil27", "lap2d", "block4x4"])
    ap.add_argument("--tables", default=None, help="directory holding benchmark/*.ben")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the tables' methods and graph launches")
    ap.add_argument("--traffic", default=None, help="JSON of measured HBM bytes per launch by kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); nothing is measured on more "
              "GPUs in this build.", file=sys.stderr)
        return 2

    import numpy as np
    import clspmv_mic_amd as cs
    from clspmv_mic_amd import matrices

    t0 = time.time()
    A, desc = build_matrix(args.workload)
    print(f"[bench] matrix {args.workload}: {A.height}x{A.width} nnz={A.nnz} ({time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    tables = args.tables or cs.TABLES_DIR
    t0 = time.time()
    ck = cs.analyze_matrix(A.copy(), tables)
    t_analyze = time.time() - t0
    summary = ck.summary()
    print(f"[bench] partition {summary} ({t_analyze:.1f}s)", file=sys.stderr, flush=True)
    x = matrices.uniform_vector(A.width, seed=1)
    rt = cs.CocktailRuntime(ck, x)
    graph = True
    if not args.no_tune:
        t0 = time.time()
        graph = rt.tune(20)
        print(f"[bench] tuned methods, launch mode {'graph' if graph else 'direct'} ({time.time() - t0:.1f}s)",
              file=sys.stderr, flush=True)

    # One pass from y = 0 for the CPU cross-check, then warm-up and timing.
    rt.do_spmv(1)
    y_one = rt.get_y()
    rt.do_spmv(args.warmup)
    rt.sync()
    t0 = time.perf_counter()
    rt.do_spmv_async(args.steps)
    rt.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    gflops = 2.0 * A.nnz / (ms_per_step * 1e-3) / 1e9

    parts, graph_ms = rt.profile(max(10, min(100, args.steps)))
    for p, (name, s) in zip(parts, ck.parts()):
        if name in ("bcsr",):
            p["block_index_bytes"] = 4 * s.b4csr_block_num + 4 * (s.b4csr_row_num + 1)
        elif name == "bell":
            p["block_index_bytes"] = 4 * s.b4ell_row_num * s.b4ell_block_num
        elif name == "sbell":
            p["block_index_bytes"] = 4 * (s.sbell_slice_ptr[s.sbell_slice_num] if s.sbell_slice_num else 0)
    for p in parts:
        p["bytes"] = part_bytes(p, A.height, A.width)
        p["gbps"] = p["bytes"] / (p["ms"] * 1e-3) / 1e9 if p["ms"] > 0 else 0.0
    dom = max(parts, key=lambda p: p["ms"])
    total_bytes = 8 * A.nnz + 4 * (A.height + 1) + 4 * A.width + 8 * A.height
    traffic = None
    if args.traffic and os.path.exists(args.traffic):
        meas = json.load(open(args.traffic))
        traffic = meas.get(dom["format"])
    rt.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(A, x, y_one)

    out = {
        "metric": "SpMV GFLOP/s & HBM GB/s (% roofline); served requests/s + p99 latency",
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "rows": A.height, "cols": A.width, "nnz": A.nnz,
                   "partition": summary, "tables": os.path.relpath(tables, ROOT),
                   "launch": "graph" if graph else "direct", "tuned": not args.no_tune},
        "hbm_gbps_algorithmic": round(total_bytes / (ms_per_step * 1e-3) / 1e9, 2),
        "roofline": {
            "bound": "hbm",
            "kernel": f"{dom['format']} method {dom['method']}",
            "achieved": round(dom["gbps"], 2),
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": round(dom["gbps"] / PEAK_HBM_GBPS, 4),
            "traffic": traffic,
            "algorithmic_bytes": dom["bytes"],
            "kernel_ms": round(dom["ms"], 6),
        },
        "parts": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in p.items()} for p in parts],
        "graph_ms": round(graph_ms, 6),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
