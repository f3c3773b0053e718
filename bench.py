"""Benchmark of the clspmv MI355X engine: autotuned Cocktail SpMV.

One step = one SpMV y += A x over all parts of the partitioned matrix.  The
default workload is the 27-point stencil on a 128^3 grid (2,097,152 rows,
55.7 M nonzeros, N(0,1) off-diagonals, diagonally dominant centre); the
same JSON line carries the 1,000,000-row power-law matrix (row lengths
Zipf(1.8) clipped to [1, 5000], uniform random columns, N(0,1) values,
united with a band at offsets {0, +-1, +-1000}) under "powerlaw".
--workload selects one configuration alone, or the SpMM, serving and
per-format sweeps.  Matrix, x (uniform[-1,1]) and y are resident in HBM
before timing starts.

Prints one JSON line: GFLOP/s of the whole SpMV (2 * nnz / t), the roofline
of the dominant kernel (the bytes its format must move / its event-timed
duration, against 8 TB/s; the byte count that also credits a column index
per nonzero is kept beside it), the same pass with the caches scrubbed
before every pass, and the CPU baseline on the same matrix: the
reference's serial csr_spmv on one core and a row-parallel CSR SpMV over
all host threads.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0


SWEEP_FORMATS = ["csr_scalar", "csr_vector", "coo", "sell", "ell", "dia", "bdia", "bcsr", "bell", "sbell"]


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
    """Algorithmic HBM bytes of one part per SpMV as the roofline counts
    them: 8 bytes per true nonzero (value + column index) + 4(m+1) + x read
    once + y read and written once; block formats count their dense blocks'
    values, 4 bytes per block and 4(mb+1) instead.  Padding is never
    credited."""
    nnz = part["nnz"]
    fmt = part["format"]
    vec = 4 * cols + 8 * rows
    if fmt in ("bcsr", "bell", "sbell"):
        return 4 * nnz + part.get("block_index_bytes", 0) + vec
    return 8 * nnz + 4 * (rows + 1) + vec


def part_min_bytes(part, rows, cols):
    """Bytes the part's own format cannot avoid (DIA stores no column
    indices, COO a row index per nonzero): for reading the roofline next to
    the measured traffic."""
    nnz = part["nnz"]
    fmt = part["format"]
    vec = 4 * cols + 8 * rows
    if fmt in ("dia", "bdia"):
        return 4 * nnz + vec
    if fmt in ("bcsr", "bell", "sbell"):
        return 4 * nnz + part.get("block_index_bytes", 0) + vec
    return 8 * nnz + 4 * (rows + 1) + vec


def cpu_baseline(A, x, y_gpu, budget_s=6.0):
    """The CPU path timed on this host over the whole matrix: the
    reference's serial csr_spmv (include/spmv_serial.h:199) compiled from
    its sources (oracle/_ref; the oracle's restatement of it where that is
    absent) on one core, and the same row loop over all host threads
    (OpenMP, rows split statically) -- each for as many calls as fit in
    budget_s.  The reported value is the faster (all-thread) leg."""
    import numpy as np
    import clspmv_mic_amd as cs
    import oracle
    csr = cs.coo2csr(A.copy())
    arrs = csr.arrays()
    y = np.zeros(A.height, np.float32)
    orc = oracle.Oracle()
    kind = "port"
    if oracle.reference_available():
        ref = oracle.Reference()
        serial = lambda: ref.spmv("csr", csr.s, x, y)  # noqa: E731
        kind = "reference"
    else:
        serial = lambda: orc.csr_spmv(arrs["rp"], arrs["col"], arrs["val"], x, y)  # noqa: E731
    threads = orc.max_threads()
    parallel = lambda: orc.csr_spmv_threads(threads, arrs["rp"], arrs["col"], arrs["val"], x, y)  # noqa: E731

    def timed(run):
        t0 = time.perf_counter()
        run()
        single = max(time.perf_counter() - t0, 1e-6)
        iters = max(2, min(500, int(budget_s / single)))
        t0 = time.perf_counter()
        for _ in range(iters):
            run()
        return (time.perf_counter() - t0) / iters, iters

    per1, it1 = timed(serial)
    pern, itn = timed(parallel)
    # GPU result against the FP64 oracle, scaled by |A||x| per row.
    y64, bound = orc.spmv_f64(A.row, A.col, A.val, x, A.height)
    err, _ = orc.scaled_error(y_gpu, y64, bound)
    gf = lambda per: round(2.0 * A.nnz / per / 1e9, 4)  # noqa: E731
    return {
        "value": gf(pern),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"whole matrix ({A.nnz} nnz): {itn} CSR SpMVs over {threads} threads "
                  f"({pern * 1e3:.2f} ms each)",
        "single_core": {"value": gf(per1), "unit": "GFLOP/s", "cores": 1, "kind": kind,
                        "sample": f"whole matrix: {it1} serial csr_spmv calls ({per1 * 1e3:.2f} ms each)"},
        "gpu_scaled_error_vs_fp64": err,
    }


def measured_traffic(workload, kernel_prefix):
    """HBM bytes per launch of a kernel from the counter table committed in
    profiles/traffic_counters.md (rows: | workload | kernel | FETCH bytes |
    WRITE bytes | traffic bytes | ...), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == workload and cells[1].startswith(kernel_prefix):
            try:
                return int(cells[4].replace(",", ""))
            except ValueError:
                return None
    return None


# Kernel-name prefix of each format's dominant kernel (rocprofv3 names).
KERNEL_OF = {"dia": "dia", "bdia": "bdia", "coo": "panel_pass", "csr": "panel_pass", "bcsr": "bcsr",
             "bell": "bell_kernel", "sbell": "sbell_kernel", "ell": "ell", "sell": "sell_kernel"}


def measure_cocktail(args, name):
    """Builds, partitions, tunes and times one Cocktail SpMV workload;
    returns its JSON object (without the run-level keys)."""
    import clspmv_mic_amd as cs
    from clspmv_mic_amd import matrices
    t0 = time.time()
    A, desc = build_matrix(name)
    print(f"[bench] matrix {name}: {A.height}x{A.width} nnz={A.nnz} ({time.time() - t0:.1f}s)",
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
    modes = rt.launch_modes()
    cold_ms = rt.time_cold(max(5, min(50, args.steps // 4)), 1 << 30)
    for p, (pname, s) in zip(parts, ck.parts()):
        if pname in ("bcsr",):
            p["block_index_bytes"] = 4 * s.b4csr_block_num + 4 * (s.b4csr_row_num + 1)
        elif pname == "bell":
            p["block_index_bytes"] = 4 * s.b4ell_row_num * s.b4ell_block_num
        elif pname == "sbell":
            p["block_index_bytes"] = 4 * (s.sbell_slice_ptr[s.sbell_slice_num] if s.sbell_slice_num else 0)
    for p in parts:
        p["bytes"] = part_bytes(p, A.height, A.width)
        p["format_bytes"] = part_min_bytes(p, A.height, A.width)
        p["gbps"] = p["format_bytes"] / (p["ms"] * 1e-3) / 1e9 if p["ms"] > 0 else 0.0
    dom = max(parts, key=lambda p: p["ms"])
    # The whole pass: every part's format bytes, x and y counted once.
    vec = 4 * A.width + 8 * A.height
    pass_format = sum(p["format_bytes"] - vec for p in parts) + vec
    pass_survey = 8 * A.nnz + 4 * (A.height + 1) + vec
    rt.close()

    cpu = None if args.no_cpu else cpu_baseline(A, x, y_one)
    frac = lambda b, ms: round(b / (ms * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4) if ms > 0 else 0.0  # noqa: E731
    return {
        "metric": "SpMV GFLOP/s & HBM GB/s (% roofline); served requests/s + p99 latency",
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "ms_per_step": round(ms_per_step, 6),
        "config": {"workload": desc, "rows": A.height, "cols": A.width, "nnz": A.nnz,
                   "partition": summary, "tables": os.path.relpath(tables, ROOT),
                   "launch": "graph" if graph else "direct", "overlap": modes["overlap"],
                   "dia_in_first_pass": modes["fuse"], "tuned": not args.no_tune},
        "roofline": {
            "bound": "hbm",
            "kernel": f"{dom['format']} method {dom['method']}",
            # Bytes the part's format must move per launch: true nonzeros'
            # values (+ a column index per nonzero for CSR/COO, + block
            # indices for block formats), x read once, y read and written.
            "achieved": round(dom["gbps"], 2),
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": frac(dom["format_bytes"], dom["ms"]),
            "traffic": measured_traffic(name, KERNEL_OF.get(dom["format"], dom["format"])),
            "algorithmic_bytes": dom["format_bytes"],
            "kernel_ms": round(dom["ms"], 6),
            # The count that credits 4 bytes of column index per nonzero to
            # every format (DIA and block formats read none, so this frac
            # can pass 1).
            "index_credit_bytes": dom["bytes"],
            "index_credit_frac": frac(dom["bytes"], dom["ms"]),
        },
        "pass_roofline": {"bytes": pass_format, "frac": frac(pass_format, ms_per_step),
                          "index_credit_bytes": pass_survey, "index_credit_frac": frac(pass_survey, ms_per_step)},
        "cold_cache": {"ms_per_step": round(cold_ms, 6),
                       "value": round(2.0 * A.nnz / (cold_ms * 1e-3) / 1e9, 3),
                       "pass_frac": frac(pass_format, cold_ms),
                       "note": "each pass timed alone after a 1 GiB read-only stream that evicts the L2s "
                               "and the 256 MB memory-side cache"},
        "parts": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in p.items()} for p in parts],
        "graph_ms": round(graph_ms, 6),
        "cpu_baseline": cpu,
    }


def spmm_bytes(blocks, brows, rows, cols, k, s_v=2, s_x=2, y_read=True):
    """Algorithmic bytes of Y = A X (y_read False) or Y += A X with 4x4
    blocks: dense block values, block column ids and block row pointers, X
    once, Y written (and read)."""
    return blocks * 16 * s_v + 4 * blocks + 4 * (brows + 1) + s_x * cols * k + (2 if y_read else 1) * 4 * rows * k


def run_spmm(args, collect=None):
    """Mixed-precision MFMA block SpMM on the 4x4-block matrix (524,288 block
    rows x 12 blocks): one JSON line per k (appended to `collect` instead of
    printed when given)."""
    import numpy as np
    import clspmv_mic_amd as cs
    from clspmv_mic_amd import matrices
    dtype = {"bf16": cs.DT_BF16, "fp16": cs.DT_F16}[args.dtype]
    t0 = time.time()
    A = matrices.block_stencil(524_288, 4, 4, seed=4)
    bc = cs.coo2bcsr(A.copy(), 4, 4)
    hdr = bc.header()
    blocks, brows = hdr["bcsr_block_num"], hdr["bcsr_row_num"]
    ctx = cs.Context(0)
    M = cs.BlockSpmm(ctx, bc, dtype)
    print(f"[bench] spmm matrix {A.height}x{A.width} nnz={A.nnz} blocks={blocks} ({time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    vals_s = cs.convert_to_f32(cs.convert_from_f32(A.val, dtype), dtype)
    # FP32 CSR references for the FP64 verification SpMV on the GPU: the
    # values before rounding and the values as stored.
    ref_in = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=0)
    ref_st = cs.DeviceMatrix(ctx, cs.coo2csr(cs.Coo(A.row, A.col, vals_s, A.height, A.width)), panel=0)
    matrix_error = M.error_report(ref_in, seed=1)
    for k in args.k:
        ldx, ldy = A.width, A.height
        rng = np.random.default_rng(5)
        X = rng.uniform(-1, 1, size=(k, ldx)).astype(np.float32)
        Xb = cs.convert_from_f32(X, dtype)
        Xs = cs.convert_to_f32(Xb, dtype)
        dX = cs.DeviceBuffer(ctx, Xb)
        dY = cs.DeviceVector(ctx, k * ldy, np.zeros(k * ldy, np.float32))
        tiled = args.x_layout == "tiled"
        pack_ms = None
        if tiled:  # X handed over in 4-row tiles: packed once, outside the timed region
            dX4 = cs.spmm_pack_x(ctx, k, A.width, dX, ldx)
            # what packing a column-major X costs (host-timed, 20 calls)
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(20):
                cs.check(cs.lib.clspmv_spmm_pack_x(ctx.h, k, A.width, dX.p, ldx, dX4.p), "spmm_pack_x")
            ctx.sync()
            pack_ms = (time.perf_counter() - t0) * 1e3 / 20

        def product(accumulate=False):
            if tiled:
                M.spmm_tiled(k, dX4, dY, ldy, accumulate=accumulate)
            else:
                M.spmm(k, dX, ldx, dY, ldy)  # Y += A X only

        overwrite = tiled  # Y = A X, as the workload defines it
        product()
        ctx.sync()
        # Error per right-hand side against the FP64 SpMV on the GPU: of the
        # stored (rounded) values and x -- the kernel's own error -- and of
        # the FP32 inputs before rounding -- what 16-bit storage costs.
        errs, qerrs = [], []
        for j in range(min(k, args.err_columns)):
            yj = dY.view(j * ldy)
            errs.append(cs.verify_spmv(ctx, ref_st, cs.DeviceVector(ctx, ldx, Xs[j]), yj)["max_scaled_error"])
            qerrs.append(cs.verify_spmv(ctx, ref_in, cs.DeviceVector(ctx, ldx, X[j]), yj)["max_scaled_error"])
        for _ in range(args.warmup):
            product()
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            product()
        ctx.sync()
        ms = (time.perf_counter() - t1) * 1e3 / args.steps
        # Y += A X (Y read as well), timed the same way.
        acc_ms = None
        if tiled and not args.no_accumulate:
            product(accumulate=True)
            ctx.sync()
            t2 = time.perf_counter()
            for _ in range(args.steps):
                product(accumulate=True)
            ctx.sync()
            acc_ms = (time.perf_counter() - t2) * 1e3 / args.steps
        # The other X layout, timed the same way, for comparison.
        other_ms = None
        if tiled:
            M.spmm(k, dX, ldx, dY, ldy)
            ctx.sync()
            t2 = time.perf_counter()
            for _ in range(args.steps):
                M.spmm(k, dX, ldx, dY, ldy)
            ctx.sync()
            other_ms = (time.perf_counter() - t2) * 1e3 / args.steps
        flops = 2.0 * A.nnz * k
        nbytes = spmm_bytes(blocks, brows, A.height, A.width, k, y_read=not overwrite)
        acc_bytes = spmm_bytes(blocks, brows, A.height, A.width, k)
        out = {
            "metric": "block SpMM GFLOP/s (BF16/FP16 values, FP32 accumulation on MFMA)",
            "value": round(flops / (ms * 1e-3) / 1e9, 3), "unit": "GFLOP/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 6),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
            "data": "synthetic",
            "config": {"workload": "4x4-block matrix, 524,288 block rows x 12 blocks, X uniform[-1,1]",
                       "rows": A.height, "nnz": A.nnz, "blocks": blocks, "k": k,
                       "x_layout": "4-row tiles" if tiled else "column-major",
                       "product": "Y = A X" if overwrite else "Y += A X"},
            "roofline": {"bound": "hbm", "achieved": round(nbytes / (ms * 1e-3) / 1e9, 2), "peak": PEAK_HBM_GBPS,
                         "unit": "GB/s", "frac": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4),
                         # counters of the BF16 tiled kernel at this k, where measured
                         "traffic": measured_traffic(f"spmm-set-k{k}" if overwrite else f"spmm-k{k}",
                                                     "bspmm4")
                         if tiled and args.dtype == "bf16" else None,
                         "algorithmic_bytes": nbytes},
            "y_accumulate": None if acc_ms is None else {
                "ms": round(acc_ms, 6), "algorithmic_bytes": acc_bytes,
                "frac": round(acc_bytes / (acc_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4),
                "traffic": measured_traffic(f"spmm-k{k}", "bspmm4") if args.dtype == "bf16" else None},
            "mfma_tflops": round(flops / (ms * 1e-3) / 1e12, 3),
            "column_major_x_ms": round(other_ms, 6) if other_ms is not None else None,
            "pack_x_ms": round(pack_ms, 6) if pack_ms is not None else None,
            "error_vs_fp64_stored_values": [float(f"{e:.3e}") for e in errs],
            "error_vs_fp64_fp32_inputs": [float(f"{e:.3e}") for e in qerrs],
            "matrix_error_vs_fp64": {kk: (float(f"{v:.3e}") if isinstance(v, float) else v)
                                     for kk, v in matrix_error.items()},
            "cpu_baseline": None,
        }
        if collect is None:
            print(json.dumps(out), flush=True)
        else:
            collect.append(out)
    return 0


def run_serve(args):
    """Serving loop over 64 resident matrices (32 stencils, 32 power-law,
    10k..2M rows log-uniform, FP32 and BF16 alternating) under a Poisson
    stream of requests with 1..16 right-hand sides."""
    import numpy as np
    import clspmv_mic_amd as cs
    from clspmv_mic_amd import matrices
    nmat = args.matrices
    ctx = cs.Context(0)
    server = cs.Server(streams=args.streams, max_batch=64, hbm_budget=16 << 30)
    dev, ids, kinds, errs16 = [], [], [], []
    t0 = time.time()
    rng, mix = matrices.serving_mix(nmat, seed=6)
    for kind, bf16, A in mix:
        dt = cs.DT_BF16 if bf16 else cs.DT_F32
        if kind == "stencil":
            d = cs.DeviceMatrix(ctx, cs.coo2dia(A.copy()), dtype=dt)
        else:
            # CSR with both encodings: the server runs a batch as one
            # multi-vector pass per 16 columns (own arrays) or as x-panel
            # SpMVs per column, whichever measured faster at registration.
            csr = cs.coo2csr(A.copy())
            d = cs.DeviceMatrix(ctx, csr, panel=cs.PANEL_BOTH, dtype=dt)
        if bf16:  # per-matrix error of the 16-bit storage against FP64 on the GPU
            ref = cs.DeviceMatrix(ctx, csr if kind == "powerlaw" else cs.coo2csr(A.copy()), panel=0)
            errs16.append(d.error_report(ref, seed=100 + len(dev))["max_scaled_error"])
            del ref
        dev.append(d)
        ids.append(server.add(d, 0, kmax=16))
        kinds.append((kind, "bf16" if bf16 else "f32", A.height, A.nnz))
    print(f"[bench] serve: {nmat} matrices resident ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    n = args.requests
    mat = rng.integers(0, nmat, size=n)
    ks = rng.integers(1, 17, size=n)
    nnz_of = np.array([kk[3] for kk in kinds])

    def trace(rate):
        t = np.cumsum(rng.exponential(1.0 / rate, size=n)) if rate else np.zeros(n)
        return [(ids[mat[i]], int(ks[i]), 10_000 + i, float(t[i]), None) for i in range(n)]

    # Capacity: every request waiting at t = 0.
    lat0, wall0 = server.replay(trace(None), mode=0)
    cap = n / wall0
    # Plain per-request launches (one at a time, each synchronised), on a
    # sample of the same requests.
    nb = min(n, args.baseline_requests)
    latb, wallb = server.replay(trace(None)[:nb], mode=1)
    plain = nb / wallb
    # Poisson arrivals at a fraction of capacity.
    rate = args.load * cap
    lat, wall = server.replay(trace(rate), mode=0)
    # The highest arrival rate whose p99 latency stays within the SLO:
    # bisection over the load (fraction of capacity) on shorter traces.
    ns = min(n, args.slo_requests)

    def slo_load(mode, full):
        lo, hi = 0.0, 1.0
        for _ in range(args.slo_steps):
            mid = 0.5 * (lo + hi)
            l, _ = server.replay(trace(mid * full)[:ns], mode=mode)
            if float(np.percentile(l, 99)) * 1e3 <= args.slo_ms:
                lo = mid
            else:
                hi = mid
        return lo

    slo_load_served = slo_load(0, cap)
    slo_rate = slo_load_served * cap
    # The same for the plain launches, whose requests wait for each other.
    slo_rate_plain = slo_load(1, plain) * plain
    stats = server.stats()
    server.close()
    bound = serve_kernel_time(cs, np, ctx, dev, kinds, mat, ks, wall0 / n, wallb / nb, nb)
    served = n / wall
    flops = 2.0 * float(np.sum(nnz_of[mat] * ks))
    out = {
        "metric": "served requests/s + p99 latency",
        "value": round(served, 1), "unit": "requests/s", "n_gpus": 1, "steps": n, "warmup": 0,
        "ms_per_step": round(wall * 1e3 / n, 6), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic",
        "config": {"workload": f"{nmat} resident matrices (stencils + power-law, 10k..2M rows, FP32/BF16), "
                               f"{n} Poisson requests, k in 1..16",
                   "arrival_rate": round(rate, 1), "streams": args.streams},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                       "max": round(float(lat.max()) * 1e3, 3)},
        "capacity_requests_per_s": round(cap, 1),
        "plain_per_request_requests_per_s": round(plain, 1),
        "speedup_vs_plain": round(cap / plain, 2),
        "slo": {"p99_ms": args.slo_ms, "requests_per_s": round(slo_rate, 1),
                "load": round(slo_load_served, 4),
                "plain_requests_per_s": round(slo_rate_plain, 1),
                "speedup_vs_plain": round(slo_rate / slo_rate_plain, 2) if slo_rate_plain > 0 else None,
                "requests_per_probe": ns},
        "plain_latency_ms_p50": round(float(np.percentile(latb, 50)) * 1e3, 3),
        "server_stats": stats,
        "kernel_time": bound,
        "gflops_served": round(flops / wall / 1e9, 2),
        "bf16_matrix_error_vs_fp64": {"matrices": len(errs16),
                                      "max": float(f"{max(errs16):.3e}") if errs16 else None,
                                      "median": float(f"{float(np.median(errs16)):.3e}") if errs16 else None},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def serve_kernel_time(cs, np, ctx, dev, kinds, mat, ks, server_wall, plain_wall, nb, kmax=16):
    """Where a request's time goes: per matrix, the HIP-event time of one
    single-vector SpMV (its best method) and of one multi-vector pass for
    each k = 1..kmax, summed over the requests of the trace.

    - plain: k single SpMVs per request (what the plain launches run);
    - best: the cheaper of the two per request (what the server runs);
    - coalesced: every matrix's columns of the whole trace in full
      kmax-column passes at the best cost, i.e. the kernel time of a server
      that batched perfectly.
    Kernels run one after another here; the server's streams can overlap
    the passes of small matrices, which none of these sums credit."""
    nmat = len(dev)
    rows = max(k[2] for k in kinds)
    cols = max(d.info.cols for d in dev)
    dX = cs.DeviceVector(ctx, cols * kmax, np.full(cols * kmax, 0.5, np.float32))
    dY = cs.DeviceVector(ctx, rows * kmax, np.zeros(rows * kmax, np.float32))
    t1 = np.zeros(nmat)
    cost = np.zeros((nmat, kmax + 1))
    multi_wins = 0
    for m, d in enumerate(dev):
        t1[m] = d.time(dX, dY, warmup=2, ntimes=10)
        for k in range(1, kmax + 1):
            tm = d.time_multi(k, dX, dY, d.info.rows, warmup=1, ntimes=5)
            cost[m, k] = min(k * t1[m], tm)
            multi_wins += tm < k * t1[m]
    plain = float(np.mean(ks[:nb] * t1[mat[:nb]]))
    best = float(np.mean(cost[mat, ks]))
    total = np.bincount(mat, weights=ks, minlength=nmat).astype(np.int64)
    coal = float(np.sum(total // kmax * cost[:, kmax] + cost[np.arange(nmat), total % kmax]) / len(mat))
    stencil = np.array([k[0] == "stencil" for k in kinds])
    per_kind = {}
    for name, sel in (("stencil", stencil), ("powerlaw", ~stencil)):
        req = sel[mat]
        per_kind[name] = {"requests": int(req.sum()),
                          "plain_kernel_ms": round(float(np.sum(ks[req] * t1[mat[req]])) / len(mat), 5),
                          "best_kernel_ms": round(float(np.sum(cost[mat[req], ks[req]])) / len(mat), 5)}
    return {"plain_wall_ms_per_request": round(plain_wall * 1e3, 5),
            "plain_kernel_ms_per_request": round(plain, 5),
            "server_wall_ms_per_request": round(server_wall * 1e3, 5),
            "best_kernel_ms_per_request": round(best, 5),
            "coalesced_kernel_ms_per_request": round(coal, 5),
            "capacity_bound_vs_plain": round(plain_wall * 1e3 / coal, 3) if coal > 0 else None,
            # A request's own kernel time at the best cost: the latency it
            # has with an idle GPU, whatever the scheduler.
            "request_kernel_ms": {"p50": round(float(np.percentile(cost[mat, ks], 50)), 5),
                                  "p99": round(float(np.percentile(cost[mat, ks], 99)), 5),
                                  "max": round(float(np.max(cost[mat, ks])), 5)},
            "multi_pass_wins": f"{multi_wins}/{nmat * kmax}",
            "by_kind_ms_per_request": per_kind}


def _driver_quiet(cs, name, coo, ntimes):
    """run_driver with the driver's own progress lines (C stdout) sent to
    stderr, so that stdout carries only the JSON lines."""
    import ctypes
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        rep, _ = cs.run_driver(name, coo, ntimes=ntimes, want_y=False)
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
    return rep


def run_sweep(args):
    """The per-format drivers (spmv_<fmt>, include/spmv_<fmt>.h:7; every
    format by default) on one matrix (the power-law 1M + band matrix by
    default): one JSON line per format with the best variant's GFLOP/s, or
    the reason the format was refused."""
    import numpy as np
    import clspmv_mic_amd as cs
    A, workload = build_matrix(args.matrix)
    longest = int(np.bincount(A.row, minlength=A.height).max())
    ell_entries = longest * ((A.height + 31) // 32 * 32)
    for name in args.formats:
        out = {"metric": "SpMV GFLOP/s per format (best variant of the format's driver)",
               "value": None, "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": 3,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic",
               "config": {"workload": workload, "nnz": A.nnz, "format": name}}
        t0 = time.time()
        if name == "ell" and ell_entries > (1 << 28):
            out["config"]["status"] = (f"refused: ELL pads every row to the longest ({longest}), "
                                       f"{ell_entries} entries")
        else:
            try:
                rep = _driver_quiet(cs, name, A.copy(), args.steps)
                out["value"] = round(rep.best_gflops, 3)
                out["ms_per_step"] = round(rep.best_ms, 6)
                out["config"].update({"status": "ok", "variants": rep.variants, "best_method": rep.best_method,
                                      "best_block": [rep.best_bheight, rep.best_bwidth]})
            except cs.ClspmvError as e:
                out["config"]["status"] = f"refused: {e}"
        out["config"]["seconds"] = round(time.time() - t0, 1)
        print(json.dumps(out), flush=True)
    return 0


def run_format(args):
    """One storage format on its own (clspmv_upload_<fmt>, then clspmv_spmv
    per method, HIP-event timed) on one matrix: one JSON line per method with
    GB/s on the bytes the format must move and the row-scaled error of its y
    against the FP64 SpMV on the GPU."""
    import numpy as np
    import clspmv_mic_amd as cs
    from clspmv_mic_amd import matrices
    A, workload = build_matrix(args.matrix)
    fmt, (bh, bw) = args.format, args.block
    dt = {"f32": None, "bf16": cs.DT_BF16, "fp16": cs.DT_F16}[args.values]
    if dt is not None and fmt not in ("dia", "csr"):
        raise SystemExit("16-bit values are stored for the DIA and CSR formats")
    owned = {"csr": lambda: cs.coo2csr(A.copy()), "ell": lambda: cs.coo2ell(A.copy()),
             "sell": lambda: cs.coo2sell(A.copy()), "dia": lambda: cs.coo2dia(A.copy()),
             "bdia": lambda: cs.coo2bdia(A.copy()),
             # Block layouts without the reference's per-buffer cap (MAX_MEM_OBJ, sized
             # for its accelerators): the 4x4 matrix's blocks take 403 MB.
             "bcsr": lambda: cs.coo2b4csr(A.copy(), bw, bh, cap=False),
             "bell": lambda: cs.coo2b4ell(A.copy(), bw, bh, cap=False),
             "sbell": lambda: cs.coo2sbell(A.copy(), bw, bh, cap=False)}[fmt]()
    if owned is None:
        raise SystemExit(f"{fmt} refused for this matrix")
    h = owned.header()
    part = {"format": fmt, "nnz": A.nnz}
    if fmt == "bcsr":
        part["block_index_bytes"] = 4 * h["b4csr_block_num"] + 4 * (h["b4csr_row_num"] + 1)
    elif fmt == "bell":
        part["block_index_bytes"] = 4 * h["b4ell_row_num"] * h["b4ell_block_num"]
    elif fmt == "sbell":
        part["block_index_bytes"] = 4 * int(owned.arrays()["ptr"][-1])
    nbytes = part_min_bytes(part, A.height, A.width)
    ctx = cs.Context(0)
    x = cs.DeviceVector(ctx, A.width, matrices.uniform_vector(A.width, seed=3))
    ref = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=0)
    if dt is None:
        dA = cs.DeviceMatrix(ctx, owned)
        ref_st = ref
    else:
        # 16-bit values (FP32 sums): 2 bytes per stored value instead of 4;
        # the error is reported against the values as stored (the kernel's
        # own) and against the FP32 inputs (what the storage costs).
        dA = cs.DeviceMatrix(ctx, owned, panel=0 if fmt == "csr" else None, dtype=dt)
        nbytes -= 2 * A.nnz
        As = cs.Coo(A.row, A.col, cs.convert_to_f32(cs.convert_from_f32(A.val, dt), dt), A.height, A.width)
        ref_st = cs.DeviceMatrix(ctx, cs.coo2csr(As), panel=0)
    for method in range(dA.methods()):
        y = cs.DeviceVector(ctx, A.height, np.zeros(A.height, np.float32))
        try:
            dA.spmv(x, y, method)
        except cs.ClspmvError as e:  # a method this matrix's shape rules out
            print(f"[bench] {fmt} method {method} refused: {e}", file=sys.stderr, flush=True)
            continue
        err = cs.verify_spmv(ctx, ref_st, x, y)["max_scaled_error"]
        err_in = cs.verify_spmv(ctx, ref, x, y)["max_scaled_error"] if dt is not None else None
        ms = dA.time(x, y, method, warmup=args.warmup, ntimes=args.steps)
        gbps = nbytes / (ms * 1e-3) / 1e9
        print(json.dumps({
            "metric": "SpMV GB/s of one storage format (% roofline)", "value": round(gbps, 2), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 6),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32" if dt is None else f"{args.values} values, f32 sums", "data": "synthetic",
            "config": {"workload": workload, "format": fmt, "block": [bh, bw], "method": method, "nnz": A.nnz},
            "roofline": {"bound": "hbm", "achieved": round(gbps, 2), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                         "frac": round(gbps / PEAK_HBM_GBPS, 4), "traffic": None, "algorithmic_bytes": nbytes},
            "gflops": round(2.0 * A.nnz / (ms * 1e-3) / 1e9, 2), "error_vs_fp64": err,
            "error_vs_fp64_fp32_inputs": err_in}), flush=True)
    return 0


def run_probe(args):
    """Bandwidth ceilings of this GPU: the fastest copy and read-only
    streams at footprints below and above the 256 MB memory-side cache, and
    random gathers from x vectors of growing size."""
    import clspmv_mic_amd as cs
    ctx = cs.Context(0)
    copy = {f"{mb}MiB": round(ctx.bandwidth_probe(mb << 20, 20)[1], 1) for mb in (64, 256, 1024, 4096)}
    read = {f"{mb}MiB": round(ctx.read_probe(mb << 20, 20)[1], 1) for mb in (64, 256, 1024, 4096)}
    gather = {f"nx{nx}": round(ctx.gather_probe(nx, 1 << 25, 10)[1], 1)
              for nx in (1 << 16, 1 << 18, 1 << 20, 1 << 22, 1 << 24)}
    out = {"metric": "HBM bandwidth ceilings (copy counts read + write)", "value": max(read.values()),
           "unit": "GB/s", "n_gpus": 1, "higher_is_better": True, "data": "synthetic",
           "config": {"workload": "probes: float4 copy, float4 read, random gathers"},
           "copy_gbps": copy, "read_gbps": read, "gather_g_per_s": gather, "device": ctx.info()[0]}
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "powerlaw", "stencil27", "lap2d", "block4x4", "spmm", "serve", "sweep",
                             "probe", "format"],
                    help="headline: the 27-point stencil 128^3 with the 1M power-law matrix in the same line")
    ap.add_argument("--dtype", default="bf16", choices=["bf16", "fp16"], help="spmm: value type")
    ap.add_argument("--k", type=int, nargs="+", default=[8, 16, 32], help="spmm: right-hand sides")
    ap.add_argument("--err-columns", type=int, default=4, help="spmm: columns checked against FP64")
    ap.add_argument("--no-accumulate", action="store_true",
                    help="spmm: time only Y = A X, not also Y += A X")
    ap.add_argument("--x-layout", default="tiled", choices=["tiled", "column"],
                    help="spmm: X in 4-row tiles (clspmv_spmm_tiled) or column-major (clspmv_spmm)")
    ap.add_argument("--matrices", type=int, default=64, help="serve: resident matrices")
    ap.add_argument("--requests", type=int, default=20000, help="serve: requests")
    ap.add_argument("--baseline-requests", type=int, default=2000, help="serve: plain-launch sample")
    ap.add_argument("--load", type=float, default=0.7, help="serve: Poisson rate / capacity")
    ap.add_argument("--streams", type=int, default=8, help="serve: HIP streams")
    ap.add_argument("--slo-ms", type=float, default=2.0, help="serve: p99 latency target")
    ap.add_argument("--slo-requests", type=int, default=4000, help="serve: requests per SLO probe")
    ap.add_argument("--slo-steps", type=int, default=6, help="serve: bisection steps for the SLO rate")
    ap.add_argument("--matrix", default="powerlaw", choices=["powerlaw", "stencil27", "lap2d", "block4x4"],
                    help="sweep, format: the matrix")
    ap.add_argument("--formats", nargs="+", default=SWEEP_FORMATS, choices=SWEEP_FORMATS,
                    help="sweep: the drivers to run")
    ap.add_argument("--values", default="f32", choices=["f32", "bf16", "fp16"],
                    help="format: stored value type (16-bit for dia and csr)")
    ap.add_argument("--format", default="bell", choices=["csr", "ell", "sell", "dia", "bdia", "bcsr", "bell", "sbell"],
                    help="format: the storage format")
    ap.add_argument("--block", type=int, nargs=2, default=[4, 4], metavar=("BH", "BW"),
                    help="format: block height and width")
    ap.add_argument("--tables", default=None, help="directory holding benchmark/*.ben")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-spmm", action="store_true", help="headline: skip the block SpMM summary")
    ap.add_argument("--no-tune", action="store_true", help="keep the tables' methods and graph launches")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the engine runs one device "
              "per process and has no multi-GPU path to measure.", file=sys.stderr)
        return 2

    if args.workload == "spmm":
        return run_spmm(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "sweep":
        return run_sweep(args)
    if args.workload == "probe":
        return run_probe(args)
    if args.workload == "format":
        return run_format(args)

    names = ["stencil27", "powerlaw"] if args.workload == "headline" else [args.workload]
    results = [measure_cocktail(args, n) for n in names]
    out = {
        "metric": results[0]["metric"],
        "value": results[0]["value"],
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results[0]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
    }
    for k, v in results[0].items():
        if k not in out:
            out[k] = v
    if len(results) > 1:  # the power-law matrix's numbers, measured the same way
        extra = {k: v for k, v in results[1].items() if k != "metric"}
        out["powerlaw"] = extra
    if args.workload == "headline" and not args.no_spmm:
        # The BF16 block SpMM (Y = A X, X in 4-row tiles) at k = 8 / 16 / 32, in
        # brief (bench.py --workload spmm prints the full lines).
        sp_args = argparse.Namespace(**{**vars(args), "k": [8, 16, 32], "err_columns": 1, "dtype": "bf16",
                                        "no_accumulate": False, "x_layout": "tiled"})
        lines = []
        run_spmm(sp_args, collect=lines)
        out["spmm"] = [{"k": d["config"]["k"], "ms_y_eq_ax": d["ms_per_step"], "tflops": d["mfma_tflops"],
                        "frac": d["roofline"]["frac"], "traffic": d["roofline"]["traffic"],
                        "ms_y_plus_ax": (d["y_accumulate"] or {}).get("ms"),
                        "frac_y_plus_ax": (d["y_accumulate"] or {}).get("frac"),
                        "max_error_vs_fp64": max(d["error_vs_fp64_stored_values"])} for d in lines]
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
