"""The serving workload at its full shape: 64 resident matrices (stencils
and power-law matrices of 10k to 2M rows, FP32 and BF16 values, as the
serving benchmark builds them) and a few hundred requests of 1 to 16
right-hand sides.  Every request gives the same y bit for bit alone (one at
a time) and in the mix (all queued at once, coalesced per matrix, several
streams), and a sample matches the FP64 oracle within 1e-5 of |A||x| (on
the values as stored for BF16 matrices)."""

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices

pytestmark = pytest.mark.gpu

TOL = 1e-5
NREQ = 256


def test_serving_mix_alone_equals_mix(ctx, oracle):
    rng, mix = matrices.serving_mix(64, seed=6)
    server = cs.Server(streams=8, max_batch=64, hbm_budget=16 << 30, host_staging=2 << 30)
    dev, ids = [], []
    for kind, bf16, A in mix:
        dt = cs.DT_BF16 if bf16 else cs.DT_F32
        if kind == "stencil":
            d = cs.DeviceMatrix(ctx, cs.coo2dia(A.copy()), dtype=dt)
        else:
            d = cs.DeviceMatrix(ctx, cs.coo2csr(A.copy()), panel=cs.PANEL_BOTH, dtype=dt)
        dev.append(d)
        ids.append(server.add(d, 0, kmax=16))
    print(f"serving mix: {len(mix)} matrices, {sum(A.nnz for _, _, A in mix)} nonzeros resident", flush=True)
    req = rng.integers(0, len(mix), size=NREQ)
    ks = rng.integers(1, 17, size=NREQ)
    alone = [np.zeros(mix[m][2].height * k, np.float32) for m, k in zip(req, ks)]
    mixed = [np.zeros_like(y) for y in alone]
    for i, (m, k) in enumerate(zip(req, ks)):
        server.wait(server.submit(ids[m], int(k), 7000 + i, alone[i]))
    before = server.stats()
    lat, wall = server.replay([(ids[m], int(k), 7000 + i, 0.0, mixed[i]) for i, (m, k) in enumerate(zip(req, ks))])
    after = server.stats()
    server.close()
    print(f"mix: {NREQ} requests in {wall * 1e3:.1f} ms, {after['passes'] - before['passes']} passes", flush=True)
    assert after["passes"] - before["passes"] < NREQ
    for i in range(NREQ):
        np.testing.assert_array_equal(alone[i], mixed[i], err_msg=f"request {i} (matrix {req[i]}, k={ks[i]})")
    for i in range(0, NREQ, NREQ // 12):
        kind, bf16, A = mix[req[i]]
        vals = cs.convert_to_f32(cs.convert_from_f32(A.val, cs.DT_BF16), cs.DT_BF16) if bf16 else A.val
        for j in sorted({0, int(ks[i]) - 1}):
            x = cs.request_x(7000 + i, j, A.width)
            y64, bound = oracle.spmv_f64(A.row, A.col, vals, x, A.height)
            y = mixed[i][j * A.height:(j + 1) * A.height]
            err, worst = oracle.scaled_error(y.copy(), y64, bound)
            assert err <= TOL, f"request {i} rhs {j} ({kind}, bf16={bf16}): {err:.3e} at row {worst}"
