"""GPU tests of the serving loop (clspmv_serve.h): every request's result is
the same alone (submitted and finished one at a time) and in the mix
(requests of one matrix coalesced into one pass, passes of different
matrices on several streams, graphs replayed), bit for bit; results and the
plain per-request baseline match the FP64 oracle within 1e-5 of |A||x|;
caller-supplied x (host or device) gives the same y as the same values
generated from the seed; a pool too small for every matrix's buffers
evicts and still serves everything identically; one device matrix
registered twice (in one server and in two) is never run by two passes at
once."""

import numpy as np
import pytest

import clspmv_mic_amd as cs
from clspmv_mic_amd import matrices

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def resident(ctx):
    """(coo, device matrix, method, dtype) of matrices of different formats
    and value types."""
    out = []
    lap = matrices.laplacian_2d(64)
    out.append((lap, cs.DeviceMatrix(ctx, cs.coo2dia(lap.copy())), 0, None))
    pl = matrices.power_law(20000, seed=2, max_len=400)
    out.append((pl, cs.DeviceMatrix(ctx, pl.sorted(), panel=cs.PANEL_AUTO), 0, None))
    rnd = matrices.random_sparse(3000, 2500, 0.004, seed=11, long_rows=3, long_len=900)
    out.append((rnd, cs.DeviceMatrix(ctx, cs.coo2csr(rnd.copy())), 1, None))
    lap16 = matrices.laplacian_2d(40)
    out.append((lap16, cs.DeviceMatrix(ctx, cs.coo2dia(lap16.copy()), dtype=cs.DT_BF16), 0, cs.DT_BF16))
    pl16 = matrices.power_law(9000, seed=7, max_len=300, band=())
    out.append((pl16, cs.DeviceMatrix(ctx, cs.coo2csr(pl16.copy()), panel=0, dtype=cs.DT_BF16), 0, cs.DT_BF16))
    blk = matrices.block_stencil(500, 4, 4, seed=4, offsets=(0, 1, -1, 9))
    out.append((blk, cs.BlockSpmm(ctx, cs.coo2bcsr(blk.copy(), 4, 4), cs.DT_BF16), 0, cs.DT_BF16))
    assert out[1][1].info.panel_active == 1
    return out


def _server(resident, **kw):
    s = cs.Server(**kw)
    ids = [s.add(d, m, kmax=8) for (_, d, m, _) in resident]
    return s, ids


def _requests(resident, ids, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        m = int(rng.integers(len(ids)))
        k = int(rng.integers(1, 5))
        A = resident[m][0]
        reqs.append((ids[m], k, 1000 + i, 0.0, np.zeros(A.height * k, np.float32)))
    return reqs


def _oracle_check(oracle, resident, m, k, seed, y):
    A, dm, _, dtype = resident[m]
    vals = A.val if dtype is None else cs.convert_to_f32(cs.convert_from_f32(A.val, dtype), dtype)
    block = isinstance(dm, cs.BlockSpmm)  # x itself is rounded to 16 bits there
    for j in range(k):
        x = cs.request_x(seed, j, A.width)
        if block:
            x = cs.convert_to_f32(cs.convert_from_f32(x, dtype), dtype)
        y64, bound = oracle.spmv_f64(A.row, A.col, vals, x, A.height)
        err, worst = oracle.scaled_error(y[j * A.height:(j + 1) * A.height].copy(), y64, bound)
        assert err <= TOL, f"matrix {m} rhs {j}: {err:.3e} at row {worst}"


def test_alone_equals_mix_and_oracle(ctx, oracle, resident):
    """Each request alone on the server (submitted and waited for one at a
    time) and in the mix (all at once: requests of a matrix coalesced into
    one pass, several streams) gives the same result bit for bit; the plain
    per-request baseline (one SpMV launch per right-hand side) agrees with
    both within the tolerance."""
    s, ids = _server(resident, streams=3, max_batch=8)
    mixed = _requests(resident, ids, 48, seed=1)
    alone = [(m, k, sd, t, np.zeros_like(y)) for (m, k, sd, t, y) in mixed]
    plain = [(m, k, sd, t, np.zeros_like(y)) for (m, k, sd, t, y) in mixed]
    for (m, k, sd, _, y) in alone:
        s.wait(s.submit(m, k, sd, y))
    before = s.stats()
    lat0, _ = s.replay(mixed, mode=0)
    after = s.stats()
    lat1, _ = s.replay(plain, mode=1)
    s.close()
    assert before["passes"] == len(alone)
    assert after["requests"] - before["requests"] == len(mixed)
    # All 48 arrive at once: each matrix's requests share passes.
    assert after["passes"] - before["passes"] < len(mixed), after
    assert np.all(lat0 >= 0) and np.all(lat1 >= 0)
    for (m, k, sd, _, ya), (_, _, _, _, ym) in zip(alone, mixed):
        np.testing.assert_array_equal(ya, ym, err_msg=f"matrix {m} seed {sd}")
    for (m, k, sd, _, y), (_, _, _, _, yp) in zip(mixed, plain):
        _oracle_check(oracle, resident, m, k, sd, y)
        _oracle_check(oracle, resident, m, k, sd, yp)


def _slot_bound(resident, kmax):
    """Bytes that hold any one matrix's working buffers for kmax columns."""
    return max(4 * kmax * (2 * ((A.width + 3) // 4 * 4) + (A.height + 3) // 4 * 4) + 3 * 256
               for (A, _, _, _) in resident)


@pytest.mark.parametrize("graphs", [False, True])
def test_graphs_and_budget(ctx, resident, graphs):
    """Plain launches and graph replays give the same results; a pool that
    holds one matrix's buffers at a time evicts and re-admits them and still
    serves every request identically."""
    ref, ids = _server(resident, streams=2, max_batch=4)
    reqs = _requests(resident, ids, 24, seed=2)
    ref.replay(reqs, mode=0)
    ref.close()
    s, _ = _server(resident, streams=4, max_batch=16, hbm_budget=_slot_bound(resident, 8), use_graphs=graphs)
    got = [(m, k, sd, 1e-4 * i, np.zeros_like(y)) for i, (m, k, sd, _, y) in enumerate(reqs)]
    s.replay(got, mode=0)
    st = s.stats()
    s.close()
    assert st["evictions"] > 0 and st["pool_used"] <= st["pool_bytes"], st
    for (_, _, _, _, a), (_, _, _, _, b) in zip(reqs, got):
        np.testing.assert_array_equal(a, b)


def test_caller_x_host_and_device(ctx, resident):
    """x handed over by the caller, from host memory or from a device
    buffer, gives the same y bit for bit as the same values generated from
    the seed; y can also land in a device buffer."""
    s, ids = _server(resident, streams=2, max_batch=8)
    for i, (A, _, _, _) in enumerate(resident):
        k = 3
        seed = 500 + i
        x = np.concatenate([cs.request_x(seed, j, A.width) for j in range(k)]).astype(np.float32)
        y_seed = np.zeros(A.height * k, np.float32)
        y_host = np.zeros_like(y_seed)
        y_dev = np.zeros_like(y_seed)
        dx = cs.DeviceVector(ctx, x.size, x)
        dy = cs.DeviceVector(ctx, y_seed.size, np.full(y_seed.size, 7.0, np.float32))
        t = [s.submit(ids[i], k, seed, y_seed), s.submit(ids[i], k, 1, y_host, x=x),
             s.submit(ids[i], k, 2, y_dev, x=dx, device_y=dy)]
        for tk in t:
            s.wait(tk)
        np.testing.assert_array_equal(y_host, y_seed, err_msg=f"matrix {i}: host x")
        np.testing.assert_array_equal(y_dev, y_seed, err_msg=f"matrix {i}: device x")
        np.testing.assert_array_equal(dy.download(), y_seed, err_msg=f"matrix {i}: device y")
    s.close()


def test_matrix_registered_twice(ctx, resident):
    """One device matrix registered twice in one server and once more in a
    second server: interleaved requests from both servers never run two
    passes on it at once (its x-panel scratch is shared), so every y equals
    the request alone."""
    A, dm, meth, _ = resident[1]  # the x-panel power-law matrix: pass 1 writes the matrix's scratch
    s1 = cs.Server(streams=3, max_batch=4)
    s2 = cs.Server(streams=3, max_batch=4)
    a = s1.add(dm, meth, kmax=8)
    b = s1.add(dm, meth, kmax=8)
    c = s2.add(dm, meth, kmax=8)
    rng = np.random.default_rng(9)
    plan = [((s1, a), (s1, b), (s2, c))[int(rng.integers(3))] for _ in range(36)]
    ks = [int(rng.integers(1, 5)) for _ in plan]
    alone = []
    for i, ((srv, mid), k) in enumerate(zip(plan, ks)):
        y = np.zeros(A.height * k, np.float32)
        srv.wait(srv.submit(mid, k, 900 + i, y))
        alone.append(y)
    mixed = [np.zeros_like(y) for y in alone]
    tickets = [srv.submit(mid, k, 900 + i, y) for i, ((srv, mid), k, y) in enumerate(zip(plan, ks, mixed))]
    for (srv, _), t in zip(plan, tickets):
        srv.wait(t)
    s1.close()
    s2.close()
    for i, (ya, ym) in enumerate(zip(alone, mixed)):
        np.testing.assert_array_equal(ya, ym, err_msg=f"request {i}")


def test_submit_and_wait(ctx, resident):
    s, ids = _server(resident, streams=2, max_batch=4)
    A = resident[0][0]
    ys = [np.zeros(A.height * 2, np.float32) for _ in range(5)]
    tickets = [s.submit(ids[0], 2, 77, y) for y in ys]
    for t in tickets:
        assert s.wait(t) >= 0
    for y in ys[1:]:
        np.testing.assert_array_equal(y, ys[0])
    with pytest.raises(cs.ClspmvError):
        s.submit(len(ids), 1, 1)        # no such matrix
    with pytest.raises(cs.ClspmvError):
        s.submit(ids[0], 9, 1)          # k above kmax
    s.close()
