"""Continuous-batching server: results equal the oracle and are the same alone
and in a mix; admission refuses what can never fit; malformed requests are
rejected at submission."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

EXT_Q = 4294957057


def _grid(min_ln, max_ln, limbs):
    import primus_amd as pa

    return {ln: pa.ntt_primes_below(59, 2 << ln, limbs) for ln in range(min_ln, max_ln + 1)}


def _mix(grid, ext_key_np, count, seed, max_limbs):
    from primus_amd import serving as sv
    from test_rlwe import samples

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        kind = [sv.POLY_MUL, sv.RLWE_ENCRYPT, sv.EXT_PRODUCT][i % 3]
        if kind == sv.POLY_MUL:
            ln = int(rng.integers(10, 15))
            L = int(rng.integers(1, max_limbs + 1))
            n = 1 << ln
            a = np.concatenate([rng.integers(0, grid[ln][j], size=n, dtype=np.uint64) for j in range(L)])
            b = np.concatenate([rng.integers(0, grid[ln][j], size=n, dtype=np.uint64) for j in range(L)])
            reqs.append(sv.Request(kind, ln, L, in0=a, in1=b))
        elif kind == sv.RLWE_ENCRYPT:
            ln = int(rng.integers(10, 15))
            L = int(rng.integers(1, max_limbs + 1))
            n = 1 << ln
            a = np.concatenate([rng.integers(0, grid[ln][j], size=n, dtype=np.uint64) for j in range(L)])
            s, e, m = samples(n, rng)
            reqs.append(sv.Request(kind, ln, L, in0=a, s=s, e=e, m=m))
        else:
            x = rng.integers(0, EXT_Q, size=2 * 1024, dtype=np.uint64)
            reqs.append(sv.Request(kind, 10, 1, in0=x))
    return reqs


def _expected(req, grid, ext_key_np):
    from primus_amd import serving as sv
    from test_rlwe import delta_of, oracle_encrypt

    n = 1 << req.log_n
    if req.kind == sv.POLY_MUL:
        out = []
        for j in range(req.limbs):
            q = grid[req.log_n][j]
            t = O.Table(req.log_n, q)
            a, b = req.in0[j * n:(j + 1) * n], req.in1[j * n:(j + 1) * n]
            out.append(t.inverse_transform(O.pointwise_mul(t.transform(a), t.transform(b), q)))
        return np.concatenate(out)
    if req.kind == sv.RLWE_ENCRYPT:
        mods = grid[req.log_n][:req.limbs]
        return np.concatenate([oracle_encrypt(O.Table(req.log_n, q), req.s, req.e, req.m, delta_of(mods, q),
                                              req.in0[j * n:(j + 1) * n]) for j, q in enumerate(mods)])
    t = O.Table(10, EXT_Q)
    return O.ext_product(t, O.Basis.new(EXT_Q, 8, 3), ext_key_np, req.in0.reshape(2, 1024)).ravel()


def _server(max_batch, streams, budget=64 << 30, max_limbs=4, per_poly=False):
    import primus_amd as pa
    from primus_amd import serving as sv
    from primus_amd.fhe import ExtKey

    grid = _grid(10, 14, max_limbs)
    rng = np.random.default_rng(99)
    key_np = rng.integers(0, EXT_Q, size=(2, 3, 2, 1024), dtype=np.uint64)
    ext_tab = pa.UintNttTable.new(10, EXT_Q)
    key = ExtKey(ext_tab, 8, 3, key_np)
    srv = sv.Server(grid, 10, 14, max_limbs, ext_key=key, hbm_budget_bytes=budget, max_batch=max_batch,
                    streams=streams, per_poly_launches=per_poly)
    return srv, grid, key_np, key


def test_server_mix_matches_oracle_and_alone():
    from primus_amd import serving as sv

    srv, grid, key_np, key = _server(max_batch=64, streams=3)
    reqs = _mix(grid, key_np, 36, seed=5, max_limbs=4)
    tickets = [srv.submit(r) for r in reqs]
    for t in tickets:
        assert srv.wait(t) == 0
    mixed = [r.out.copy() for r in reqs]
    for r, got in zip(reqs, mixed):
        np.testing.assert_array_equal(got, _expected(r, grid, key_np))
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] >= 1
    srv.close()
    # the same requests one at a time, one launch per limb-polynomial (the
    # plain path): identical results
    alone, grid2, _, key2 = _server(max_batch=1, streams=1, per_poly=True)
    for r, got in zip(reqs, mixed):
        r.out[:] = 0
        assert alone.wait(alone.submit(r)) == 0
        np.testing.assert_array_equal(r.out, got)
    n_poly_limbs = sum(r.limbs for r in reqs if r.kind == sv.POLY_MUL)
    assert alone.stats()["launches"] == n_poly_limbs + sum(r.kind != sv.POLY_MUL for r in reqs)
    alone.close()


def test_server_page_locked_requests():
    """Requests in page-locked memory (copied without staging) mixed with
    ordinary ones, in one batch and alone: identical results, equal to the oracle."""
    srv, grid, key_np, key = _server(max_batch=64, streams=2)
    reqs = _mix(grid, key_np, 24, seed=8, max_limbs=4)
    mixed = [r.pinned_copy() if i % 2 else r for i, r in enumerate(reqs)]
    assert all(isinstance(r.out, np.ndarray) for r in mixed)
    tickets = [srv.submit(r) for r in mixed]
    for t in tickets:
        assert srv.wait(t) == 0
    for r in mixed:
        np.testing.assert_array_equal(r.out, _expected(r, grid, key_np))
    # all page-locked: a batch without any staging copy
    pinned = [r.pinned_copy() for r in reqs]
    for t in [srv.submit(r) for r in pinned]:
        assert srv.wait(t) == 0
    for r, m in zip(pinned, mixed):
        np.testing.assert_array_equal(r.out, m.out)
    srv.close()


def test_server_poisson_load_and_budget():
    srv, grid, key_np, key = _server(max_batch=32, streams=2, budget=256 << 20)
    reqs = _mix(grid, key_np, 60, seed=6, max_limbs=4)
    lat, elapsed = srv.run_poisson(reqs, rate_per_s=2000.0, seed=1)
    assert np.all(lat > 0) and elapsed > 0
    assert srv.stats()["peak_inflight_bytes"] <= 256 << 20
    # batches of up to 32 requests exceed the size above which staging copies
    # are split over helper threads: every result is checked
    for r in reqs:
        np.testing.assert_array_equal(r.out, _expected(r, grid, key_np))
    srv.close()


def test_server_rejects():
    from primus_amd import serving as sv

    srv, grid, key_np, key = _server(max_batch=8, streams=1, budget=1 << 20)
    n = 1 << 14
    big = sv.Request(sv.POLY_MUL, 14, 4, in0=np.zeros(4 * n, dtype=np.uint64), in1=np.zeros(4 * n, dtype=np.uint64))
    with pytest.raises(Exception):
        srv.submit(big)  # 1.5 MiB of device memory can never fit a 1 MiB budget
    bad = sv.Request(sv.POLY_MUL, 16, 1, in0=np.zeros(1 << 16, dtype=np.uint64),
                     in1=np.zeros(1 << 16, dtype=np.uint64))
    with pytest.raises(ValueError):
        srv.submit(bad)  # ring degree outside the server's grid
    wrong_ext = sv.Request(sv.EXT_PRODUCT, 11, 1, in0=np.zeros(2 << 11, dtype=np.uint64))
    with pytest.raises(ValueError):
        srv.submit(wrong_ext)  # the key is for N = 2^10
    ok = sv.Request(sv.EXT_PRODUCT, 10, 1, in0=np.zeros(2 << 10, dtype=np.uint64))
    assert srv.wait(srv.submit(ok)) == 0
    srv.close()


def test_server_full_range_matches_oracle_and_alone():
    """The serving workload's whole range: ring degrees up to 2^16 (two-pass
    transforms, RLWE scratch) and 5-16 limbs in one mix, against the oracle
    and against each request run alone."""
    import primus_amd as pa
    from primus_amd import serving as sv
    from primus_amd.fhe import ExtKey
    from test_rlwe import samples

    grid = _grid(10, 16, 16)
    rng = np.random.default_rng(123)
    key_np = rng.integers(0, EXT_Q, size=(2, 3, 2, 1024), dtype=np.uint64)
    key = ExtKey(pa.UintNttTable.new(10, EXT_Q), 8, 3, key_np)
    shapes = [(sv.POLY_MUL, 15, 5), (sv.POLY_MUL, 16, 16), (sv.POLY_MUL, 16, 7), (sv.POLY_MUL, 15, 16),
              (sv.RLWE_ENCRYPT, 16, 16), (sv.RLWE_ENCRYPT, 15, 12), (sv.RLWE_ENCRYPT, 16, 5),
              (sv.POLY_MUL, 10, 9), (sv.RLWE_ENCRYPT, 11, 16), (sv.EXT_PRODUCT, 10, 1), (sv.POLY_MUL, 13, 14),
              (sv.EXT_PRODUCT, 10, 1)]
    reqs = []
    for kind, ln, L in shapes:
        n = 1 << ln
        if kind == sv.EXT_PRODUCT:
            reqs.append(sv.Request(kind, 10, 1, in0=rng.integers(0, EXT_Q, size=2048, dtype=np.uint64)))
            continue
        a = np.concatenate([rng.integers(0, grid[ln][j], size=n, dtype=np.uint64) for j in range(L)])
        if kind == sv.POLY_MUL:
            b = np.concatenate([rng.integers(0, grid[ln][j], size=n, dtype=np.uint64) for j in range(L)])
            reqs.append(sv.Request(kind, ln, L, in0=a, in1=b))
        else:
            s, e, m = samples(n, rng)
            reqs.append(sv.Request(kind, ln, L, in0=a, s=s, e=e, m=m))
    print("full-range mix:", [(r.kind, r.log_n, r.limbs) for r in reqs], flush=True)
    srv = sv.Server(grid, 10, 16, 16, ext_key=key, hbm_budget_bytes=8 << 30, max_batch=64, streams=2)
    srv.pause()  # every request joins one batch
    tickets = [srv.submit(r) for r in reqs]
    srv.pause(False)
    for t in tickets:
        assert srv.wait(t) == 0
    assert srv.stats()["batches"] == 1
    mixed = [r.out.copy() for r in reqs]
    srv.close()
    for r, got in zip(reqs, mixed):
        want = _expected(r, grid, key_np)
        np.testing.assert_array_equal(got, want, err_msg="kind %d N=2^%d L=%d" % (r.kind, r.log_n, r.limbs))
    alone = sv.Server(grid, 10, 16, 16, ext_key=key, hbm_budget_bytes=8 << 30, max_batch=1, streams=1,
                      per_poly_launches=True)
    for r, got in zip(reqs, mixed):
        r.out[:] = 0
        assert alone.wait(alone.submit(r)) == 0
        np.testing.assert_array_equal(r.out, got)
    alone.close()


def test_server_cancel_queued_requests():
    """Requests cancelled while queued complete at once with PH_ERR_CANCELLED
    and their output buffers are never written; the rest run normally."""
    srv, grid, key_np, key = _server(max_batch=16, streams=2)
    reqs = _mix(grid, key_np, 9, seed=21, max_limbs=4)
    srv.pause()
    tickets = [srv.submit(r) for r in reqs]
    cancelled = set(tickets[1::3])
    for t in cancelled:
        assert srv.cancel(t)
    assert not srv.cancel(tickets[1])  # cancelling twice is refused
    srv.pause(False)
    for t, r in zip(tickets, reqs):
        st = srv.wait(t)
        if t in cancelled:
            assert st == 9  # PH_ERR_CANCELLED
            assert not r.out.any()
        else:
            assert st == 0
            np.testing.assert_array_equal(r.out, _expected(r, grid, key_np))
    assert not srv.cancel(tickets[0])  # finished
    assert srv.stats()["cancelled"] == len(cancelled)
    srv.close()


def test_server_injected_fault_reports_per_request_status():
    """A batch whose device work fails reports PH_ERR_DEVICE on every request
    in it and leaves their outputs untouched; the server then goes on."""
    srv, grid, key_np, key = _server(max_batch=16, streams=2)
    reqs = _mix(grid, key_np, 6, seed=22, max_limbs=3)
    srv.inject_fault(1)
    srv.pause()
    bad = [srv.submit(r) for r in reqs[:3]]
    srv.pause(False)
    for t, r in zip(bad, reqs[:3]):
        assert srv.wait(t) == 6  # PH_ERR_DEVICE
        assert not r.out.any()
    good = [srv.submit(r) for r in reqs[3:]]
    for t, r in zip(good, reqs[3:]):
        assert srv.wait(t) == 0
        np.testing.assert_array_equal(r.out, _expected(r, grid, key_np))
    srv.close()


def test_server_free_drains_outstanding_work():
    """Freeing the server right after submitting still completes every
    request, with its results in the caller's buffers."""
    srv, grid, key_np, key = _server(max_batch=4, streams=2)
    reqs = _mix(grid, key_np, 20, seed=23, max_limbs=4)
    for r in reqs:
        srv.submit(r)
    srv.close()
    for r in reqs:
        np.testing.assert_array_equal(r.out, _expected(r, grid, key_np))
