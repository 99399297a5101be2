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
