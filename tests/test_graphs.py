"""HIP-graph capture of the batched forms: every replay of a captured graph
equals the eager result.  Between replays the outputs are cleared by eager
kernels (torch zero_) and the stream is synchronised, so each replay is
checked on its own.  The graphs hold kernel launches and device-to-device
copies only (no memset nodes)."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

REPLAYS = 3


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _replay_and_check(torch, graph, stream, outputs, expected):
    for r in range(REPLAYS):
        for o in outputs:
            o.zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            graph.replay()
        stream.synchronize()
        for o, want in zip(outputs, expected):
            got = o.cpu().numpy().view(np.uint64)
            np.testing.assert_array_equal(got, want, err_msg="replay %d" % r)


def test_graph_fused_batches():
    """Three fused forward -> multiply -> inverse batches on separate buffers, each
    restored from a pristine copy inside the graph."""
    torch = _torch()
    import primus_amd as pa

    log_n, L, count = 12, 4, 96
    n = 1 << log_n
    moduli = pa.ntt_primes_below(50, 2 * n, L)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    rng = np.random.default_rng(21)
    polys = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    opv = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in moduli])
    op = pa.Operand(tab, opv)
    dev = torch.device("cuda:0")
    src = torch.from_numpy(polys.view(np.int64)).to(dev)
    bufs = [torch.empty_like(src) for _ in range(3)]
    stream = torch.cuda.Stream(device=dev)
    # eager reference
    eager = src.clone()
    pa.fwd_mul_inv_device(tab, _vp(eager), count, op, stream=stream.cuda_stream)
    stream.synchronize()
    want = eager.cpu().numpy().view(np.uint64)
    tables = [O.Table(log_n, q) for q in moduli]
    for i in (0, count - 1):
        t = tables[i % L]
        np.testing.assert_array_equal(want[i], t.inverse_transform(O.pointwise_mul(t.transform(polys[i]),
                                                                                 opv[i % L], t.q)))
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=stream):
        for b in bufs:
            b.copy_(src)
            pa.fwd_mul_inv_device(tab, _vp(b), count, op, stream=stream.cuda_stream)
    _replay_and_check(torch, g, stream, bufs, [want] * 3)


def test_graph_rlwe_plan():
    """An uneven RLWE batch through a prepared plan, captured and replayed."""
    torch = _torch()
    import primus_amd as pa
    from primus_amd.fhe import RlweContext, rlwe_layout

    rng = np.random.default_rng(22)
    shapes = [(int(rng.integers(10, 15)), int(rng.integers(1, 5))) for _ in range(24)]
    grid = {ln: pa.ntt_primes_below(59, 2 << ln, 4) for ln in range(10, 15)}
    ctx = RlweContext(10, 14, 4, grid)
    reqs, a_words, small = rlwe_layout(shapes)
    a = np.zeros(a_words, dtype=np.uint64)
    for (ln, lb, ao, so) in reqs:
        for i in range(lb):
            a[ao + (i << ln): ao + ((i + 1) << ln)] = rng.integers(0, grid[ln][i], size=1 << ln, dtype=np.uint64)
    s = rng.integers(-1, 2, size=small).astype(np.int8)
    e = np.clip(np.rint(rng.normal(0, 3.2, size=small)), -19, 19).astype(np.int8)
    m = rng.integers(0, 256, size=small).astype(np.uint8)
    dev = torch.device("cuda:0")
    da = torch.from_numpy(a.view(np.int64)).to(dev)
    ds, de, dm = (torch.from_numpy(v.copy()).to(dev) for v in (s, e, m))
    dout = torch.zeros_like(da)
    stream = torch.cuda.Stream(device=dev)
    ctx.encrypt_device(reqs, _vp(da), _vp(ds), _vp(de), _vp(dm), _vp(dout), stream=stream.cuda_stream)
    stream.synchronize()
    want = dout.cpu().numpy().view(np.uint64).copy()
    ln, lb, ao, so = reqs[0]
    mods = grid[ln][:lb]
    np.testing.assert_array_equal(
        want[ao:ao + (1 << ln)],
        O.rlwe_encrypt(O.Table(ln, mods[0]), s[so:so + (1 << ln)], e[so:so + (1 << ln)], m[so:so + (1 << ln)],
                       O.rlwe_delta(mods, mods[0]), a[ao:ao + (1 << ln)]))
    plan = ctx.plan(reqs)
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=stream):
        plan.encrypt_device(_vp(da), _vp(ds), _vp(de), _vp(dm), _vp(dout), stream=stream.cuda_stream)
    _replay_and_check(torch, g, stream, [dout], [want])


def test_graph_ext_product():
    torch = _torch()
    import primus_amd as pa
    from primus_amd.fhe import ExtKey

    q, log_n, count = 4294957057, 10, 64
    n = 1 << log_n
    rng = np.random.default_rng(23)
    key_np = rng.integers(0, q, size=(2, 3, 2, n), dtype=np.uint64)
    key = ExtKey(pa.UintNttTable.new(log_n, q), 8, 3, key_np)
    x = rng.integers(0, q, size=(count, 2, n), dtype=np.uint64)
    dev = torch.device("cuda:0")
    din = torch.from_numpy(x.view(np.int64)).to(dev)
    dout = torch.zeros_like(din)
    stream = torch.cuda.Stream(device=dev)
    key.product_device(_vp(din), _vp(dout), count, stream=stream.cuda_stream)
    stream.synchronize()
    want = dout.cpu().numpy().view(np.uint64).copy()
    np.testing.assert_array_equal(want[0], O.ext_product(O.Table(log_n, q), O.Basis.new(q, 8, 3), key_np, x[0]))
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=stream):
        key.product_device(_vp(din), _vp(dout), count, stream=stream.cuda_stream)
    _replay_and_check(torch, g, stream, [dout], [want])


@pytest.mark.parametrize("bits,L", [(59, 2), (50, 3), (62, 2)])
def test_graph_ext_product_rns(bits, L):
    """The fused external product over several limbs (one launch, no
    allocation), captured and replayed, for each of its arithmetics."""
    torch = _torch()
    import primus_amd as pa
    from primus_amd.fhe import ExtKeyRns

    log_n, count, l = 10, 24, 3
    n = 1 << log_n
    moduli = pa.ntt_primes_below(bits, 2 * n, L)
    table = pa.UintCrtNttTable.new(log_n, moduli)
    rng = np.random.default_rng(29 + L)
    qcol = np.array(moduli, dtype=np.uint64).reshape(-1, 1)
    key_np = np.stack([rng.integers(0, 1 << 63, size=(L, n), dtype=np.uint64) % qcol
                       for _ in range(4 * l)]).reshape(2, l, 2, L, n)
    key = ExtKeyRns(table, 8, l, key_np)
    x = np.stack([rng.integers(0, 1 << 63, size=(L, n), dtype=np.uint64) % qcol
                  for _ in range(2 * count)]).reshape(count, 2, L, n)
    dev = torch.device("cuda:0")
    din = torch.from_numpy(x.view(np.int64)).to(dev)
    dout = torch.zeros_like(din)
    stream = torch.cuda.Stream(device=dev)
    key.product_device(_vp(din), _vp(dout), count, stream=stream.cuda_stream)
    stream.synchronize()
    want = dout.cpu().numpy().view(np.uint64).copy()
    tables = [O.Table(log_n, q) for q in moduli]
    np.testing.assert_array_equal(want[count - 1], O.ext_product_rns(tables, O.RnsBasis(moduli, 8, l), key_np,
                                                                    x[count - 1]))
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=stream):
        key.product_device(_vp(din), _vp(dout), count, stream=stream.cuda_stream)
    _replay_and_check(torch, g, stream, [dout], [want])


def test_graph_rlwe_sample_and_encrypt():
    """Device sampling and encryption of an uneven batch, both from one plan,
    captured together and replayed."""
    torch = _torch()
    import primus_amd as pa
    from primus_amd.fhe import RlweContext, rlwe_layout

    shapes = [(10, 2), (12, 3), (14, 1), (11, 4), (13, 2)]
    grid = {ln: pa.ntt_primes_below(59, 2 << ln, 4) for ln in range(10, 15)}
    ctx = RlweContext(10, 14, 4, grid)
    reqs, a_words, small = rlwe_layout(shapes)
    plan = ctx.plan(reqs)
    dev = torch.device("cuda:0")
    da = torch.zeros(a_words, dtype=torch.int64, device=dev)
    dout = torch.zeros_like(da)
    ds, de = (torch.zeros(small, dtype=torch.int8, device=dev) for _ in range(2))
    dm = torch.zeros(small, dtype=torch.uint8, device=dev)
    stream = torch.cuda.Stream(device=dev)
    seed = 0xFEEDF00D
    plan.sample_device(seed, _vp(da), _vp(ds), _vp(de), _vp(dm), stream=stream.cuda_stream)
    plan.encrypt_device(_vp(da), _vp(ds), _vp(de), _vp(dm), _vp(dout), stream=stream.cuda_stream)
    stream.synchronize()
    want = [t.cpu().numpy().copy() for t in (da, dout)]
    ln, lb, ao, so = reqs[1]
    mods = grid[ln][:lb]
    s_, e_, m_, a_ = O.rlwe_sample(seed, 1, ln, mods)
    np.testing.assert_array_equal(want[0].view(np.uint64)[ao:ao + lb * (1 << ln)], a_.ravel())
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=stream):
        plan.sample_device(seed, _vp(da), _vp(ds), _vp(de), _vp(dm), stream=stream.cuda_stream)
        plan.encrypt_device(_vp(da), _vp(ds), _vp(de), _vp(dm), _vp(dout), stream=stream.cuda_stream)
    _replay_and_check(torch, g, stream, [da, dout], [w.view(np.uint64) for w in want])
