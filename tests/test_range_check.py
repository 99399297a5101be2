"""The range-checking debug variant of the library (libprimus_hip_rc.so, built
with PH_RANGE_CHECK=1): every value at the transform kernels' register-round
boundaries is checked against its arithmetic's range -- the integer
butterflies' [0, 4q) forward and [0, 2q) inverse values, canonical [0, q)
outputs, and the wider documented ranges of the canonical-only arithmetics --
the device counterpart of the reference's debug_assert!ed ranges
(primus_ntt/src/ntt/hexl/mod.rs:452-465, 546-558).

GPU cases run each operation through the default library and through the
checked variant on the same inputs at the edges of the input contracts
(4q - 1 forward, 2q - 1 inverse), expect identical outputs and no violation.
A planted contract violation (lazy forward inputs at and above 4q, values
only: no address depends on them) must be counted, with its site.
"""
import contextlib
import ctypes
import gc

import numpy as np
import pytest

import primus_amd as pa
from primus_amd import _lib

SITES = {1: "forward input", 2: "forward round", 3: "inverse input", 4: "inverse round", 5: "inverse output"}


def test_default_build_has_no_range_checks():
    rep = (ctypes.c_uint64 * 4)()
    assert _lib.lib().ph_debug_range_report(rep, 0) == 7  # PH_ERR_UNSUPPORTED
    assert list(rep) == [0, 0, 0, 0]


def test_checked_variant_exports_the_header():
    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.RC_LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert exported == set(_lib.header_functions())


def _report(clear=True):
    rep = (ctypes.c_uint64 * 4)()
    code = _lib.lib_rc().ph_debug_range_report(rep, int(clear))
    assert code == 0, _lib.lib_rc().ph_last_error()
    return list(rep)


@contextlib.contextmanager
def _checked():
    """Runs the block on the range-checking variant; every object made inside
    must be released inside (the helpers below return plain arrays)."""
    saved = _lib._lib
    _lib._lib = _lib.lib_rc()
    try:
        pa.synchronize()
        _report(clear=True)
        yield
    finally:
        gc.collect()
        pa.synchronize()
        _lib._lib = saved


def _edge_inputs(rng, moduli, n, factor):
    """[L][n] values in [0, factor*q) with every tenth at factor*q - 1."""
    rows = []
    for q in moduli:
        r = rng.integers(0, factor * q, size=n, dtype=np.uint64)
        r[::10] = factor * q - 1
        rows.append(r)
    return np.stack(rows)


def _dcrt_ops(log_n, moduli, seed, count=3):
    """Lazy and canonical forward, inverse, fused forward->multiply->inverse and
    negacyclic product over a small batch; returns every output."""
    rng = np.random.default_rng(seed)
    n, L = 1 << log_n, len(moduli)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    fwd_in = np.concatenate([_edge_inputs(rng, moduli, n, 4) for _ in range(count)])
    inv_in = np.concatenate([_edge_inputs(rng, moduli, n, 2) for _ in range(count)])
    canon = np.concatenate([_edge_inputs(rng, moduli, n, 1) for _ in range(count)])
    outs = []
    for lazy in (True, False):
        d = pa.DeviceBuffer.from_array(fwd_in)
        pa.transform_device(tab, d.ptr, count * L, lazy=lazy)
        pa.synchronize()
        outs.append(d.download())
    d = pa.DeviceBuffer.from_array(inv_in)
    pa.inverse_transform_device(tab, d.ptr, count * L)
    pa.synchronize()
    outs.append(d.download())
    op = pa.Operand(tab, _edge_inputs(rng, moduli, n, 1))
    d = pa.DeviceBuffer.from_array(canon)
    pa.fwd_mul_inv_device(tab, d.ptr, count * L, op)
    pa.synchronize()
    outs.append(d.download())
    da, db = pa.DeviceBuffer.from_array(canon), pa.DeviceBuffer.from_array(canon[::-1].copy())
    pa.negacyclic_mul_device(tab, da.ptr, db.ptr, count * L)
    pa.synchronize()
    outs.append(da.download())
    return outs


CASES = [
    (12, [50, 50, 50, 50]),   # binary64 arithmetic
    (12, [59, 59]),           # canonical-only integer arithmetic
    (12, [62, 62]),           # integer Shoup arithmetic
    (6, [50]),
    (13, [50, 50]),
    (14, [59, 59]),           # two-pass
    (16, [50]),
    (17, [61]),
]


@pytest.mark.gpu
@pytest.mark.parametrize("log_n,bits", CASES)
def test_transforms_within_ranges(log_n, bits):
    moduli = []
    for b in sorted(set(bits)):
        moduli += pa.ntt_primes_below(b, 2 << log_n, bits.count(b))
    print("range check: log_n", log_n, "moduli", moduli, flush=True)
    want = _dcrt_ops(log_n, moduli, seed=log_n)
    with _checked():
        got = _dcrt_ops(log_n, moduli, seed=log_n)
        rep = _report()
    assert rep[0] == 0, "%d violations, first at %s: value bits %#x, q %d" % (rep[0], SITES.get(rep[1]), rep[2], rep[3])
    for w, g in zip(want, got):
        np.testing.assert_array_equal(g, w)


def _hexl_lazy(log_n, q, seed):
    rng = np.random.default_rng(seed)
    t = pa.HexlNttTable.new(log_n, q)
    x = rng.integers(0, 2 * q, size=1 << log_n, dtype=np.uint64)
    x[::7] = 2 * q - 1
    t.compute_inverse(x, 2, 2)
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("log_n,bits", [(11, 60), (14, 61)])
def test_hexl_lazy_inverse_within_ranges(log_n, bits):
    q = pa.ntt_primes_below(bits, 2 << log_n, 1)[0]
    want = _hexl_lazy(log_n, q, 5)
    with _checked():
        got = _hexl_lazy(log_n, q, 5)
        rep = _report()
    assert rep[0] == 0, (rep, SITES.get(rep[1]))
    np.testing.assert_array_equal(got, want)


def _rlwe(seed):
    from primus_amd.fhe import RlweContext, rlwe_layout

    shapes = [(10, 3), (11, 1), (12, 2), (13, 2), (14, 1), (15, 2)]
    grid = {ln: pa.ntt_primes_below(59, 2 << ln, 3) for ln in range(10, 16)}
    ctx = RlweContext(10, 15, 3, grid)
    reqs, a_words, small = rlwe_layout(shapes)
    rng = np.random.default_rng(seed)
    a = np.zeros(a_words, dtype=np.uint64)
    for (log_n, limbs, ao, so) in reqs:
        n = 1 << log_n
        for i in range(limbs):
            a[ao + i * n: ao + (i + 1) * n] = rng.integers(0, grid[log_n][i], size=n, dtype=np.uint64)
    s = rng.integers(-1, 2, size=small).astype(np.int8)
    e = np.clip(np.rint(rng.normal(0, 3.2, size=small)), -19, 19).astype(np.int8)
    m = rng.integers(0, 256, size=small).astype(np.uint8)

    def words(v):
        raw = np.ascontiguousarray(v).tobytes()
        return np.frombuffer(raw.ljust((len(raw) + 7) // 8 * 8, b"\0"), dtype=np.uint64)

    da, dout = pa.DeviceBuffer.from_array(a), pa.DeviceBuffer(a_words)
    ds, de, dm = (pa.DeviceBuffer.from_array(words(v)) for v in (s, e, m))
    ctx.encrypt_device(reqs, da.ptr, ds.ptr, de.ptr, dm.ptr, dout.ptr)
    pa.synchronize()
    return dout.download()


@pytest.mark.gpu
def test_rlwe_encryption_within_ranges():
    want = _rlwe(11)
    with _checked():
        got = _rlwe(11)
        rep = _report()
    assert rep[0] == 0, (rep, SITES.get(rep[1]))
    np.testing.assert_array_equal(got, want)


def _ext_rns(bits, L, log_b, l, seed):
    from primus_amd.fhe import ExtKeyRns

    log_n, n, count = 10, 1024, 4
    moduli = pa.ntt_primes_below(bits, 2 << log_n, L)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    rng = np.random.default_rng(seed)
    key = np.stack([rng.integers(0, q, size=(2, l, 2, n), dtype=np.uint64) for q in moduli], axis=3)
    x = np.stack([rng.integers(0, q, size=(count, 2, n), dtype=np.uint64) for q in moduli], axis=2)
    k = ExtKeyRns(tab, log_b, l, key)
    din, dout = pa.DeviceBuffer.from_array(x), pa.DeviceBuffer(x.size)
    k.product_device(din.ptr, dout.ptr, count)
    pa.synchronize()
    return dout.download()


@pytest.mark.gpu
@pytest.mark.parametrize("bits,L,log_b,l", [(59, 2, 8, 14), (50, 3, 16, 9), (62, 2, 20, 6)])
def test_ext_product_rns_within_ranges(bits, L, log_b, l):
    want = _ext_rns(bits, L, log_b, l, 3)
    with _checked():
        got = _ext_rns(bits, L, log_b, l, 3)
        rep = _report()
    assert rep[0] == 0, (rep, SITES.get(rep[1]))
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
def test_planted_violation_is_counted():
    """Lazy forward inputs at and above 4q break the forward input contract:
    the checked build counts them at the forward-input site."""
    log_n = 12
    q = pa.ntt_primes_below(62, 2 << log_n, 1)[0]
    x = np.full((1, 1 << log_n), 4 * q, dtype=np.uint64)
    x[0, 1::2] = np.uint64(2 ** 63 + 12345)
    with _checked():
        tab = pa.UintCrtNttTable.new(log_n, [q])
        d = pa.DeviceBuffer.from_array(x)
        pa.transform_device(tab, d.ptr, 1, lazy=True)
        pa.synchronize()
        rep = _report()
        del d, tab
    assert rep[0] > 0
    assert rep[3] == q
    assert SITES[rep[1]] in ("forward input", "forward round")
