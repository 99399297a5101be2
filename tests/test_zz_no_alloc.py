"""ph_dcrt_negacyclic_mul_device_noalloc: the negacyclic product with its
two-pass scratch in a fixed device array instead of an allocation.  Its
outputs must equal ph_dcrt_negacyclic_mul_device's bit for bit, on two calls
in a row (the second smaller, so it runs over what the first left in the
array), and a batch beyond the array is refused before anything runs."""
import ctypes

import numpy as np
import pytest

import primus_amd as pa
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LIMIT_WORDS = (64 << 20) // 8  # PH_NOALLOC_SCRATCH_BYTES / 8


def _inputs(log_n, moduli, count, seed):
    rng = np.random.default_rng(seed)
    n = 1 << log_n
    L = len(moduli)
    a = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    b = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    return a, b


def _run(fn, table, a, b):
    da, db = pa.DeviceBuffer.from_array(a), pa.DeviceBuffer.from_array(b)
    pa._lib.check(fn(table.handle, da.ptr, db.ptr, a.shape[0], None))
    pa.synchronize()
    out = da.download(a.shape)
    np.testing.assert_array_equal(db.download(b.shape), b)  # the second operand is left intact
    return out


@pytest.mark.parametrize("log_n,bits,counts", [
    (15, 59, (64, 24)),     # two-pass, integer arithmetic
    (16, 50, (128, 40)),    # two-pass at the array's full size (128 x 2^16 words), binary64 tables
    (14, 31, (96, 7)),
    (12, 50, (300, 5)),     # single pass: no scratch used
])
def test_noalloc_equals_allocating_entry_point(log_n, bits, counts):
    L = pa.lib()
    moduli = pa.ntt_primes_below(bits, 2 << log_n, 3)
    table = pa.UintCrtNttTable.new(log_n, moduli)
    for k, count in enumerate(counts):
        print("noalloc N=2^%d bits=%d count=%d call=%d" % (log_n, bits, count, k), flush=True)
        a, b = _inputs(log_n, moduli, count, seed=1000 * log_n + k)
        want = _run(L.ph_dcrt_negacyclic_mul_device, table, a, b)
        got = _run(L.ph_dcrt_negacyclic_mul_device_noalloc, table, a, b)
        np.testing.assert_array_equal(got, want, err_msg="call %d" % k)
    # and the last row of the last call against the oracle
    q = moduli[(count - 1) % len(moduli)]
    t = O.Table(log_n, q)
    np.testing.assert_array_equal(got[-1], t.inverse_transform(O.pointwise_mul(t.transform(a[-1]),
                                                                               t.transform(b[-1]), q)))


def test_noalloc_refuses_batches_beyond_the_array():
    L = pa.lib()
    log_n = 16
    moduli = pa.ntt_primes_below(59, 2 << log_n, 1)
    table = pa.UintCrtNttTable.new(log_n, moduli)
    count = LIMIT_WORDS // (1 << log_n) + 1
    buf = pa.DeviceBuffer(count << log_n)
    rc = L.ph_dcrt_negacyclic_mul_device_noalloc(table.handle, buf.ptr, buf.ptr, count, None)
    assert rc == 5  # PH_ERR_INVALID, refused on the host
    assert b"PH_NOALLOC_SCRATCH_BYTES" in L.ph_last_error()
    assert L.ph_dcrt_negacyclic_mul_device_noalloc(table.handle, buf.ptr, buf.ptr, 0, None) == 0
