"""RNS base conversion: GPU parity with the oracle (which tests/test_oracle.py
pins against big-integer CRT), bit for bit, fast and exact."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P59 = [576460752303421441, 576460752303419393, 576460752303415297, 576460752303403009]
P50 = [1125899906826241, 1125899906629633, 1125899906424833, 1125899906260993]

CASES = [
    ([31, 37, 41, 43], [47, 53, 59, 61]),  # primus_rns/src/converter.rs:331-332
    (P59[:3], P50[:2]),
    (P50, P59),
    ([2147473409, 4294957057, 1152921504606830593, 12289, 17], [576460752303421441, 97]),
    ([P59[0]], [P50[0], 65537]),  # one input modulus: the factor is one, inputs are reduced
    # six moduli just below 2^62: the integer kernel sums at most 4 products per
    # reduction (floor(2^64 / max q)), so each dot product is reduced in two parts
    ([4611686018427387847, 4611686018427387817, 4611686018427387787, 4611686018427387761,
      4611686018427387751, 4611686018427387737], [P59[0], P50[0]]),
]


def _run(conv_np, conv_dev, x, log_n, count, exact):
    import primus_amd as pa

    n = 1 << log_n
    rows_out = 1 if exact else len(conv_np.obase)
    din = pa.DeviceBuffer.from_array(x)
    dout = pa.DeviceBuffer(count * rows_out * n)
    (conv_dev.exact_convert_device if exact else conv_dev.fast_convert_device)(din.ptr, dout.ptr, log_n, count)
    pa.synchronize()
    got = dout.download((count, rows_out, n))
    for c in range(count):
        want = (conv_np.exact_convert if exact else conv_np.fast_convert)(x[c])
        np.testing.assert_array_equal(got[c], want, err_msg="poly %d" % c)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("log_n,count", [(1, 3), (10, 4), (12, 2), (16, 1)])
def test_fast_convert(case, log_n, count):
    from primus_amd.rns import BaseConverter

    ibase, obase = CASES[case]
    n = 1 << log_n
    rng = np.random.default_rng(case * 100 + log_n)
    x = np.stack([np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in ibase]) for _ in range(count)])
    if len(ibase) == 1:
        x[..., : n // 2] = rng.integers(0, 2**63, size=x[..., : n // 2].shape, dtype=np.uint64)  # non-canonical
    _run(O.BaseConverter(ibase, obase), BaseConverter(ibase, obase), x, log_n, count, exact=False)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("log_n,count", [(3, 2), (12, 2), (15, 1)])
def test_exact_convert(case, log_n, count):
    from primus_amd.rns import BaseConverter

    ibase, obase = CASES[case]
    n = 1 << log_n
    rng = np.random.default_rng(case * 1000 + log_n)
    x = np.stack([np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in ibase]) for _ in range(count)])
    x[0, :, 0] = 0
    x[0, :, 1] = [q - 1 for q in ibase]
    _run(O.BaseConverter(ibase, obase[:1]), BaseConverter(ibase, obase[:1]), x, log_n, count, exact=True)


def test_converter_rejects():
    import primus_amd as pa
    from primus_amd.rns import BaseConverter

    with pytest.raises(ValueError):
        BaseConverter([6, 9], [7])  # not coprime
    with pytest.raises(ValueError):
        BaseConverter([1 << 62], [7])  # modulus too large
    c = BaseConverter([31, 37], [41, 43])
    d = pa.DeviceBuffer(16)
    with pytest.raises(ValueError):
        c.exact_convert_device(d.ptr, pa.DeviceBuffer(16).ptr, 2, 1)  # two output moduli
