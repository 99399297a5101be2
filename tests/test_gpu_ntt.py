"""GPU parity of the HIP NTT path against the oracle and the golden vectors.

Every comparison is bit-exact: the lazy forward output too, since the kernels
run the reference's Harvey butterflies one for one.
"""
import ctypes

import numpy as np
import pytest

import primus_amd as pa
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BATCH_PRIMES_50 = [1125899906826241, 1125899906629633, 1125899906424833, 1125899906260993]


def _rand(q, shape, seed):
    return np.random.default_rng(seed).integers(0, q, size=shape, dtype=np.uint64)


# ---------------------------------------------------------------- golden vectors

@pytest.mark.parametrize("integer_only", [False, True])
def test_golden_slices(golden, integer_only):
    for label, c in golden.items():
        q, log_n = int(c["q"][0]), int(c["log_n"][0])
        print("golden", label, integer_only, flush=True)
        t = pa.UintNttTable.new(log_n, q, integer_only=integer_only)
        assert t.arith() == ("f64" if (q < 1 << 50 and not integer_only) else "int")
        assert t.root() == int(c["psi"][0]), label
        x = c["a"].copy()
        t.transform_slice(x)
        np.testing.assert_array_equal(x, c["fwd_a"], err_msg=label)
        t.inverse_transform_slice(x)
        np.testing.assert_array_equal(x, c["a"], err_msg=label)
        if "inv_in" in c:
            y = c["inv_in"].copy()
            t.inverse_transform_slice(y)
            np.testing.assert_array_equal(y, c["inv_out"], err_msg=label)
        coeff, deg = (int(v) for v in c["mono_params"])
        m = np.zeros(1 << log_n, dtype=np.uint64)
        t.transform_monomial(coeff, deg, m)
        np.testing.assert_array_equal(m, c["mono"], err_msg=label)
        t.transform_coeff_one_monomial(deg, m)
        np.testing.assert_array_equal(m, c["mono_one"], err_msg=label)
        t.transform_coeff_minus_one_monomial(deg, m)
        np.testing.assert_array_equal(m, c["mono_minus_one"], err_msg=label)


@pytest.mark.parametrize("integer_only", [False, True])
def test_golden_negacyclic_device(golden, integer_only):
    for label, c in golden.items():
        if "prod" not in c:
            continue
        q, log_n = int(c["q"][0]), int(c["log_n"][0])
        tab = pa.UintCrtNttTable.new(log_n, [q], integer_only=integer_only)
        da, db = pa.DeviceBuffer.from_array(c["a"]), pa.DeviceBuffer.from_array(c["b"])
        pa.negacyclic_mul_device(tab, da.ptr, db.ptr, 1)
        pa.synchronize()
        np.testing.assert_array_equal(da.download(), c["prod"], err_msg=label)
        np.testing.assert_array_equal(db.download(), c["b"], err_msg=label)


# ---------------------------------------------------------------- every ring degree

@pytest.mark.parametrize("log_n", list(range(1, 18)))
def test_all_degrees_bit_exact(log_n):
    n = 1 << log_n
    q = pa.ntt_primes_below(59, 2 * n, 1)[0]
    print("degree", log_n, q, flush=True)
    t = pa.UintNttTable.new(log_n, q)
    ot = O.Table(log_n, q)
    assert t.root() == ot.root and t.inv_root() == ot.inv_root and t.inv_n() == ot.inv_n
    rp, rpq = t.root_powers()
    orp, orpq = ot.powers(0)
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(rpq, orpq)
    rng = np.random.default_rng(log_n)
    # lazy forward from [0, 4q): bit-identical lazy values
    a = (rng.integers(0, q, size=n, dtype=np.uint64)
         + rng.integers(0, 4, size=n, dtype=np.uint64) * np.uint64(q)).astype(np.uint64)
    x = a.copy()
    t.lazy_transform_slice(x)
    np.testing.assert_array_equal(x, ot.lazy_transform(a))
    x = a.copy()
    t.transform_slice(x)
    np.testing.assert_array_equal(x, ot.transform(a))
    # inverse from [0, 2q)
    b = (rng.integers(0, q, size=n, dtype=np.uint64)
         + rng.integers(0, 2, size=n, dtype=np.uint64) * np.uint64(q)).astype(np.uint64)
    y = b.copy()
    t.lazy_inverse_transform_slice(y)
    np.testing.assert_array_equal(y, ot.lazy_inverse_transform(b))
    y = b.copy()
    t.inverse_transform_slice(y)
    np.testing.assert_array_equal(y, ot.inverse_transform(b))


@pytest.mark.parametrize("q", [17, 12289, 2147473409, 4294957057, 1125899906826241, 1125899906260993,
                               576460752303415297, 1152921504606830593, 2305843009213683713,
                               4611686018427322369])
def test_moduli_edge_cases(q):
    log_n = min(10, ((q - 1) & -(q - 1)).bit_length() - 2)
    n = 1 << log_n
    t = pa.UintNttTable.new(log_n, q)
    ot = O.Table(log_n, q)
    for a in [np.zeros(n, dtype=np.uint64), np.full(n, q - 1, dtype=np.uint64),
              np.full(n, 4 * q - 1, dtype=np.uint64), _rand(q, n, 7)]:
        x = a.copy()
        t.transform_slice(x)
        np.testing.assert_array_equal(x, ot.transform(a))
        x = a.copy()
        t.lazy_transform_slice(x)
        np.testing.assert_array_equal(x, ot.lazy_transform(a))
    b = np.full(n, 2 * q - 1, dtype=np.uint64)
    y = b.copy()
    t.inverse_transform_slice(y)
    np.testing.assert_array_equal(y, ot.inverse_transform(b))


def test_slice_length_checked():
    t = pa.UintNttTable.new(10, 2147473409)
    with pytest.raises(ValueError):
        t.transform_slice(np.zeros(512, dtype=np.uint64))


# ---------------------------------------------------------------- DCRT and batched forms

def test_dcrt_slices():
    log_n = 12
    tab = pa.UintCrtNttTable.new(log_n, BATCH_PRIMES_50)
    assert tab.moduli_count() == 4 and tab.crt_poly_length() == 4 << log_n
    x = np.concatenate([_rand(q, 1 << log_n, i) for i, q in enumerate(BATCH_PRIMES_50)])
    want = np.concatenate([O.Table(log_n, q).transform(x[i << log_n:(i + 1) << log_n])
                           for i, q in enumerate(BATCH_PRIMES_50)])
    y = x.copy()
    tab.transform_slice(y)
    np.testing.assert_array_equal(y, want)
    tab.inverse_transform_slice(y)
    np.testing.assert_array_equal(y, x)


def _fmi_reference(tables, polys, op):
    out = polys.copy()
    L = len(tables)
    for i in range(polys.shape[0]):
        t = tables[i % L]
        out[i] = t.inverse_transform(O.pointwise_mul(t.transform(polys[i]), op[i % L], t.q))
    return out


@pytest.mark.parametrize("integer_only", [False, True])
@pytest.mark.parametrize("log_n,L,count", [(12, 4, 1024), (10, 3, 30), (13, 2, 8), (14, 2, 4),
                                           (16, 1, 2), (17, 2, 2), (5, 2, 6), (6, 2, 4)])
def test_fwd_mul_inv_batch(log_n, L, count, integer_only):
    n = 1 << log_n
    moduli = pa.ntt_primes_below(50, 2 * n, L)
    print("fwd_mul_inv", log_n, L, count, integer_only, flush=True)
    tab = pa.UintCrtNttTable.new(log_n, moduli, integer_only=integer_only)
    rng = np.random.default_rng(log_n * 100 + L)
    polys = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    op_vals = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in moduli])
    op = pa.Operand(tab, op_vals)
    d = pa.DeviceBuffer.from_array(polys)
    pa.fwd_mul_inv_device(tab, d.ptr, count, op)
    pa.synchronize()
    got = d.download((count, n))
    want = _fmi_reference([O.Table(log_n, q) for q in moduli], polys, op_vals)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("bits", [59, 50, 31])
@pytest.mark.parametrize("log_n,L,count", [(10, 1, 64), (12, 3, 6), (13, 1, 3), (14, 1, 2), (15, 2, 2),
                                           (16, 1, 1), (3, 1, 5), (7, 2, 4)])
def test_negacyclic_batch(log_n, L, count, bits):
    n = 1 << log_n
    moduli = pa.ntt_primes_below(bits, 2 * n, L)
    print("negacyclic", log_n, L, count, flush=True)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    rng = np.random.default_rng(log_n)
    a = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    b = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    da, db = pa.DeviceBuffer.from_array(a), pa.DeviceBuffer.from_array(b)
    pa.negacyclic_mul_device(tab, da.ptr, db.ptr, count)
    pa.synchronize()
    got = da.download((count, n))
    tables = [O.Table(log_n, q) for q in moduli]
    for i in range(count):
        t = tables[i % L]
        want = t.inverse_transform(O.pointwise_mul(t.transform(a[i]), t.transform(b[i]), t.q))
        np.testing.assert_array_equal(got[i], want)
        if n <= 256:
            np.testing.assert_array_equal(got[i], O.schoolbook(a[i], b[i], t.q))
    np.testing.assert_array_equal(db.download((count, n)), b)


@pytest.mark.parametrize("bits", [59, 50])
@pytest.mark.parametrize("log_n", [10, 12, 13, 14, 16])
def test_negacyclic_rows_uneven_limbs(log_n, bits):
    """Requests of one ring degree with uneven limb counts in one launch: row i
    uses limb limb_of[i] of a 4-limb table; a row naming no limb (7) stays as it
    was, and b is left intact."""
    n, L = 1 << log_n, 4
    moduli = pa.ntt_primes_below(bits, 2 * n, L)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    limbs = [3, 1, 4, 2] if log_n <= 13 else [2, 1]
    limb_of = np.array([j for k in limbs for j in range(k)] + [7], dtype=np.uint8)
    rows = limb_of.size
    print("negacyclic rows", log_n, bits, rows, flush=True)
    rng = np.random.default_rng(100 + log_n)
    qs = [moduli[j] if j < L else moduli[0] for j in limb_of]
    a = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in qs])
    b = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in qs])
    da, db = pa.DeviceBuffer.from_array(a), pa.DeviceBuffer.from_array(b)
    padded = np.zeros((rows + 7) // 8 * 8, dtype=np.uint8)
    padded[:rows] = limb_of
    dm = pa.DeviceBuffer.from_array(padded.view(np.uint64))
    pa.negacyclic_mul_rows_device(tab, da.ptr, db.ptr, dm.ptr, rows)
    pa.synchronize()
    got = da.download((rows, n))
    tables = [O.Table(log_n, q) for q in moduli]
    for i, j in enumerate(limb_of):
        if j >= L:
            np.testing.assert_array_equal(got[i], a[i], err_msg="row %d without a limb" % i)
            continue
        t = tables[j]
        want = t.inverse_transform(O.pointwise_mul(t.transform(a[i]), t.transform(b[i]), t.q))
        np.testing.assert_array_equal(got[i], want, err_msg="row %d limb %d" % (i, j))
    np.testing.assert_array_equal(db.download((rows, n)), b)


@pytest.mark.parametrize("bits", [62, 50, 40])
def test_device_forward_inverse_batch(bits):
    log_n, L, count = 11, 3, 12
    n = 1 << log_n
    moduli = pa.ntt_primes_below(bits, 2 * n, L)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    tables = [O.Table(log_n, q) for q in moduli]
    rng = np.random.default_rng(3)
    x = np.stack([rng.integers(0, 4 * moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    d = pa.DeviceBuffer.from_array(x)
    pa.transform_device(tab, d.ptr, count, lazy=True)
    pa.synchronize()
    lz = d.download((count, n))
    for i in range(count):
        np.testing.assert_array_equal(lz[i], tables[i % L].lazy_transform(x[i]))
    d.upload(x)
    pa.transform_device(tab, d.ptr, count, lazy=False)
    pa.synchronize()
    canon = d.download((count, n))
    for i in range(count):
        np.testing.assert_array_equal(canon[i], tables[i % L].transform(x[i]))
    # inverse from [0, 2q)
    y = (canon + np.stack([rng.integers(0, 2, size=n, dtype=np.uint64) * np.uint64(moduli[i % L])
                           for i in range(count)])).astype(np.uint64)
    d.upload(y)
    pa.inverse_transform_device(tab, d.ptr, count)
    pa.synchronize()
    back = d.download((count, n))
    for i in range(count):
        np.testing.assert_array_equal(back[i], x[i] % np.uint64(moduli[i % L]))


def test_pointwise_device():
    log_n, L, count = 10, 2, 6
    n = 1 << log_n
    moduli = pa.ntt_primes_below(62, 2 * n, L)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    rng = np.random.default_rng(9)
    a = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    b = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    c = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    da, db, dc = (pa.DeviceBuffer.from_array(v) for v in (a, b, c))
    dout = pa.DeviceBuffer(count * n)
    pa.pointwise_mul_device(tab, da.ptr, db.ptr, dout.ptr, count)
    pa.pointwise_mul_add_device(tab, dc.ptr, da.ptr, db.ptr, count)
    pa.synchronize()
    out, acc = dout.download((count, n)), dc.download((count, n))
    for i in range(count):
        q = moduli[i % L]
        np.testing.assert_array_equal(out[i], O.pointwise_mul(a[i], b[i], q))
        np.testing.assert_array_equal(acc[i], O.pointwise_mul_add(c[i], a[i], b[i], q))


@pytest.mark.parametrize("q", [1125899906826241, 1125899906260993, 2147473409])
def test_f64_extreme_operands(q):
    """Largest magnitudes the binary64 path can meet: every input and operand q-1 (or 4q-1)."""
    log_n = 12 if q > 1 << 40 else 10
    n = 1 << log_n
    tab = pa.UintCrtNttTable.new(log_n, [q])
    assert tab.arith() == "f64"
    ot = O.Table(log_n, q)
    rng = np.random.default_rng(q % 997)
    cases = [np.full(n, q - 1, dtype=np.uint64), np.full(n, 4 * q - 1, dtype=np.uint64),
             np.where(rng.integers(0, 2, size=n) == 1, q - 1, 0).astype(np.uint64)]
    for a in cases:
        for opv in [np.full(n, q - 1, dtype=np.uint64), rng.integers(0, q, size=n, dtype=np.uint64)]:
            op = pa.Operand(tab, opv)
            d = pa.DeviceBuffer.from_array(a)
            pa.fwd_mul_inv_device(tab, d.ptr, 1, op)
            pa.synchronize()
            want = ot.inverse_transform(O.pointwise_mul(ot.transform(a), opv, q))
            np.testing.assert_array_equal(d.download(), want)
        b = np.minimum(a, np.uint64(q - 1))
        da, db = pa.DeviceBuffer.from_array(b), pa.DeviceBuffer.from_array(np.full(n, q - 1, dtype=np.uint64))
        pa.negacyclic_mul_device(tab, da.ptr, db.ptr, 1)
        pa.synchronize()
        want = ot.inverse_transform(O.pointwise_mul(ot.transform(b), ot.transform(np.full(n, q - 1, dtype=np.uint64)), q))
        np.testing.assert_array_equal(da.download(), want)


@pytest.mark.parametrize("q", [576460752303415297, 2305843009213554689, 4611686018427322369, 1125899906826241])
def test_integer_extreme_operands(q):
    """The integer butterflies at their largest magnitudes: the approximate-
    quotient path (q < 2^61, intermediates up to 8q) at a 59-bit prime and at
    the largest prime below 2^61 with q = 1 mod 8192, and the exact path above
    it, with every input 4q - 1 (or q - 1) and every operand q - 1."""
    log_n = 12
    n = 1 << log_n
    tab = pa.UintCrtNttTable.new(log_n, [q], integer_only=True)
    ot = O.Table(log_n, q)
    rng = np.random.default_rng(q % 991)
    cases = [np.full(n, q - 1, dtype=np.uint64), np.full(n, 4 * q - 1, dtype=np.uint64),
             np.where(rng.integers(0, 2, size=n) == 1, 4 * q - 1, 0).astype(np.uint64)]
    for a in cases:
        fw = ot.transform(a)
        x = pa.DeviceBuffer.from_array(a)
        pa.transform_device(tab, x.ptr, 1)
        pa.synchronize()
        np.testing.assert_array_equal(x.download(), fw)
        for opv in [np.full(n, q - 1, dtype=np.uint64), rng.integers(0, q, size=n, dtype=np.uint64)]:
            op = pa.Operand(tab, opv)
            d = pa.DeviceBuffer.from_array(a)
            pa.fwd_mul_inv_device(tab, d.ptr, 1, op)
            pa.synchronize()
            np.testing.assert_array_equal(d.download(), ot.inverse_transform(O.pointwise_mul(fw, opv, q)))
    b = np.full(n, 2 * q - 1, dtype=np.uint64)
    y = pa.DeviceBuffer.from_array(b)
    pa.inverse_transform_device(tab, y.ptr, 1)
    pa.synchronize()
    np.testing.assert_array_equal(y.download(), ot.inverse_transform(b))
    qm = np.full(n, q - 1, dtype=np.uint64)
    da, db = pa.DeviceBuffer.from_array(qm), pa.DeviceBuffer.from_array(qm)
    pa.negacyclic_mul_device(tab, da.ptr, db.ptr, 1)
    pa.synchronize()
    np.testing.assert_array_equal(da.download(), ot.inverse_transform(O.pointwise_mul(ot.transform(qm),
                                                                                      ot.transform(qm), q)))


@pytest.mark.parametrize("log_n", [10, 11, 12, 13])
@pytest.mark.parametrize("bits", [50, 59])
def test_launch_shapes_agree(log_n, bits):
    """Every autotuner launch shape gives the oracle's outputs."""
    from primus_amd import tuning

    n = 1 << log_n
    L, count = 2, 6
    moduli = pa.ntt_primes_below(bits, 2 * n, L)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    arith = tab.arith()
    tables = [O.Table(log_n, q) for q in moduli]
    rng = np.random.default_rng(log_n + bits)
    a = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    b = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    opv = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in moduli])
    op = pa.Operand(tab, opv)
    try:
        for shape in tuning.shapes(log_n, "fwd_mul_inv"):
            for name in ("forward", "inverse", "fwd_mul_inv", "negacyclic_mul", "forward_lazy"):
                tuning.set_shape(log_n, name, arith, shape)
                tuning.set_shape(log_n, name, "int", shape)
            d = pa.DeviceBuffer.from_array(a)
            pa.fwd_mul_inv_device(tab, d.ptr, count, op)
            pa.synchronize()
            got = d.download((count, n))
            for i in range(count):
                t = tables[i % L]
                np.testing.assert_array_equal(got[i], t.inverse_transform(O.pointwise_mul(t.transform(a[i]),
                                                                                        opv[i % L], t.q)))
            d.upload(a)
            db = pa.DeviceBuffer.from_array(b)
            pa.negacyclic_mul_device(tab, d.ptr, db.ptr, count)
            pa.synchronize()
            got = d.download((count, n))
            for i in range(count):
                t = tables[i % L]
                np.testing.assert_array_equal(got[i], t.inverse_transform(O.pointwise_mul(
                    t.transform(a[i]), t.transform(b[i]), t.q)))
            d.upload(a)
            pa.transform_device(tab, d.ptr, count, lazy=True)
            pa.synchronize()
            got = d.download((count, n))
            for i in range(count):
                np.testing.assert_array_equal(got[i], tables[i % L].lazy_transform(a[i]))
            d.upload(a)
            pa.transform_device(tab, d.ptr, count)
            pa.inverse_transform_device(tab, d.ptr, count)
            pa.synchronize()
            np.testing.assert_array_equal(d.download((count, n)), a)
        shape, ms = tuning.autotune(tab, "fwd_mul_inv", 64, reps=3)
        assert shape in tuning.shapes(log_n, "fwd_mul_inv") and ms > 0
    finally:
        for name in ("forward", "inverse", "fwd_mul_inv", "negacyclic_mul", "forward_lazy"):
            tuning.set_shape(log_n, name, "f64", 0)
            tuning.set_shape(log_n, name, "int", 0)


@pytest.mark.parametrize("log_n", [10, 11, 12, 13])
@pytest.mark.parametrize("bits", [50, 59])
def test_streaming_shapes_long_batch(log_n, bits):
    """The persistent streaming kernels on batches several resident generations
    long (every workgroup walks many polynomials): forward canonical, forward
    lazy and fused forward -> multiply -> inverse equal the oracle."""
    from primus_amd import tuning

    n = 1 << log_n
    L = 3
    count = (1 << 23) // n + 5  # 8 Mi coefficients: above four generations of every shape
    moduli = pa.ntt_primes_below(bits, 2 * n, L)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    tables = [O.Table(log_n, q) for q in moduli]
    rng = np.random.default_rng(1000 + log_n + bits)
    a = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    opv = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in moduli])
    op = pa.Operand(tab, opv)
    sample = sorted(set(range(6)) | set(range(count - 6, count)) |
                    set(int(i) for i in rng.integers(0, count, size=12)))
    d = pa.DeviceBuffer.from_array(a)
    print("streaming", log_n, bits, count, flush=True)
    try:
        for shape in [e | tuning.STREAM | tuning.lpb(k) for e in (3, 4) for k in range(4)]:
            for name in ("forward", "fwd_mul_inv", "forward_lazy"):
                tuning.set_shape(log_n, name, tab.arith(), shape)
                tuning.set_shape(log_n, name, "int", shape)
            for kind in ("forward", "forward_lazy", "fwd_mul_inv"):
                d.upload(a)
                if kind == "fwd_mul_inv":
                    pa.fwd_mul_inv_device(tab, d.ptr, count, op)
                else:
                    pa.transform_device(tab, d.ptr, count, lazy=(kind == "forward_lazy"))
                pa.synchronize()
                got = d.download((count, n))
                for i in sample:
                    t = tables[i % L]
                    if kind == "forward":
                        want = t.transform(a[i])
                    elif kind == "forward_lazy":
                        want = t.lazy_transform(a[i])
                    else:
                        want = t.inverse_transform(O.pointwise_mul(t.transform(a[i]), opv[i % L], t.q))
                    np.testing.assert_array_equal(got[i], want, err_msg="%s shape %d poly %d" % (kind, shape, i))
    finally:
        for name in ("forward", "fwd_mul_inv", "forward_lazy"):
            tuning.set_shape(log_n, name, tab.arith(), 0)
            tuning.set_shape(log_n, name, "int", 0)


@pytest.mark.parametrize("log_n", [14, 15, 16, 17])
def test_two_pass_tiles_agree(log_n):
    """Every two-pass shape the autotuner tries (LDS tile of 2^11, 2^12 or
    2^13 coefficients, 8 or 16 per thread in the chunk pass) gives the
    oracle's outputs: lazy and canonical forward, inverse, fused forward ->
    multiply -> inverse and the negacyclic product."""
    from primus_amd import tuning

    n = 1 << log_n
    L, count = 2, 3
    moduli = pa.ntt_primes_below(59, 2 * n, L)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    tables = [O.Table(log_n, q) for q in moduli]
    rng = np.random.default_rng(2000 + log_n)
    a = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    b = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    opv = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in moduli])
    op = pa.Operand(tab, opv)
    fwd = [tables[i % L].transform(a[i]) for i in range(count)]
    lazy = [tables[i % L].lazy_transform(a[i]) for i in range(count)]
    fmi = [tables[i % L].inverse_transform(O.pointwise_mul(fwd[i], opv[i % L], moduli[i % L])) for i in range(count)]
    prod = [tables[i % L].inverse_transform(O.pointwise_mul(fwd[i], tables[i % L].transform(b[i]), moduli[i % L]))
            for i in range(count)]
    shapes = tuning.shapes(log_n, "fwd_mul_inv")
    assert len(shapes) == (6 if log_n < 17 else 4)
    d, db = pa.DeviceBuffer.from_array(a), pa.DeviceBuffer.from_array(b)
    names = ("forward", "inverse", "fwd_mul_inv", "negacyclic_mul", "forward_lazy")
    try:
        for shape in shapes:
            print("two-pass", log_n, "shape", shape, flush=True)
            for name in names:
                tuning.set_shape(log_n, name, tab.arith(), shape)
                tuning.set_shape(log_n, name, "int", shape)
            for kind, want in (("forward", fwd), ("forward_lazy", lazy)):
                d.upload(a)
                pa.transform_device(tab, d.ptr, count, lazy=(kind == "forward_lazy"))
                pa.synchronize()
                got = d.download((count, n))
                for i in range(count):
                    np.testing.assert_array_equal(got[i], want[i], err_msg="%s shape %d" % (kind, shape))
            d.upload(np.stack(fwd))
            pa.inverse_transform_device(tab, d.ptr, count)
            pa.synchronize()
            np.testing.assert_array_equal(d.download((count, n)), a, err_msg="inverse shape %d" % shape)
            d.upload(a)
            pa.fwd_mul_inv_device(tab, d.ptr, count, op)
            pa.synchronize()
            np.testing.assert_array_equal(d.download((count, n)), np.stack(fmi), err_msg="fused shape %d" % shape)
            d.upload(a)
            pa.negacyclic_mul_device(tab, d.ptr, db.ptr, count)
            pa.synchronize()
            np.testing.assert_array_equal(d.download((count, n)), np.stack(prod), err_msg="product shape %d" % shape)
    finally:
        for name in names:
            tuning.set_shape(log_n, name, "f64", 0)
            tuning.set_shape(log_n, name, "int", 0)


@pytest.mark.parametrize("log_n", [1, 3, 9, 12, 13, 16, 17])
def test_automorphism_device(log_n):
    """NTT-domain automorphisms on limb-major batches equal the oracle's
    permutation, row by row, for odd degrees including 1, 2N-1 and N+1."""
    n = 1 << log_n
    L, count = 2, 5
    moduli = pa.ntt_primes_below(59, 2 * n, L)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    rng = np.random.default_rng(log_n + 77)
    x = np.stack([rng.integers(0, moduli[i % L], size=n, dtype=np.uint64) for i in range(count)])
    din, dout = pa.DeviceBuffer.from_array(x), pa.DeviceBuffer(count * n)
    for k in sorted({1, 3, 2 * n - 1, n + 1, 4 * n + 5, int(rng.integers(0, n)) * 2 + 1}):
        pa.automorphism_device(tab, din.ptr, dout.ptr, count, k)
        pa.synchronize()
        got = dout.download((count, n))
        for i in range(count):
            np.testing.assert_array_equal(got[i], O.ntt_automorphism(x[i], k), err_msg="k=%d row %d" % (k, i))


@pytest.mark.parametrize("log_n", [1, 10, 11, 12])
@pytest.mark.parametrize("aligned", [True, False])
def test_automorphism_device_many_rows(log_n, aligned):
    """Batches with more rows than the persistent grid of the LDS-staged form
    (so workgroups walk several rows, the last pass partial), 16-byte aligned
    rows and rows offset by one word (8-byte accesses): every row equals the
    oracle's permutation of it."""
    n = 1 << log_n
    rows = (512 << (12 - log_n)) + 37  # the launcher's grid for this degree, plus a partial pass
    q = pa.ntt_primes_below(50, 2 * n, 1)[0]
    tab = pa.UintCrtNttTable.new(log_n, [q])
    rng = np.random.default_rng(log_n * 7 + aligned)
    x = rng.integers(0, q, size=(rows, n), dtype=np.uint64)
    off = 0 if aligned else 1
    din = pa.DeviceBuffer.from_array(np.concatenate([np.zeros(off, np.uint64), x.ravel()]))
    dout = pa.DeviceBuffer(rows * n + off)
    for k in (5, 2 * n - 1):
        perm = O.ntt_automorphism(np.arange(n, dtype=np.uint64), k).astype(np.int64)
        pa.automorphism_device(tab, ctypes.c_void_p(din.ptr.value + 8 * off), ctypes.c_void_p(dout.ptr.value + 8 * off),
                               rows, k)
        pa.synchronize()
        got = dout.download()[off:].reshape(rows, n)
        np.testing.assert_array_equal(got, x[:, perm], err_msg="k=%d" % k)
        np.testing.assert_array_equal(got[rows - 1], O.ntt_automorphism(x[rows - 1], k), err_msg="k=%d last row" % k)


def test_automorphism_rejects():
    n = 1 << 10
    tab = pa.UintCrtNttTable.new(10, pa.ntt_primes_below(59, 2 * n, 1))
    d = pa.DeviceBuffer(4 * n)
    with pytest.raises(ValueError):
        pa.automorphism_device(tab, d.ptr, pa.DeviceBuffer(n).ptr, 1, 4)  # even degree
    with pytest.raises(ValueError):
        pa.automorphism_device(tab, d.ptr, ctypes.c_void_p(d.ptr.value + 8 * n), 2, 3)  # overlapping rows


@pytest.mark.parametrize("bit_shift,moduli", [
    (64, [576460752303415297, 1125899906826241]),   # ShoupFactor (mul_factor_assign)
    (52, [562949953392641, 1125899906826241]),      # HEXL's IFMA factor form
    (32, [536813569, 2147473409]),                  # HEXL's 32-bit factor form
])
def test_mul_factor_device(bit_shift, moduli):
    """MultiplyFactor's lazy_mul_modulo / mul_modulo (primus_factor/src/multiply.rs:50-70)
    on the device, per limb, against the oracle's restatement (pinned by exact
    remainders in test_oracle.py)."""
    log_n, count = 10, 6  # every modulus here is 1 mod 2048
    n = 1 << log_n
    L = len(moduli)
    tab = pa.UintCrtNttTable.new(log_n, moduli)
    rng = np.random.default_rng(bit_shift)
    bound = [min(4 * q, 1 << bit_shift) if bit_shift < 64 else 4 * q for q in moduli]
    x = np.stack([rng.integers(0, bound[i % L], size=n, dtype=np.uint64) for i in range(count)])
    x[0, :3] = [0, moduli[0] - 1, bound[0] - 1]
    factors = [int(rng.integers(0, q)) for q in moduli]
    factors[0] = moduli[0] - 1
    lib = O.lib()
    for lazy in (True, False):
        d = pa.DeviceBuffer.from_array(x)
        pa.mul_factor_device(tab, d.ptr, count, factors, bit_shift=bit_shift, lazy=lazy)
        pa.synchronize()
        got = d.download((count, n))
        for i in range(count):
            q, c = moduli[i % L], factors[i % L]
            bf = lib.po_multiply_factor(c, bit_shift, q)
            want = np.array([lib.po_multiply_mod_lazy(int(v), c, bf, bit_shift, q) for v in x[i]], dtype=np.uint64)
            if not lazy:
                want = np.where(want >= q, want - np.uint64(q), want)
            np.testing.assert_array_equal(got[i], want, err_msg="row %d lazy %s" % (i, lazy))
            assert int(got[i].max()) < (2 if lazy else 1) * q
    with pytest.raises(ValueError):
        pa.mul_factor_device(tab, d.ptr, count, factors, bit_shift=48)
    with pytest.raises(ValueError):
        pa.mul_factor_device(tab, d.ptr, count, [moduli[0] + 1] + factors[1:], bit_shift=bit_shift)

