"""External product (RLWE x RGSW, k = 1): oracle pinning on the CPU, GPU parity.

The gadget decomposition is pinned by its defining property (signed digits
that reconstruct the centred value up to the dropped, rounded low bits), the
product by the schoolbook identity out_j = sum_{p,l} d_{p,l} * INTT(key[p][l][j])
negacyclically mod q.
"""
import numpy as np
import pytest

from oracle import oracle as O

EXT_Q32 = 4294957057  # largest prime < 2^32 with q = 1 mod 2048


def _basis_reference(q, log_basis, l):
    """The basis constants, re-derived from primus_decompose/src/big_integer/basis.rs:34-195."""
    bits = q.bit_length()
    drop = bits - l * log_basis
    if log_basis == 1:
        split = None if drop == 0 else (((1 << (l + 1)) - 1) << (drop - 1))
    else:
        v = 0
        for _ in range(l):
            v = (v << log_basis) | ((1 << log_basis) - 1) >> 1
        v = ((v << 1) | 1) << (drop - 1) if drop > 0 else v + 1
        split = v
    if split is not None and split >= q:
        split = None
    return drop, split, (1 << bits) - q


@pytest.mark.parametrize("q,log_basis,l", [(EXT_Q32, 8, 3), (EXT_Q32, 8, 4), (EXT_Q32, 4, 5), (132120577, 7, 3),
                                           (1125899906826241, 10, 4), (576460752303421441, 20, 2),
                                           (12289, 1, 13), (12289, 1, 10)])
def test_basis_constants(q, log_basis, l):
    b = O.Basis.new(q, log_basis, l)
    drop, split, p2 = _basis_reference(q, log_basis, l)
    assert b.drop_bits == drop and b.pow2_sub_modulus == p2 and b.modulus_bits == q.bit_length()
    assert bool(b.has_split) == (split is not None)
    if split is not None:
        assert b.split == split
    assert b.carry_mask == (2 if log_basis == 1 else (1 << log_basis) | (1 << (log_basis - 1)))


@pytest.mark.parametrize("q,log_basis,l", [(EXT_Q32, 8, 3), (EXT_Q32, 4, 6), (132120577, 7, 3),
                                           (1125899906826241, 10, 4), (576460752303421441, 20, 2)])
def test_decomposition_reconstructs(q, log_basis, l):
    """sum_l s_l 2^(drop + l*B) = centred(v) + delta with |delta| <= 2^(drop-1), digits |s_l| <= 2^(B-1)."""
    b = O.Basis.new(q, log_basis, l)
    rng = np.random.default_rng(q % 1000 + l)
    edge = [0, 1, 2, q - 1, q - 2, q // 2, q // 2 + 1, b.split - 1 if b.has_split else 3,
            b.split if b.has_split else 4, (1 << (b.drop_bits - 1)) if b.drop_bits else 5]
    vals = np.array(edge + list(rng.integers(0, q, size=2000)), dtype=np.uint64)
    res = b.decompose(vals)
    B, drop = 1 << log_basis, b.drop_bits
    for i, v in enumerate(vals.tolist()):
        s = [int(r) if int(r) < B // 2 else int(r) - q for r in res[:, i]]
        assert all(-B // 2 <= x < B // 2 for x in s)
        rec = sum(x << (drop + k * log_basis) for k, x in enumerate(s))
        centred = v - q if (b.has_split and v >= b.split) else v
        delta = rec - centred
        # the top carry is dropped: equality holds modulo 2^bits(q)
        m = 1 << q.bit_length()
        delta = (delta + m // 2) % m - m // 2
        assert abs(delta) <= (1 << (drop - 1) if drop else 0), (v, s, delta)


def _negacyclic(a, b, q):
    n = len(a)
    c = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            if k < n:
                c[k] += int(a[i]) * int(b[j])
            else:
                c[k - n] -= int(a[i]) * int(b[j])
    return np.array([x % q for x in c], dtype=np.uint64)


@pytest.mark.parametrize("q,log_n,log_basis,l", [(12289, 4, 4, 3), (132120577, 5, 7, 3), (EXT_Q32, 5, 8, 3)])
def test_oracle_ext_product_schoolbook(q, log_n, log_basis, l):
    n = 1 << log_n
    t = O.Table(log_n, q)
    b = O.Basis.new(q, log_basis, l)
    rng = np.random.default_rng(log_n)
    key = rng.integers(0, q, size=(2, l, 2, n), dtype=np.uint64)
    x = rng.integers(0, q, size=(2, n), dtype=np.uint64)
    got = O.ext_product(t, b, key, x)
    want = np.zeros((2, n), dtype=np.uint64)
    for p in range(2):
        digits = b.decompose(x[p])
        for d in range(l):
            for j in range(2):
                kc = t.inverse_transform(key[p, d, j])
                want[j] = (want[j] + _negacyclic(digits[d], kc, q)) % np.uint64(q)
    np.testing.assert_array_equal(got, want)
    # NTT-form output is the forward transform of the coefficient-form one
    got_ntt = O.ext_product(t, b, key, x, ntt_output=True)
    np.testing.assert_array_equal(got_ntt, t.transform(got))


@pytest.mark.parametrize("q,log_basis,l", [(EXT_Q32, 8, 3), (EXT_Q32, 8, 4), (EXT_Q32, 1, 32), (12289, 14, 1),
                                           (12289, 4, 3), (1125899906826241, 10, 5), (576460752303421441, 20, 2)])
def test_unsigned_digits_reconstruct(q, log_basis, l):
    """The literal unsigned digits of v < q: each below the basis, and their
    weighted sum is v with its low drop_bits cleared, exactly (no carry, no
    split); residues are the digits mod q."""
    b = O.Basis.new(q, log_basis, l)
    rng = np.random.default_rng(q % 101 + l)
    v = rng.integers(0, q, size=400, dtype=np.uint64)
    v[:6] = [0, 1, q - 1, q - 2, q // 2, (1 << b.drop_bits) - 1 if b.drop_bits else 3]
    res = b.decompose_unsigned(v)
    raw = [[(int(x) >> (b.drop_bits + d * log_basis)) & ((1 << log_basis) - 1) for x in v] for d in range(l)]
    for k, x in enumerate(v):
        total = sum(raw[d][k] << (b.drop_bits + d * log_basis) for d in range(l))
        assert total == int(x) >> b.drop_bits << b.drop_bits
        for d in range(l):
            assert raw[d][k] < (1 << log_basis)
            assert int(res[d, k]) == raw[d][k] % q


@pytest.mark.parametrize("q,log_n,log_basis,l", [(12289, 4, 4, 3), (132120577, 5, 7, 3), (EXT_Q32, 5, 8, 3)])
def test_oracle_ext_product_unsigned_schoolbook(q, log_n, log_basis, l):
    """The unsigned-digit external product equals the sum of schoolbook
    negacyclic products of each digit polynomial with its key row."""
    n = 1 << log_n
    t = O.Table(log_n, q)
    b = O.Basis.new(q, log_basis, l)
    rng = np.random.default_rng(log_n + 7)
    key = rng.integers(0, q, size=(2, l, 2, n), dtype=np.uint64)
    x = rng.integers(0, q, size=(2, n), dtype=np.uint64)
    got = O.ext_product(t, b, key, x, unsigned_digits=True)
    want = np.zeros((2, n), dtype=np.uint64)
    for p in range(2):
        for d in range(l):
            digits = np.array([(int(v) >> (b.drop_bits + d * log_basis)) & ((1 << log_basis) - 1) for v in x[p]],
                              dtype=np.uint64)
            for j in range(2):
                kc = t.inverse_transform(key[p, d, j])
                want[j] = (want[j] + _negacyclic(digits, kc, q)) % np.uint64(q)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(O.ext_product(t, b, key, x, ntt_output=True, unsigned_digits=True),
                                  t.transform(got))


# ---------------------------------------------------------------- GPU parity

@pytest.mark.gpu
@pytest.mark.parametrize("integer_only", [False, True])
@pytest.mark.parametrize("q,log_n,log_basis,l,count", [
    (EXT_Q32, 10, 8, 3, 64),                  # cfg 4's parameters, read as literal unsigned digits
    (EXT_Q32, 10, 8, 4, 5),                   # every bit a digit bit (drop_bits 0)
    (12289, 6, 14, 1, 3),                     # basis above q: digits up to 2q, reduced by the transform
    (1125899906826241, 12, 10, 4, 3),
    (576460752303421441, 10, 20, 2, 4),       # > 2^50: integer path either way
])
def test_gpu_ext_product_unsigned_digits(q, log_n, log_basis, l, count, integer_only):
    """PH_DIGITS_UNSIGNED against the oracle's restatement of the literal
    unsigned digits, coefficient and NTT-form outputs."""
    import primus_amd as pa
    from primus_amd.fhe import ExtKey

    n = 1 << log_n
    print("ext_product unsigned", q, log_n, log_basis, l, count, integer_only, flush=True)
    tab = pa.UintNttTable.new(log_n, q, integer_only=integer_only)
    rng = np.random.default_rng(q % 977 + l)
    key = rng.integers(0, q, size=(2, l, 2, n), dtype=np.uint64)
    x = rng.integers(0, q, size=(count, 2, n), dtype=np.uint64)
    x[0, 0, :6] = [0, 1, q - 1, q - 2, q // 2, q // 2 + 1]
    ek = ExtKey(tab, log_basis, l, key, unsigned_digits=True)
    got = ek.product(x)
    got_ntt = ek.product(x, ntt_output=True)
    ot = O.Table(log_n, q)
    ob = O.Basis.new(q, log_basis, l)
    for i in range(count):
        np.testing.assert_array_equal(got[i], O.ext_product(ot, ob, key, x[i], unsigned_digits=True), err_msg=str(i))
        np.testing.assert_array_equal(got_ntt[i], O.ext_product(ot, ob, key, x[i], ntt_output=True,
                                                                unsigned_digits=True), err_msg=str(i))
    # the reference's digits give a different product on the same inputs,
    # except where both represent every value exactly (one digit, no bits dropped)
    if l > 1 or ob.drop_bits > 0:
        assert not np.array_equal(ExtKey(tab, log_basis, l, key).product(x[:1])[0], got[0])


@pytest.mark.gpu
@pytest.mark.parametrize("integer_only", [False, True])
@pytest.mark.parametrize("q,log_n,log_basis,l,count", [
    (EXT_Q32, 10, 8, 3, 2048),                 # the full batch of 2048 products
    (EXT_Q32, 10, 8, 4, 7),
    (132120577, 8, 7, 3, 5),
    (1125899906826241, 12, 10, 4, 3),
    (1125899906826241, 10, 5, 9, 2),          # 18 key products: binary64 accumulators reduced twice
    (576460752303421441, 10, 20, 2, 4),       # > 2^50: integer path either way
    (12289, 6, 1, 13, 3),                     # binary basis
])
def test_gpu_ext_product(q, log_n, log_basis, l, count, integer_only):
    import primus_amd as pa
    from primus_amd.fhe import ExtKey

    n = 1 << log_n
    print("ext_product", q, log_n, log_basis, l, count, integer_only, flush=True)
    tab = pa.UintNttTable.new(log_n, q, integer_only=integer_only)
    rng = np.random.default_rng(q % 991 + l)
    key = rng.integers(0, q, size=(2, l, 2, n), dtype=np.uint64)
    x = rng.integers(0, q, size=(count, 2, n), dtype=np.uint64)
    x[0, 0, :8] = [0, 1, q - 1, q - 2, q // 2, q // 2 + 1, 2, 3]
    ek = ExtKey(tab, log_basis, l, key)
    got = ek.product(x)
    got_ntt = ek.product(x, ntt_output=True)
    ot = O.Table(log_n, q)
    ob = O.Basis.new(q, log_basis, l)
    check = range(count) if count <= 64 else list(range(16)) + list(range(count - 16, count))
    for i in check:
        np.testing.assert_array_equal(got[i], O.ext_product(ot, ob, key, x[i]), err_msg=str(i))
        np.testing.assert_array_equal(got_ntt[i], O.ext_product(ot, ob, key, x[i], ntt_output=True), err_msg=str(i))


@pytest.mark.gpu
def test_gpu_ext_product_device_batch_properties():
    """Full configuration batch: linearity in the key (a size-independent property)."""
    import primus_amd as pa
    from primus_amd.fhe import ExtKey

    q, log_n, log_basis, l, count = EXT_Q32, 10, 8, 3, 2048
    n = 1 << log_n
    tab = pa.UintNttTable.new(log_n, q)
    rng = np.random.default_rng(44)
    k1 = rng.integers(0, q, size=(2, l, 2, n), dtype=np.uint64)
    k2 = rng.integers(0, q, size=(2, l, 2, n), dtype=np.uint64)
    x = rng.integers(0, q, size=(count, 2, n), dtype=np.uint64)
    ks = ((k1.astype(object) + k2.astype(object)) % q).astype(np.uint64)
    r1 = ExtKey(tab, log_basis, l, k1).product(x, ntt_output=True)
    r2 = ExtKey(tab, log_basis, l, k2).product(x, ntt_output=True)
    rs = ExtKey(tab, log_basis, l, ks).product(x, ntt_output=True)
    np.testing.assert_array_equal(rs, ((r1.astype(object) + r2.astype(object)) % q).astype(np.uint64))
