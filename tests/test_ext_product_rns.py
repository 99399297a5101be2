"""Multi-limb external product (CrtGlwe::mul_dcrt_ggsw_inplace,
primus_lattice/src/glwe/crt.rs:201-229): the oracle's big-integer restatement
pinned on the CPU, then the device path against it bit for bit.

Pins of the oracle (no reference vectors exist for this path):
* with one limb it equals the C oracle's single-modulus basis and external
  product (po_basis / po_ext_product, pinned in test_ext_product.py);
* the composition is the CRT value: congruent to every residue, below Q;
* the digits reconstruct the composed value: sum of signed digits times
  their weights is within 2^drop_bits of it modulo Q, and each limb's digit
  residue is congruent to the signed digit.
"""
import numpy as np
import pytest

import primus_amd as pa
from oracle import oracle as O

EXT_Q = 4294957057


def _rng(seed):
    return np.random.default_rng(seed)


@pytest.mark.parametrize("q,log_b,l", [(EXT_Q, 8, 3), (EXT_Q, 8, 4), (EXT_Q, 1, 20), (1125899906826241, 10, 5),
                                       (576460752303421441, 4, 14), (12289, 2, 7)])
def test_rns_basis_one_limb_equals_c_oracle(q, log_b, l):
    v = _rng(q % 97 + l).integers(0, q, size=512, dtype=np.uint64)
    v[:4] = [0, 1, q - 1, q // 2]
    want = O.Basis.new(q, log_b, l).decompose(v)
    got = O.RnsBasis([q], log_b, l).decompose(v.reshape(1, -1))
    np.testing.assert_array_equal(got[:, 0, :], want)


def _moduli(L, bits=59, log_n=10):
    return pa.ntt_primes_below(bits, 2 << log_n, L)


@pytest.mark.parametrize("L,bits", [(2, 59), (3, 50), (4, 62), (2, 31)])
def test_rns_compose_is_crt(L, bits):
    moduli = _moduli(L, bits)
    b = O.RnsBasis(moduli, 8, 3)
    r = np.stack([_rng(L + i).integers(0, q, size=256, dtype=np.uint64) for i, q in enumerate(moduli)])
    vals = b.compose(r)
    for k, v in enumerate(vals):
        assert 0 <= v < b.Q
        for i, q in enumerate(moduli):
            assert v % q == int(r[i, k])


@pytest.mark.parametrize("L,bits,log_b,l", [(2, 59, 8, 3), (2, 59, 8, 14), (3, 50, 16, 9), (4, 62, 20, 12),
                                            (2, 31, 1, 61), (3, 40, 7, 2),
                                            (2, 59, 2, 59), (2, 31, 1, 62)])  # drop_bits == 0
def test_rns_digits_reconstruct(L, bits, log_b, l):
    moduli = _moduli(L, bits)
    b = O.RnsBasis(moduli, log_b, l)
    r = np.stack([_rng(7 * L + i).integers(0, q, size=128, dtype=np.uint64) for i, q in enumerate(moduli)])
    r[:, 0] = 0
    r[:, 1] = [q - 1 for q in moduli]
    digits = b.decompose(r)
    half = (b.basis + 1) // 2
    for k, value in enumerate(b.compose(r)):
        total = 0
        for d in range(l):
            dig = int(digits[d, 0, k])
            dig = dig if (b.basis == 2 or dig < half) else dig - (moduli[0] - b.basis)
            signed = dig - b.basis if (b.basis != 2 and dig >= half) else dig
            for i, q in enumerate(moduli):
                assert int(digits[d, i, k]) % q == signed % q
            total += signed << (b.drop_bits + d * log_b)
        err = (total - value) % b.Q
        err = err - b.Q if err > b.Q // 2 else err
        assert abs(err) <= (1 << b.drop_bits), (k, err, b.drop_bits)


def test_rns_ext_product_one_limb_equals_c_oracle():
    log_n, n, l = 10, 1024, 3
    rng = _rng(31)
    key = rng.integers(0, EXT_Q, size=(2, l, 2, 1, n), dtype=np.uint64)
    x = rng.integers(0, EXT_Q, size=(2, 1, n), dtype=np.uint64)
    t = O.Table(log_n, EXT_Q)
    got = O.ext_product_rns([t], O.RnsBasis([EXT_Q], 8, l), key, x)
    want = O.ext_product(t, O.Basis.new(EXT_Q, 8, l), np.ascontiguousarray(key[:, :, :, 0, :]),
                         np.ascontiguousarray(x[:, 0, :]))
    np.testing.assert_array_equal(got[:, 0, :], want)


@pytest.mark.gpu
@pytest.mark.parametrize("L,bits,log_b,l,log_n,count", [
    (1, 32, 8, 3, 10, 3),     # one limb: the 32-bit external-product modulus
    (2, 59, 8, 3, 10, 3),     # approximate digits, top 24 bits
    (2, 59, 8, 14, 10, 2),    # every digit
    (3, 50, 16, 9, 11, 2),    # binary64 tables
    (4, 62, 20, 12, 10, 2),   # four words
    (2, 31, 1, 61, 8, 2),     # binary digits
    (2, 59, 8, 3, 14, 1),     # two-pass transforms
    (2, 59, 2, 59, 8, 1),     # Q of 118 bits: every bit a digit bit (drop_bits 0, split + 1)
    (2, 31, 1, 62, 8, 1),     # binary digits, drop_bits 0 (no split value)
    # the fused kernel (digits within one 64-bit window) in each arithmetic
    # and word count, at the smallest and largest ring degrees it serves
    (2, 62, 8, 3, 10, 2),     # integer Shoup, limbs >= 2^60
    (4, 62, 20, 3, 12, 2),    # four words, window of 61 bits
    (3, 50, 16, 3, 11, 3),    # binary64, three words
    (2, 50, 8, 3, 6, 5),      # N = 2^6
    (2, 59, 8, 3, 12, 9),     # approximate-quotient integer path, N = 2^12
    (1, 59, 20, 2, 9, 4),     # one limb, one word
])
def test_gpu_ext_product_rns(L, bits, log_b, l, log_n, count):
    from primus_amd.fhe import ExtKeyRns
    n = 1 << log_n
    moduli = [EXT_Q] if bits == 32 else pa.ntt_primes_below(bits, 2 * n, L)
    print("ext rns", L, bits, log_b, l, log_n, count, flush=True)
    table = pa.UintCrtNttTable.new(log_n, moduli)
    rng = _rng(1000 * L + l)
    qcol = np.array(moduli, dtype=np.uint64).reshape(-1, 1)
    key = np.stack([rng.integers(0, 1 << 63, size=(L, n), dtype=np.uint64) % qcol for _ in range(4 * l)])
    key = key.reshape(2, l, 2, L, n)
    x = np.stack([rng.integers(0, 1 << 63, size=(L, n), dtype=np.uint64) % qcol for _ in range(2 * count)])
    x = x.reshape(count, 2, L, n)
    x[0, 0, :, 0] = 0
    x[0, 1, :, 1] = qcol[:, 0] - 1
    k = ExtKeyRns(table, log_b, l, key)
    din, dout = pa.DeviceBuffer.from_array(x), pa.DeviceBuffer(x.size)
    tables = [O.Table(log_n, q) for q in moduli]
    basis = O.RnsBasis(moduli, log_b, l)
    for ntt_output in (False, True):
        k.product_device(din.ptr, dout.ptr, count, ntt_output=ntt_output)
        pa.synchronize()
        got = dout.download((count, 2, L, n))
        for j in range(count):
            want = O.ext_product_rns(tables, basis, key, x[j], ntt_output=ntt_output)
            np.testing.assert_array_equal(got[j], want, err_msg="product %d ntt_output %s" % (j, ntt_output))
    np.testing.assert_array_equal(din.download((count, 2, L, n)), x)


def _edge_values(b, n, rng):
    """n composed values below Q: the ends of [0, Q), the split point and its
    neighbours, the first limb's modulus (where Garner's first mixed-radix
    digit wraps), the boundaries of every digit's bit window and of the
    rounding carry, then uniform values."""
    Q = b.Q
    vals = [0, 1, 2, Q - 1, Q - 2, b.moduli[0] - 1, b.moduli[0], b.moduli[0] + 1, Q // 2, Q // 2 + 1]
    if len(b.moduli) > 2:
        vals += [b.moduli[0] * b.moduli[1] - 1, b.moduli[0] * b.moduli[1]]
    if b.split is not None:
        vals += [b.split - 1, b.split, b.split + 1]
    for d in range(b.l + 1):
        edge = 1 << (b.drop_bits + d * b.log_basis)
        half = edge >> 1
        vals += [edge - 1, edge, edge + 1, edge + half - 1, edge + half, Q - edge, Q - half]
    vals = [v % Q for v in vals]
    while len(vals) < n:
        vals.append(int.from_bytes(rng.bytes(40), "little") % Q)
    return vals[:n]


@pytest.mark.gpu
@pytest.mark.parametrize("L,bits,log_b,l,log_n", [(2, 59, 8, 3, 10), (2, 50, 8, 3, 10), (3, 50, 16, 3, 11),
                                                  (4, 62, 20, 3, 12), (3, 59, 7, 9, 9), (2, 59, 2, 29, 8)])
def test_gpu_ext_product_rns_composition_edges(L, bits, log_b, l, log_n):
    """The fused kernel's composition (Garner's mixed-radix form, then the
    split adjustment and the digit window) on inputs whose composed values sit
    on every boundary they can cross, against the oracle's limb-by-limb sum
    (compose_inplace, primus_rns/src/base.rs:396-420)."""
    from primus_amd.fhe import ExtKeyRns
    n = 1 << log_n
    moduli = pa.ntt_primes_below(bits, 2 * n, L)
    print("ext rns edges", L, bits, log_b, l, log_n, flush=True)
    basis = O.RnsBasis(moduli, log_b, l)
    rng = _rng(77 * L + bits)
    x = np.zeros((1, 2, L, n), dtype=np.uint64)
    for p in range(2):
        vals = _edge_values(basis, n, rng)
        if p:
            vals = vals[::-1]
        for i, q in enumerate(moduli):
            x[0, p, i] = [v % q for v in vals]
    assert basis.compose(x[0, 0])[:3] == [0, 1, 2]
    qcol = np.array(moduli, dtype=np.uint64).reshape(-1, 1)
    key = np.stack([rng.integers(0, 1 << 63, size=(L, n), dtype=np.uint64) % qcol for _ in range(4 * l)])
    key = key.reshape(2, l, 2, L, n)
    k = ExtKeyRns(pa.UintCrtNttTable.new(log_n, moduli), log_b, l, key)
    din, dout = pa.DeviceBuffer.from_array(x), pa.DeviceBuffer(x.size)
    tables = [O.Table(log_n, q) for q in moduli]
    for ntt_output in (True, False):
        k.product_device(din.ptr, dout.ptr, 1, ntt_output=ntt_output)
        pa.synchronize()
        want = O.ext_product_rns(tables, basis, key, x[0], ntt_output=ntt_output)
        np.testing.assert_array_equal(dout.download((1, 2, L, n))[0], want, err_msg="ntt_output %s" % ntt_output)


@pytest.mark.gpu
def test_gpu_ext_key_rns_rejects():
    from primus_amd.fhe import ExtKeyRns
    moduli = pa.ntt_primes_below(59, 2048, 5)
    n = 1024
    with pytest.raises(Exception):
        ExtKeyRns(pa.UintCrtNttTable.new(10, moduli), 8, 3, np.zeros((2, 3, 2, 5, n), dtype=np.uint64))
    t = pa.UintCrtNttTable.new(10, moduli[:2])
    with pytest.raises(ValueError):  # 118-bit Q holds at most 14 digits of 8 bits
        ExtKeyRns(t, 8, 15, np.zeros((2, 15, 2, 2, n), dtype=np.uint64))
    bad = np.zeros((2, 3, 2, 2, n), dtype=np.uint64)
    bad[1, 2, 1, 1, 5] = moduli[1]
    with pytest.raises(ValueError):
        ExtKeyRns(t, 8, 3, bad)
    with pytest.raises(ValueError):  # a short key array is refused before the library reads it
        ExtKeyRns(t, 8, 3, np.zeros((2, 3, 2, 2, n - 1), dtype=np.uint64))
    # 2^40 exceeds the 31-bit limb moduli: digits >= 2^39 have no residue q_i - B + d
    small = pa.ntt_primes_below(31, 2048, 2)
    with pytest.raises(ValueError):
        ExtKeyRns(pa.UintCrtNttTable.new(10, small), 40, 1, np.zeros((2, 1, 2, 2, n), dtype=np.uint64))


def test_rns_basis_rejects_basis_above_a_modulus():
    small = pa.ntt_primes_below(31, 2048, 2)
    with pytest.raises(ValueError):
        O.RnsBasis(small, 40, 1)
    O.RnsBasis(small, 30, 2)  # 2^30 is below both moduli
