"""Pins the CPU oracle (oracle/primus_oracle.c) before anything is compared to it.

Checks come from outside the oracle: the golden vectors of
tests/golden/make_golden.py (Python big-int evaluation, brute-force roots,
schoolbook products), closed forms (Shoup/Barrett/MultiplyFactor factors),
Python's exact `%`, and invariants (round trips, lazy ranges, congruences).
"""
import random

import numpy as np
import pytest

from oracle import oracle as O

MASK64 = (1 << 64) - 1


def rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


# ---------------------------------------------------------------- golden vectors

def test_golden_minimal_root(golden):
    for label, c in golden.items():
        q, log_n = int(c["q"][0]), int(c["log_n"][0])
        assert O.minimal_primitive_root(log_n + 1, q) == int(c["psi"][0]), label
        assert O.Table(log_n, q).root == int(c["psi"][0]), label


def test_golden_forward_inverse(golden):
    for label, c in golden.items():
        q, log_n = int(c["q"][0]), int(c["log_n"][0])
        t = O.Table(log_n, q)
        np.testing.assert_array_equal(t.transform(c["a"]), c["fwd_a"], err_msg=label)
        np.testing.assert_array_equal(t.inverse_transform(c["fwd_a"]), c["a"], err_msg=label)
        if "inv_in" in c:
            np.testing.assert_array_equal(t.inverse_transform(c["inv_in"]), c["inv_out"], err_msg=label)


def test_golden_negacyclic_product(golden):
    for label, c in golden.items():
        if "prod" not in c:
            continue
        q, log_n = int(c["q"][0]), int(c["log_n"][0])
        t = O.Table(log_n, q)
        fa, fb = t.transform(c["a"]), t.transform(c["b"])
        got = t.inverse_transform(O.pointwise_mul(fa, fb, q))
        np.testing.assert_array_equal(got, c["prod"], err_msg=label)
        if (1 << log_n) <= 1024:
            np.testing.assert_array_equal(O.schoolbook(c["a"], c["b"], q), c["prod"], err_msg=label)


def test_golden_monomials(golden):
    for label, c in golden.items():
        q, log_n = int(c["q"][0]), int(c["log_n"][0])
        t = O.Table(log_n, q)
        coeff, deg = (int(x) for x in c["mono_params"])
        np.testing.assert_array_equal(t.transform_monomial(coeff, deg), c["mono"], err_msg=label)
        np.testing.assert_array_equal(t.transform_coeff_one_monomial(deg), c["mono_one"], err_msg=label)
        np.testing.assert_array_equal(t.transform_coeff_minus_one_monomial(deg), c["mono_minus_one"],
                                      err_msg=label)
        # the monomial closed form agrees with transforming the monomial itself
        x = np.zeros(1 << log_n, dtype=np.uint64)
        x[deg] = coeff
        np.testing.assert_array_equal(t.transform(x), c["mono"], err_msg=label)


# ---------------------------------------------------------------- lazy ranges

@pytest.mark.parametrize("q,log_n", [(2147473409, 10), (1125899906826241, 8), (1152921504606830593, 9),
                                     (4611686018427322369, 7)])
def test_lazy_ranges_and_congruence(q, log_n):
    n = 1 << log_n
    t = O.Table(log_n, q)
    rng = np.random.default_rng(q % 1000)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    canon = t.transform(a)
    # inputs anywhere in [0, 4q) give the same canonical transform
    shift = rng.integers(0, 4, size=n, dtype=np.uint64)
    a4 = (a + shift * np.uint64(q)).astype(np.uint64)
    assert int(a4.max()) < 4 * q
    lz = t.lazy_transform(a4)
    assert int(lz.max()) < 4 * q
    np.testing.assert_array_equal(lz % np.uint64(q), canon)
    np.testing.assert_array_equal(t.transform(a4), canon)
    # lazy inverse from [0, 2q): output fully reduced (the Uint variant folds N^-1 with full reductions)
    b2 = (canon + rng.integers(0, 2, size=n, dtype=np.uint64) * np.uint64(q)).astype(np.uint64)
    li = t.lazy_inverse_transform(b2)
    assert int(li.max()) < q
    np.testing.assert_array_equal(li, a)


def test_extreme_inputs():
    q, log_n = 1152921504606830593, 6
    n = 1 << log_n
    t = O.Table(log_n, q)
    top = np.full(n, 4 * q - 1, dtype=np.uint64)
    np.testing.assert_array_equal(t.transform(top), t.transform(np.full(n, q - 1, dtype=np.uint64)))
    z = np.zeros(n, dtype=np.uint64)
    np.testing.assert_array_equal(t.transform(z), z)
    np.testing.assert_array_equal(t.inverse_transform(z), z)
    d = z.copy()
    d[0] = 1
    np.testing.assert_array_equal(t.transform(d), np.ones(n, dtype=np.uint64))


# ---------------------------------------------------------------- table layout

@pytest.mark.parametrize("q,log_n", [(12289, 9), (1125899906826241, 6)])
def test_table_layout(q, log_n):
    n = 1 << log_n
    t = O.Table(log_n, q)
    psi = t.root
    rp, rpq = t.powers(0)
    irp, irpq = t.powers(1)
    op, opq = t.powers(2)
    for k in range(2 * n):
        assert int(op[k]) == pow(psi, k, q)
        assert int(opq[k]) == (int(op[k]) << 64) // q
    for k in range(n):
        assert int(rp[rev(k, log_n)]) == pow(psi, k, q)
        assert int(rpq[rev(k, log_n)]) == (pow(psi, k, q) << 64) // q
    assert int(irp[0]) == 1
    for k in range(n - 1):
        assert int(irp[rev(k, log_n) + 1]) == pow(psi, -(k + 1), q)
    inv_n, inv_n_q = t.inv_n
    assert inv_n * n % q == 1 and inv_n_q == (inv_n << 64) // q
    assert t.inv_root * psi % q == 1


def test_table_errors():
    with pytest.raises(O.OracleError) as e:
        O.Table(10, 12289 + 2)  # 12291 = 3 * 17 * 241: 2N does not divide q-1
    assert e.value.code == 1
    with pytest.raises(O.OracleError) as e:
        O.Table(4, 1)
    assert e.value.code == 5
    with pytest.raises(O.OracleError) as e:
        O.Table(4, (1 << 62) + 1)  # BarrettModulus::new rejects bit length >= 63
    assert e.value.code == 5
    with pytest.raises(O.OracleError) as e:
        O.Table(12, 12289)  # 8192 does not divide 12288
    assert e.value.code == 1


def test_minimal_root_bruteforce():
    for q, log_n in [(17, 3), (97, 4), (257, 7), (7681, 8), (12289, 10)]:
        n = 1 << log_n
        roots = [x for x in range(1, q) if pow(x, n, q) == q - 1]
        assert O.minimal_primitive_root(log_n + 1, q) == min(roots)
        assert O.lib().po_is_primitive_root(min(roots), log_n + 1, q) == 1
        assert O.lib().po_is_primitive_root(1, log_n + 1, q) == 0


def test_reverse_lsbs():
    L = O.lib()
    assert L.po_reverse_lsbs(0b1101, 4) == 0b1011  # primus_ntt/src/reverse.rs:122
    assert L.po_reverse_lsbs(0b01101101, 4) == 0b1011  # reverse.rs:123
    assert L.po_reverse_lsbs(0, 5) == 0 and L.po_reverse_lsbs(7, 0) == 0
    for x in range(64):
        assert L.po_reverse_lsbs(x, 6) == rev(x, 6)


# ---------------------------------------------------------------- scalar arithmetic

PRIMES = [3, 17, 12289, 536813569, 2147473409, 4294957057, 562949953392641, 1125899906826241,
          576460752303421441, 1152921504606830593, 4611686018427322369]


def _samples(q, rng, k=300):
    edge = [0, 1, q - 1, q, 2 * q - 1, 4 * q - 1, MASK64]
    return [x for x in edge if x <= MASK64] + [rng.randrange(1 << 64) for _ in range(k)]


def test_shoup_against_exact():
    L = O.lib()
    rng = random.Random(1)
    for q in PRIMES:
        for w in [0, 1, q - 1] + [rng.randrange(q) for _ in range(20)]:
            wq = L.po_shoup_quotient(w, q)
            assert wq == (w << 64) // q  # ShoupFactor::new closed form
            for b in _samples(q, rng, 50):
                lz = L.po_shoup_lazy_mul(w, wq, b, q)
                assert lz < 2 * q and lz % q == w * b % q
                assert L.po_shoup_mul(w, wq, b, q) == w * b % q


def test_barrett_against_exact():
    L = O.lib()
    rng = random.Random(2)
    for q in PRIMES + [6, 1 << 40, (1 << 62) - 1]:
        r = O.barrett_ratio(q)
        assert int(r[0]) + (int(r[1]) << 64) == (1 << 128) // q  # mod.rs:68-75 closed form
        for _ in range(300):
            a, b = rng.randrange(q), rng.randrange(q)
            c = rng.randrange(q)
            assert L.po_barrett_reduce_mul(a, b, q, O.ptr(r)) == a * b % q
            assert L.po_barrett_reduce_mul_add(a, b, c, q, O.ptr(r)) == (a * b + c) % q
            v = rng.randrange(1 << 64)
            lz = L.po_barrett_lazy_reduce(v, q, O.ptr(r))
            assert lz < 2 * q and lz % q == v % q
            lo, hi = rng.randrange(1 << 64), rng.randrange(q)
            lz2 = L.po_barrett_lazy_reduce_2limb(lo, hi, q, O.ptr(r))
            assert lz2 < 2 * q and lz2 % q == ((hi << 64) + lo) % q
        for a, b in [(0, 0), (q - 1, q - 1), (q - 1, 1), (1, 0)]:
            assert L.po_barrett_reduce_mul(a, b, q, O.ptr(r)) == a * b % q


def test_barrett_rejects():
    for q in [0, 1, 1 << 62, (1 << 63) + 1]:
        with pytest.raises(O.OracleError):
            O.barrett_ratio(q)


def test_multiply_factor():
    """MultiplyFactor (primus_factor/src/multiply.rs:84-135 checks it against u128 %)."""
    L = O.lib()
    rng = random.Random(3)
    for shift, qbits in [(32, 31), (52, 50), (64, 62)]:
        for q in [p for p in PRIMES if p.bit_length() <= qbits][-3:]:
            for _ in range(200):
                op = rng.randrange(q)
                f = L.po_multiply_factor(op, shift, q)
                assert f == (op << shift) // q
                x = rng.randrange(min(q, 1 << shift))
                r = L.po_multiply_mod_lazy(x, op, f, shift, q)
                assert r < 2 * q and r % q == x * op % q


def test_pow_inv():
    L = O.lib()
    rng = random.Random(4)
    for q in PRIMES[2:]:
        for _ in range(50):
            a, e = rng.randrange(1, q), rng.randrange(1 << 64)
            assert L.po_pow_mod(a, e, q) == pow(a, e, q)
            assert L.po_inv_mod(a, q) == pow(a, -1, q)
    assert L.po_inv_mod(6, 9) == 0


def test_batch_matches_single():
    q, log_n = 1125899906826241, 7
    n = 1 << log_n
    tables = [O.Table(log_n, q), O.Table(log_n, 1125899906629633)]
    rng = np.random.default_rng(5)
    polys = rng.integers(0, 1125899906629633, size=(6, n), dtype=np.uint64)
    op = np.stack([rng.integers(0, t.q, size=n, dtype=np.uint64) for t in tables])
    opq = np.stack([O.shoup_quotients(op[i], tables[i].q) for i in range(2)])
    want = polys.copy()
    for i in range(6):
        t = tables[i % 2]
        f = t.transform(want[i])
        want[i] = t.inverse_transform(O.pointwise_mul(f, op[i % 2], t.q))
    got = O.fwd_mul_inv_batch(tables, polys.copy(), op, opq, threads=3)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("q,log_n", [(17, 3), (12289, 9), (1125899906826241, 12), (576460752303421441, 10)])
def test_ntt_automorphism_pinned_by_transform(q, log_n):
    """The NTT-domain permutation (dcrt.rs:43-61) is pinned through the transform:
    NTT(a(X^k) mod X^N+1) == permuted NTT(a) for odd k, including k = 1, 2N-1
    (conjugation) and k = N+1."""
    n = 1 << log_n
    t = O.Table(log_n, q)
    rng = np.random.default_rng(log_n)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    fa = t.transform(a)
    for k in sorted({1, 3, 5, 2 * n - 1, n + 1, 2 * n + 5, int(rng.integers(0, n)) * 2 + 1}):
        np.testing.assert_array_equal(O.ntt_automorphism(fa, k), t.transform(O.coeff_automorphism(a, k, q)),
                                      err_msg="k=%d" % k)
    # composition: sigma_k1 after sigma_k2 is sigma_(k1*k2)
    np.testing.assert_array_equal(O.ntt_automorphism(O.ntt_automorphism(fa, 3), 5),
                                  O.ntt_automorphism(fa, 15 % (2 * n)))


def test_coeff_automorphism_brute_force():
    """a(X^k) mod (X^N + 1) by expanding every monomial with Python integers."""
    q, n = 97, 8
    rng = np.random.default_rng(5)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    for k in (1, 3, 5, 7, 9, 15):
        want = [0] * n
        for j, c in enumerate(a.tolist()):
            e = j * k
            sign = -1 if (e // n) % 2 else 1  # X^N = -1
            want[e % n] = (want[e % n] + sign * c) % q
        np.testing.assert_array_equal(O.coeff_automorphism(a, k, q), np.array(want, dtype=np.uint64))


def _crt_value(residues, moduli):
    Q = 1
    for q in moduli:
        Q *= int(q)
    x = 0
    for r, q in zip(residues, moduli):
        qi = Q // int(q)
        x += int(r) * qi * pow(qi, -1, int(q))
    return x % Q, Q


@pytest.mark.parametrize("ibase,obase", [
    ([31, 37, 41, 43], [47, 53, 59, 61]),  # primus_rns/src/converter.rs:331-332
    ([576460752303421441, 576460752303419393, 576460752303415297], [1125899906826241, 1125899906629633]),
    ([2147473409, 4294957057, 1152921504606830593, 12289, 17], [576460752303421441]),
])
def test_fast_base_conversion_pinned(ibase, obase):
    """fast_convert_array against big integers: with y_i = [x_i (Q/q_i)^-1]_{q_i},
    S = sum_i y_i Q/q_i is congruent to the CRT value X mod Q, 0 <= S - X < len(ibase) Q,
    and every output residue is S mod p_j."""
    c = O.BaseConverter(ibase, obase)
    rng = np.random.default_rng(len(ibase))
    n = 64
    x = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in ibase])
    out = c.fast_convert(x)
    for k in range(n):
        X, Q = _crt_value(x[:, k], ibase)
        S = 0
        for i, q in enumerate(ibase):
            qi = Q // q
            S += (int(x[i, k]) * pow(qi, -1, q) % q) * qi
        assert S % Q == X and 0 <= S - X < len(ibase) * Q
        for j, p in enumerate(obase):
            assert int(out[j, k]) == S % p


@pytest.mark.parametrize("ibase,p", [
    ([31, 37, 41, 43], 47),
    ([576460752303421441, 576460752303419393, 576460752303415297, 576460752303403009], 1125899906826241),
    ([1125899906826241, 1125899906629633], 4294957057),
])
def test_exact_base_conversion_pinned(ibase, p):
    """exact_convert_array equals the centred CRT value (X - Q when X > Q/2: v
    rounds S/Q to nearest) reduced mod p, for inputs whose fraction X/Q is not
    within rounding of 1/2 (all random ones)."""
    c = O.BaseConverter(ibase, [p])
    rng = np.random.default_rng(p % 1000)
    n = 256
    x = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in ibase])
    x[:, 0] = 0  # zero
    x[:, 1] = [q - 1 for q in ibase]  # Q - 1
    out = c.exact_convert(x)[0]
    for k in range(n):
        X, Q = _crt_value(x[:, k], ibase)
        centred = X - Q if 2 * X > Q else X
        assert int(out[k]) == centred % p


def test_base_converter_rejects():
    with pytest.raises(O.OracleError):
        O.BaseConverter([6, 9], [7])  # not coprime
