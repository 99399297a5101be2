"""HexlNttTable boundary (primus_ntt/src/ntt/hexl/mod.rs:447-635) on the GPU
tables: the checks of the reference's own primus_ntt/tests/hexl.rs -- the
HEXL table's full-range forward, inverse and monomials equal UintNttTable's,
at its three primes and N = 2048 -- against the oracle's restatement of
UintNttTable, plus HEXL's input and output range factors (inputs in
[0, factor*q); lazy outputs in their range and congruent to the canonical
ones), and HEXL's own lazy values: bit for bit those of the oracle's
restatement of its radix-2 path (hexl/radix2.rs).

The oracle's radix-2 restatement is pinned on the CPU (no GPU marker): its
full-range outputs equal the UintNttTable restatement's, which the golden
vectors and brute-force evaluation pin (test_oracle.py), as the reference's
tests/hexl.rs asserts; its lazy outputs lie in [0, 4q) / [0, 2q) and are
congruent to them; its lazy forward equals UintNttTable's lazy forward (the
same butterflies, hexl/radix2.rs:40-50 and primitive.rs:225-228); and its lazy
inverse leaves the last stage's products unreduced, so at 60 bits it differs
from the canonical inverse on some of 2048 random inputs."""
import numpy as np
import pytest

import primus_amd as pa
from oracle import oracle as O

LOG_N = 11
N = 1 << LOG_N
# test_bit_shift_{32,52,64} of primus_ntt/tests/hexl.rs
PRIMES = [536813569, 562949953392641, 1152921504606830593]


@pytest.mark.parametrize("log_n", [4, 11])
@pytest.mark.parametrize("q", PRIMES)
def test_oracle_hexl_radix2_pinned(q, log_n):
    n = 1 << log_n
    rng = np.random.default_rng(q % 977 + log_n)
    ot = O.Table(log_n, q)
    x = rng.integers(0, 4 * q, size=n, dtype=np.uint64)
    x[:3] = [0, q - 1, 4 * q - 1]
    np.testing.assert_array_equal(ot.hexl_forward(x, 1), ot.transform(x))
    lazy = ot.hexl_forward(x, 4)
    assert int(lazy.max()) < 4 * q
    np.testing.assert_array_equal(lazy % np.uint64(q), ot.transform(x))
    np.testing.assert_array_equal(lazy, ot.lazy_transform(x))
    y = rng.integers(0, 2 * q, size=n, dtype=np.uint64)
    y[:2] = [0, 2 * q - 1]
    canon = ot.inverse_transform(y)
    np.testing.assert_array_equal(ot.hexl_inverse(y, 1), canon)
    ilazy = ot.hexl_inverse(y, 2)
    assert int(ilazy.max()) < 2 * q
    np.testing.assert_array_equal(ilazy % np.uint64(q), canon)
    if q > 1 << 59 and log_n == 11:
        # the lazy products' quotient is one short often enough at 60 bits
        # for HEXL's representatives to differ from the canonical ones
        assert (ilazy != canon).any()


@pytest.mark.gpu
@pytest.mark.parametrize("q,log_n", [(q, ln) for q in PRIMES for ln in (4, 11)] +
                         [(1152921504606584833, 14), (1152921504606584833, 16)])
def test_hexl_lazy_values_bit_exact(q, log_n):
    """compute_forward(.., 4) and compute_inverse(.., 2) give HEXL's own lazy
    values, single-pass and two-pass ring degrees alike."""
    n = 1 << log_n
    if (q - 1) % (2 * n):
        pytest.skip("q - 1 not divisible by 2N")
    print("hexl lazy", q, log_n, flush=True)
    rng = np.random.default_rng(q % 991 + log_n)
    h = pa.HexlNttTable.new(log_n, q)
    ot = O.Table(log_n, q)
    for fin in (1, 2, 4):
        x = (rng.integers(0, q, size=n, dtype=np.uint64) +
             rng.integers(0, fin, size=n, dtype=np.uint64) * np.uint64(q)).astype(np.uint64)
        z = x.copy()
        h.compute_forward(z, fin, 4)
        np.testing.assert_array_equal(z, ot.hexl_forward(x, 4), err_msg="forward in %d" % fin)
    for fin in (1, 2):
        y = (rng.integers(0, q, size=n, dtype=np.uint64) +
             rng.integers(0, fin, size=n, dtype=np.uint64) * np.uint64(q)).astype(np.uint64)
        w = y.copy()
        h.compute_inverse(w, fin, 2)
        np.testing.assert_array_equal(w, ot.hexl_inverse(y, 2), err_msg="inverse in %d" % fin)


@pytest.mark.gpu
@pytest.mark.parametrize("q", PRIMES)
def test_hexl_full_range_equals_uint(q):
    print("hexl", q, flush=True)
    rng = np.random.default_rng(q % 1000)
    h = pa.HexlNttTable.new(LOG_N, q)
    ot = O.Table(LOG_N, q)
    poly = rng.integers(0, q, size=N, dtype=np.uint64)
    x = poly.copy()
    h.compute_forward(x, 1, 1)
    fwd = ot.transform(poly)
    np.testing.assert_array_equal(x, fwd)
    h.compute_inverse(x, 1, 1)
    np.testing.assert_array_equal(x, ot.inverse_transform(fwd))
    np.testing.assert_array_equal(x, poly)
    # the NttTable methods route through compute_* with HEXL's factors
    y = poly.copy()
    h.transform_slice(y)
    np.testing.assert_array_equal(y, fwd)
    h.inverse_transform_slice(y)
    np.testing.assert_array_equal(y, poly)
    for degree in (0, 1, int(rng.integers(0, N)), N - 1):
        coeff = int(rng.integers(0, q))
        m = np.zeros(N, dtype=np.uint64)
        h.transform_monomial(coeff, degree, m)
        np.testing.assert_array_equal(m, ot.transform_monomial(coeff, degree))
        h.transform_coeff_one_monomial(degree, m)
        np.testing.assert_array_equal(m, ot.transform_coeff_one_monomial(degree))
        h.transform_coeff_minus_one_monomial(degree, m)
        np.testing.assert_array_equal(m, ot.transform_coeff_minus_one_monomial(degree))


@pytest.mark.gpu
@pytest.mark.parametrize("q", PRIMES)
def test_hexl_range_factors(q):
    rng = np.random.default_rng(7 + q % 1000)
    h = pa.HexlNttTable.new(LOG_N, q)
    ot = O.Table(LOG_N, q)
    base = rng.integers(0, q, size=N, dtype=np.uint64)
    fwd = ot.transform(base)
    for fin in (1, 2, 4):
        x = (base + rng.integers(0, fin, size=N, dtype=np.uint64) * np.uint64(q)).astype(np.uint64)
        assert int(x.max()) < fin * q
        y = x.copy()
        h.compute_forward(y, fin, 1)
        np.testing.assert_array_equal(y, fwd, err_msg="in %d out 1" % fin)
        z = x.copy()
        h.compute_forward(z, fin, 4)
        assert int(z.max()) < 4 * q
        np.testing.assert_array_equal(z % np.uint64(q), fwd, err_msg="in %d out 4" % fin)
    for fin in (1, 2):
        v = (fwd + rng.integers(0, fin, size=N, dtype=np.uint64) * np.uint64(q)).astype(np.uint64)
        for fout in (1, 2):
            w = v.copy()
            h.compute_inverse(w, fin, fout)
            assert int(w.max()) < fout * q
            np.testing.assert_array_equal(w % np.uint64(q), base, err_msg="inv in %d out %d" % (fin, fout))


@pytest.mark.gpu
def test_hexl_factors_rejected():
    """Factors the reference debug_asserts against are errors (the binding's
    ValueError for PH_ERR_INVALID), and the operand is left as it was."""
    h = pa.HexlNttTable.new(LOG_N, PRIMES[0])
    x = np.arange(N, dtype=np.uint64)
    for fin, fout in ((3, 1), (1, 2), (8, 4), (0, 1)):
        with pytest.raises(ValueError, match="mod_factor"):
            h.compute_forward(x, fin, fout)
    for fin, fout in ((4, 1), (1, 4), (0, 1)):
        with pytest.raises(ValueError, match="mod_factor"):
            h.compute_inverse(x, fin, fout)
    np.testing.assert_array_equal(x, np.arange(N, dtype=np.uint64))
