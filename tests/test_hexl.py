"""HexlNttTable boundary (primus_ntt/src/ntt/hexl/mod.rs:447-635) on the GPU
tables: the checks of the reference's own primus_ntt/tests/hexl.rs -- the
HEXL table's full-range forward, inverse and monomials equal UintNttTable's,
at its three primes and N = 2048 -- against the oracle's restatement of
UintNttTable, plus HEXL's input and output range factors (inputs in
[0, factor*q); lazy outputs in their range and congruent to the canonical
ones)."""
import numpy as np
import pytest

import primus_amd as pa
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LOG_N = 11
N = 1 << LOG_N
# test_bit_shift_{32,52,64} of primus_ntt/tests/hexl.rs
PRIMES = [536813569, 562949953392641, 1152921504606830593]


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
