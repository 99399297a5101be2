"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as the product path.  The C sources restate the reference's
UintNttTable<u64> (primus_ntt/src/ntt/primitive.rs) and the Shoup/Barrett
arithmetic under it; see oracle/primus_oracle.c for the citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u64 = ctypes.c_uint64
_p64 = ctypes.POINTER(ctypes.c_uint64)


def build():
    """Compile oracle/liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name, res, args in [
            ("po_shoup_quotient", _u64, [_u64, _u64]),
            ("po_shoup_lazy_mul", _u64, [_u64, _u64, _u64, _u64]),
            ("po_shoup_mul", _u64, [_u64, _u64, _u64, _u64]),
            ("po_multiply_factor", _u64, [_u64, ctypes.c_uint32, _u64]),
            ("po_multiply_mod_lazy", _u64, [_u64, _u64, _u64, ctypes.c_uint32, _u64]),
            ("po_barrett_ratio", ctypes.c_int, [_u64, _p64]),
            ("po_barrett_lazy_reduce", _u64, [_u64, _u64, _p64]),
            ("po_barrett_lazy_reduce_2limb", _u64, [_u64, _u64, _u64, _p64]),
            ("po_barrett_reduce_mul", _u64, [_u64, _u64, _u64, _p64]),
            ("po_barrett_reduce_mul_add", _u64, [_u64, _u64, _u64, _u64, _p64]),
            ("po_pow_mod", _u64, [_u64, _u64, _u64]),
            ("po_inv_mod", _u64, [_u64, _u64]),
            ("po_is_primitive_root", ctypes.c_int, [_u64, ctypes.c_uint32, _u64]),
            ("po_minimal_primitive_root", ctypes.c_int, [ctypes.c_uint32, _u64, _p64]),
            ("po_reverse_lsbs", _u64, [_u64, ctypes.c_uint32]),
            ("po_table_new", ctypes.c_int, [ctypes.c_uint32, _u64, ctypes.POINTER(ctypes.c_void_p)]),
            ("po_table_free", None, [ctypes.c_void_p]),
            ("po_table_n", ctypes.c_size_t, [ctypes.c_void_p]),
            ("po_table_modulus", _u64, [ctypes.c_void_p]),
            ("po_table_root", _u64, [ctypes.c_void_p]),
            ("po_table_inv_root", _u64, [ctypes.c_void_p]),
            ("po_table_inv_n", None, [ctypes.c_void_p, _p64, _p64]),
            ("po_table_powers", None, [ctypes.c_void_p, ctypes.c_int, _p64, _p64]),
            ("po_lazy_transform", None, [ctypes.c_void_p, _p64]),
            ("po_transform", None, [ctypes.c_void_p, _p64]),
            ("po_lazy_inverse_transform", None, [ctypes.c_void_p, _p64]),
            ("po_hexl_forward", None, [ctypes.c_void_p, _p64, ctypes.c_uint32]),
            ("po_hexl_inverse", None, [ctypes.c_void_p, _p64, ctypes.c_uint32]),
            ("po_inverse_transform", None, [ctypes.c_void_p, _p64]),
            ("po_transform_monomial", None, [ctypes.c_void_p, _u64, ctypes.c_size_t, _p64]),
            ("po_transform_coeff_one_monomial", None, [ctypes.c_void_p, ctypes.c_size_t, _p64]),
            ("po_transform_coeff_minus_one_monomial", None, [ctypes.c_void_p, ctypes.c_size_t, _p64]),
            ("po_pointwise_mul", None, [_p64, _p64, _p64, ctypes.c_size_t, _u64]),
            ("po_pointwise_mul_add", None, [_p64, _p64, _p64, ctypes.c_size_t, _u64]),
            ("po_shoup_pointwise", None, [_p64, _p64, _p64, ctypes.c_size_t, _u64]),
            ("po_schoolbook_negacyclic", None, [_p64, _p64, _p64, ctypes.c_size_t, _u64]),
            ("po_fwd_mul_inv_batch", None,
             [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint32, _p64, ctypes.c_size_t, _p64, _p64, ctypes.c_int]),
            ("po_basis_new", ctypes.c_int, [_u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
            ("po_decompose", None, [ctypes.c_void_p, _p64, ctypes.c_size_t, _p64]),
            ("po_ext_product", None, [ctypes.c_void_p, ctypes.c_void_p, _p64, _p64, _p64, ctypes.c_int]),
            ("po_decompose_unsigned", None, [ctypes.c_void_p, _p64, ctypes.c_size_t, _p64]),
            ("po_ext_product_unsigned", None, [ctypes.c_void_p, ctypes.c_void_p, _p64, _p64, _p64, ctypes.c_int]),
            ("po_ntt_automorphism", None, [_p64, _p64, ctypes.c_size_t, _u64]),
            ("po_base_converter_new", ctypes.c_int,
             [_p64, ctypes.c_uint32, _p64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]),
            ("po_base_converter_free", None, [ctypes.c_void_p]),
            ("po_fast_convert_array", None, [ctypes.c_void_p, _p64, _p64, ctypes.c_size_t]),
            ("po_exact_convert_array", None, [ctypes.c_void_p, _p64, _p64, ctypes.c_size_t]),
            ("po_coeff_automorphism", None, [_p64, _p64, ctypes.c_size_t, _u64, _u64]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_p64)


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__("oracle status %d" % code)
        self.code = code


def barrett_ratio(q):
    r = np.zeros(2, dtype=np.uint64)
    st = lib().po_barrett_ratio(q, ptr(r))
    if st:
        raise OracleError(st)
    return r


def minimal_primitive_root(log_degree, q):
    out = _u64(0)
    st = lib().po_minimal_primitive_root(log_degree, q, ctypes.byref(out))
    if st:
        raise OracleError(st)
    return out.value


class Table:
    """UintNttTable<u64> restated (primus_ntt/src/ntt/primitive.rs:34-398)."""

    def __init__(self, log_n, q):
        h = ctypes.c_void_p()
        st = lib().po_table_new(log_n, q, ctypes.byref(h))
        if st:
            raise OracleError(st)
        self._h = h
        self.log_n = log_n
        self.n = 1 << log_n
        self.q = q

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.po_table_free(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def root(self):
        return lib().po_table_root(self._h)

    @property
    def inv_root(self):
        return lib().po_table_inv_root(self._h)

    @property
    def inv_n(self):
        v, w = _u64(0), _u64(0)
        lib().po_table_inv_n(self._h, ctypes.byref(v), ctypes.byref(w))
        return v.value, w.value

    def powers(self, which):
        cnt = 2 * self.n if which == 2 else self.n
        v = np.zeros(cnt, dtype=np.uint64)
        w = np.zeros(cnt, dtype=np.uint64)
        lib().po_table_powers(self._h, which, ptr(v), ptr(w))
        return v, w

    def _run(self, fn, x):
        y = np.ascontiguousarray(x, dtype=np.uint64).copy()
        assert y.shape[-1] == self.n
        flat = y.reshape(-1, self.n)
        for row in flat:
            fn(self._h, ptr(row))
        return y

    def transform(self, x):
        return self._run(lib().po_transform, x)

    def lazy_transform(self, x):
        return self._run(lib().po_lazy_transform, x)

    def inverse_transform(self, x):
        return self._run(lib().po_inverse_transform, x)

    def lazy_inverse_transform(self, x):
        return self._run(lib().po_lazy_inverse_transform, x)

    def hexl_forward(self, x, output_mod_factor):
        """HexlNttTable::compute_forward on its radix-2 path (hexl/radix2.rs)."""
        return self._run(lambda h, p: lib().po_hexl_forward(h, p, output_mod_factor), x)

    def hexl_inverse(self, x, output_mod_factor):
        """HexlNttTable::compute_inverse on its radix-2 path (hexl/radix2.rs)."""
        return self._run(lambda h, p: lib().po_hexl_inverse(h, p, output_mod_factor), x)

    def transform_monomial(self, coeff, degree):
        out = np.zeros(self.n, dtype=np.uint64)
        lib().po_transform_monomial(self._h, coeff, degree, ptr(out))
        return out

    def transform_coeff_one_monomial(self, degree):
        out = np.zeros(self.n, dtype=np.uint64)
        lib().po_transform_coeff_one_monomial(self._h, degree, ptr(out))
        return out

    def transform_coeff_minus_one_monomial(self, degree):
        out = np.zeros(self.n, dtype=np.uint64)
        lib().po_transform_coeff_minus_one_monomial(self._h, degree, ptr(out))
        return out


def pointwise_mul(a, b, q):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros_like(a)
    lib().po_pointwise_mul(ptr(a), ptr(b), ptr(out), a.size, q)
    return out


def pointwise_mul_add(acc, a, b, q):
    acc = np.ascontiguousarray(acc, dtype=np.uint64).copy()
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    lib().po_pointwise_mul_add(ptr(acc), ptr(a), ptr(b), a.size, q)
    return acc


def shoup_quotients(w, q):
    L = lib()
    return np.array([L.po_shoup_quotient(int(x), q) for x in np.asarray(w).ravel()],
                    dtype=np.uint64).reshape(np.shape(w))


def schoolbook(a, b, q):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    c = np.zeros_like(a)
    lib().po_schoolbook_negacyclic(ptr(a), ptr(b), ptr(c), a.size, q)
    return c


def fwd_mul_inv_batch(tables, polys, operand, operand_q, threads=1):
    """Forward NTT, Shoup pointwise multiply, inverse NTT for each limb-poly in place."""
    L = len(tables)
    arr = (ctypes.c_void_p * L)(*[t.handle for t in tables])
    n = tables[0].n
    assert polys.dtype == np.uint64 and polys.flags["C_CONTIGUOUS"]
    count = polys.size // n
    lib().po_fwd_mul_inv_batch(arr, L, ptr(polys), count, ptr(operand), ptr(operand_q), threads)
    return polys


class Basis(ctypes.Structure):
    """po_basis: BigUintApproxSignedBasis for one u64 limb
    (primus_decompose/src/big_integer/basis.rs:34-195)."""

    _fields_ = [("q", ctypes.c_uint64), ("log_basis", ctypes.c_uint32), ("l", ctypes.c_uint32),
                ("drop_bits", ctypes.c_uint32), ("modulus_bits", ctypes.c_uint32), ("basis", ctypes.c_uint64),
                ("basis_minus_one", ctypes.c_uint64), ("carry_mask", ctypes.c_uint64), ("split", ctypes.c_uint64),
                ("has_split", ctypes.c_int), ("pow2_sub_modulus", ctypes.c_uint64)]

    @classmethod
    def new(cls, q, log_basis, l):
        b = cls()
        st = lib().po_basis_new(q, log_basis, l, ctypes.byref(b))
        if st:
            raise OracleError(st)
        return b

    def decompose(self, values):
        v = np.ascontiguousarray(values, dtype=np.uint64)
        out = np.zeros((self.l, v.size), dtype=np.uint64)
        lib().po_decompose(ctypes.byref(self), ptr(v), v.size, ptr(out))
        return out

    def decompose_unsigned(self, values):
        """The literal unsigned digits (po_decompose_unsigned): no carry, no
        split, no signed mapping; [l][n] residues."""
        v = np.ascontiguousarray(values, dtype=np.uint64)
        out = np.zeros((self.l, v.size), dtype=np.uint64)
        lib().po_decompose_unsigned(ctypes.byref(self), ptr(v), v.size, ptr(out))
        return out


def ext_product(table, basis, key, rlwe, ntt_output=False, unsigned_digits=False):
    """External product of one RLWE ciphertext [2][N] with a GGSW key [2][l][2][N]
    (NTT form); unsigned_digits: the literal unsigned digits instead of the
    reference's approximate signed ones."""
    key = np.ascontiguousarray(key, dtype=np.uint64)
    x = np.ascontiguousarray(rlwe, dtype=np.uint64)
    out = np.zeros(2 * table.n, dtype=np.uint64)
    fn = lib().po_ext_product_unsigned if unsigned_digits else lib().po_ext_product
    fn(table.handle, ctypes.byref(basis), ptr(key), ptr(x), ptr(out), int(bool(ntt_output)))
    return out.reshape(2, table.n)


class RnsBasis:
    """BigUintApproxSignedBasis over Q = the product of several limb moduli
    (primus_decompose/src/big_integer/basis.rs:34-195) with the RNS base's
    composition (primus_rns/src/base.rs:356-450), on Python integers: the
    multi-limb form of Basis/po_basis (which restates one u64 limb)."""

    def __init__(self, moduli, log_basis, l):
        self.moduli = [int(q) for q in moduli]
        Q = 1
        for q in self.moduli:
            Q *= q
        self.Q = Q
        bits = Q.bit_length()
        if log_basis <= 0 or log_basis >= 64 or l <= 0 or bits // log_basis < l:
            raise ValueError("gadget basis does not fit the modulus")  # basis.rs:44-59
        if any(q <= (1 << log_basis) for q in self.moduli):
            # digits d >= B/2 become q_i - B + d (primus_rns/src/base.rs:209-213)
            raise ValueError("every limb modulus must exceed the gadget basis")
        self.log_basis, self.l, self.bits = log_basis, l, bits
        self.drop_bits = bits - l * log_basis
        self.basis = 1 << log_basis
        bm1 = self.basis - 1
        self.carry_mask = 2 if log_basis == 1 else (self.basis | (self.basis >> 1))  # :75-79
        # split_value (:81-128)
        if log_basis == 1:
            if self.drop_bits == 0:
                split = None
            else:
                v = 0
                for _ in range(l + 1):
                    v = (v << 1) | 1
                v <<= self.drop_bits - 1
                split = v if v < Q else None
        else:
            v = 0
            for _ in range(l):
                v = (v << log_basis) | (bm1 >> 1)
            v = (((v << 1) | 1) << (self.drop_bits - 1)) if self.drop_bits > 0 else v + 1
            split = v if v < Q else None
        self.split = split
        self.pow2_sub_modulus = (1 << bits) - Q  # :134-144
        # compose constants (base.rs:33-70): Q/q_i and (Q/q_i)^-1 mod q_i
        self.punctured = [Q // q for q in self.moduli]
        self.inv_punctured = [pow(Q // q % q, -1, q) for q in self.moduli]

    def compose(self, residues):
        """[L][N] residues -> N Python integers in [0, Q) (compose_inplace, base.rs:396-420:
        one product per limb added with a conditional subtraction of Q)."""
        r = np.asarray(residues, dtype=np.uint64)
        out = []
        for k in range(r.shape[1]):
            value = 0
            for i, q in enumerate(self.moduli):
                value += (int(r[i, k]) * self.inv_punctured[i] % q) * self.punctured[i]
                if value >= self.Q:
                    value -= self.Q
            out.append(value)
        return out

    def decompose(self, residues):
        """[L][N] residues -> [l][L][N] wrapped digit residues: init_value_carry
        (basis.rs:307-331), the digits with their carries (common.rs:141-152,
        184-201) and wrapping_decompose_small_values_inplace (base.rs:201-236)."""
        r = np.asarray(residues, dtype=np.uint64)
        L, n = r.shape
        out = np.zeros((self.l, L, n), dtype=np.uint64)
        half = (self.basis + 1) // 2
        bm1 = self.basis - 1
        for k, value in enumerate(self.compose(r)):
            if self.split is not None and value >= self.split:
                value += self.pow2_sub_modulus
            carry = (value >> (self.drop_bits - 1)) & 1 if self.drop_bits > 0 else 0
            for d in range(self.l):
                temp = ((value >> (self.drop_bits + d * self.log_basis)) & bm1) + carry
                carry = 1 if temp & self.carry_mask else 0
                digit = temp & bm1
                for i, q in enumerate(self.moduli):
                    out[d, i, k] = digit if (self.basis == 2 or digit < half) else q - self.basis + digit
        return out


def ext_product_rns(tables, basis, key, rlwe, ntt_output=False):
    """External product of one CRT GLWE (k = 1) [2][L][N] with a DCRT GGSW key
    [2][l][2][L][N] in NTT form (CrtGlwe::mul_dcrt_ggsw_inplace, primus_lattice/src/glwe/crt.rs:201-229,
    through add_dcrt_glev_mul_crt_poly_assign, glwe/dcrt.rs:170-247): for each input
    polynomial p and digit d, the digit's residues are transformed and
    multiply-accumulated into both output polynomials.  The result is in NTT
    form; ntt_output=False applies the inverse transform per limb."""
    key = np.asarray(key, dtype=np.uint64)
    x = np.asarray(rlwe, dtype=np.uint64)
    L, n = len(tables), tables[0].n
    acc = np.zeros((2, L, n), dtype=np.uint64)
    for p in range(2):
        digits = basis.decompose(x[p])
        for d in range(basis.l):
            for i, t in enumerate(tables):
                dh = t.transform(digits[d, i])
                for c in range(2):
                    acc[c, i] = pointwise_mul_add(acc[c, i], dh, key[p, d, c, i], t.q)
    if not ntt_output:
        for c in range(2):
            for i, t in enumerate(tables):
                acc[c, i] = t.inverse_transform(acc[c, i])
    return acc


def rlwe_encrypt(table, s, e, m, delta, a_hat):
    """b_hat = NTT(e + delta*m) + a_hat (.) NTT(s) for one limb (po_rlwe_encrypt)."""
    L = lib()
    L.po_rlwe_encrypt.restype = None
    L.po_rlwe_encrypt.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_uint64, _p64, _p64]
    out = np.zeros(table.n, dtype=np.uint64)
    a_hat = np.ascontiguousarray(a_hat, dtype=np.uint64)
    vp = ctypes.c_void_p
    L.po_rlwe_encrypt(table.handle, np.ascontiguousarray(s, dtype=np.int8).ctypes.data_as(vp),
                      np.ascontiguousarray(e, dtype=np.int8).ctypes.data_as(vp),
                      np.ascontiguousarray(m, dtype=np.uint8).ctypes.data_as(vp), delta, ptr(a_hat), ptr(out))
    return out


def rlwe_delta(moduli, qi):
    """Delta = floor(Q/256) mod q_i, Q the product of the request's moduli (Python big integers)."""
    Q = 1
    for q in moduli:
        Q *= q
    return (Q // 256) % qi


def ntt_automorphism(values, degree):
    """x -> x^degree (degree odd) on one polynomial in bit-reversed NTT order."""
    x = np.ascontiguousarray(values, dtype=np.uint64)
    out = np.zeros_like(x)
    lib().po_ntt_automorphism(ptr(x), ptr(out), x.size, degree)
    return out


def coeff_automorphism(values, degree, q):
    """a(X) -> a(X^degree) mod (X^N + 1, q) in the coefficient domain."""
    x = np.ascontiguousarray(values, dtype=np.uint64)
    out = np.zeros_like(x)
    lib().po_coeff_automorphism(ptr(x), ptr(out), x.size, degree, q)
    return out


class BaseConverter:
    """po_base_converter: RNS base conversion from ibase to obase (limb-major rows)."""

    def __init__(self, ibase, obase):
        self.ibase = np.array(ibase, dtype=np.uint64)
        self.obase = np.array(obase, dtype=np.uint64)
        h = ctypes.c_void_p()
        rc = lib().po_base_converter_new(ptr(self.ibase), len(ibase), ptr(self.obase), len(obase), ctypes.byref(h))
        if rc != 0:
            raise OracleError(rc)
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().po_base_converter_free(self._h)
            self._h = None

    def _run(self, fn, x, rows_out):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        n = x.shape[-1]
        out = np.zeros((rows_out, n), dtype=np.uint64)
        fn(self._h, ptr(x), ptr(out), n)
        return out

    def fast_convert(self, x):
        """[len(ibase)][n] residues -> [len(obase)][n]."""
        return self._run(lib().po_fast_convert_array, x, len(self.obase))

    def exact_convert(self, x):
        """[len(ibase)][n] residues -> [1][n] (obase holds one modulus)."""
        return self._run(lib().po_exact_convert_array, x, 1)


def philox4x32_10(ctr, key):
    """Philox4x32-10 block of counter ctr (4 x u32) under key (2 x u32)."""
    L = lib()
    L.po_philox4x32_10.restype = None
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.po_philox4x32_10.argtypes = [u32p, u32p, u32p]
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    L.po_philox4x32_10(c.ctypes.data_as(u32p), k.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
    return out


def gauss_thresholds():
    """The 38 cumulative thresholds of the rounded Gaussian (sigma 3.2, |e| <= 19)."""
    L = lib()
    L.po_gauss_thresholds.restype = None
    L.po_gauss_thresholds.argtypes = [_p64]
    t = np.zeros(38, dtype=np.uint64)
    L.po_gauss_thresholds(ptr(t))
    return t


def rlwe_sample(seed, request, log_n, moduli):
    """One request's (s, e, m, a_hat [limbs][N]) drawn as po_rlwe_sample does."""
    L = lib()
    vp = ctypes.c_void_p
    L.po_rlwe_sample.restype = None
    L.po_rlwe_sample.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _p64,
                                 vp, vp, vp, _p64]
    n = 1 << log_n
    mods = np.ascontiguousarray(moduli, dtype=np.uint64)
    s = np.zeros(n, dtype=np.int8)
    e = np.zeros(n, dtype=np.int8)
    m = np.zeros(n, dtype=np.uint8)
    a = np.zeros(len(moduli) * n, dtype=np.uint64)
    L.po_rlwe_sample(seed, request, log_n, len(moduli), ptr(mods), s.ctypes.data_as(vp), e.ctypes.data_as(vp),
                     m.ctypes.data_as(vp), ptr(a))
    return s, e, m, a.reshape(len(moduli), n)
