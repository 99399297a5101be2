"""Host-side mirror of the reference's NttTable / DcrtTable API over libprimus_hip.so.

Method names, argument meaning, ranges and in-place semantics follow
primus_ntt/src/ntt/mod.rs:20-117 (NttTable, impl UintNttTable in
primus_ntt/src/ntt/primitive.rs) and primus_ntt/src/dcrt/mod.rs:17-129
(DcrtTable, impl UintCrtNttTable in primus_ntt/src/dcrt/primitive.rs).
Slices are numpy uint64 arrays, transformed in place on the GPU.
Constructor failures raise NttError subclasses mirroring
primus_ntt/src/error.rs:7-35; a bad modulus raises ValueError (the reference
panics in BarrettModulus::new).
"""
import ctypes

import numpy as np

from ._lib import PrimusHipError, lib

_p64 = ctypes.POINTER(ctypes.c_uint64)


class NttError(Exception):
    """primus_ntt::NttError."""


class NoPrimitiveRoot(NttError):
    pass


class DegreeConversionErr(NttError):
    pass


class DegreeTooLarge(NttError):
    pass


class NttTableErr(NttError):
    pass


_CTOR_ERRORS = {1: NoPrimitiveRoot, 2: DegreeConversionErr, 3: DegreeTooLarge, 4: NttTableErr}


def _raise(code):
    msg = lib().ph_last_error().decode(errors="replace")
    if code in _CTOR_ERRORS:
        raise _CTOR_ERRORS[code](msg)
    if code == 5:
        raise ValueError(msg)
    raise PrimusHipError(code, msg)


def _check(code):
    if code != 0:
        _raise(code)


def _slice_ptr(a):
    if not isinstance(a, np.ndarray) or a.dtype != np.uint64 or not a.flags["C_CONTIGUOUS"] \
            or not a.flags["WRITEABLE"]:
        raise TypeError("expected a writeable C-contiguous numpy uint64 array")
    return a.ctypes.data_as(_p64)


class UintNttTable:
    """NttTable over one modulus q < 2^62 with device-resident Shoup twiddles."""

    def __init__(self, handle, owner=None):
        self._h = ctypes.c_void_p(handle) if isinstance(handle, int) else handle
        self._owner = owner  # a DCRT table that owns this one, if any

    @classmethod
    def new(cls, log_n, modulus, integer_only=False):
        h = ctypes.c_void_p()
        _check(lib().ph_ntt_table_new_ex(log_n, modulus, int(bool(integer_only)), ctypes.byref(h)))
        return cls(h)

    def arith(self):
        """'f64' if canonical transforms use binary64 FMA arithmetic, else 'int'."""
        return "f64" if lib().ph_ntt_table_arith(self._h) == 1 else "int"

    def __del__(self):
        if getattr(self, "_owner", None) is None and getattr(self, "_h", None):
            try:
                lib().ph_ntt_table_free(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def poly_length(self):
        return lib().ph_ntt_table_poly_length(self._h)

    def log_n(self):
        return lib().ph_ntt_table_log_n(self._h)

    def modulus(self):
        return lib().ph_ntt_table_modulus(self._h)

    def root(self):
        return lib().ph_ntt_table_root(self._h)

    def inv_root(self):
        return lib().ph_ntt_table_inv_root(self._h)

    def inv_n(self):
        v, w = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().ph_ntt_table_inv_n(self._h, ctypes.byref(v), ctypes.byref(w)))
        return v.value, w.value

    def _powers(self, which):
        n = self.poly_length() * (2 if which == 2 else 1)
        v = np.zeros(n, dtype=np.uint64)
        w = np.zeros(n, dtype=np.uint64)
        _check(lib().ph_ntt_table_powers(self._h, which, _slice_ptr(v), _slice_ptr(w)))
        return v, w

    def root_powers(self):
        return self._powers(0)

    def inv_root_powers(self):
        return self._powers(1)

    def ordinal_root_powers(self):
        return self._powers(2)

    # -- slice transforms (in place) --
    def lazy_transform_slice(self, poly):
        _check(lib().ph_lazy_transform_slice(self._h, _slice_ptr(poly), poly.size))

    def transform_slice(self, poly):
        _check(lib().ph_transform_slice(self._h, _slice_ptr(poly), poly.size))

    def lazy_inverse_transform_slice(self, values):
        _check(lib().ph_lazy_inverse_transform_slice(self._h, _slice_ptr(values), values.size))

    def inverse_transform_slice(self, values):
        _check(lib().ph_inverse_transform_slice(self._h, _slice_ptr(values), values.size))

    def transform_monomial(self, coeff, degree, values):
        _check(lib().ph_transform_monomial(self._h, coeff, degree, _slice_ptr(values), values.size))

    def transform_coeff_one_monomial(self, degree, values):
        _check(lib().ph_transform_coeff_one_monomial(self._h, degree, _slice_ptr(values), values.size))

    def transform_coeff_minus_one_monomial(self, degree, values):
        _check(lib().ph_transform_coeff_minus_one_monomial(self._h, degree, _slice_ptr(values), values.size))

    # -- Polynomial wrappers (transform_inplace / inverse_transform_inplace) --
    def transform_inplace(self, poly):
        self.transform_slice(poly)
        return poly

    def inverse_transform_inplace(self, values):
        self.inverse_transform_slice(values)
        return values


class HexlNttTable(UintNttTable):
    """HexlNttTable (primus_ntt/src/ntt/hexl/mod.rs): the NttTable interface plus
    compute_forward / compute_inverse with HEXL's range factors, on the same
    device tables and kernels.  Full-range outputs and monomials are identical
    to UintNttTable's (the reference asserts it in primus_ntt/tests/hexl.rs);
    lazy outputs (output_mod_factor 4 forward, 2 inverse) are HEXL's own
    radix-2 values (hexl/radix2.rs), bit for bit."""

    def compute_forward(self, operand, input_mod_factor, output_mod_factor):
        _check(lib().ph_hexl_compute_forward(self._h, _slice_ptr(operand), operand.size, input_mod_factor,
                                             output_mod_factor))

    def compute_inverse(self, operand, input_mod_factor, output_mod_factor):
        _check(lib().ph_hexl_compute_inverse(self._h, _slice_ptr(operand), operand.size, input_mod_factor,
                                             output_mod_factor))

    # the reference's NttTable methods of the HEXL table (hexl/mod.rs:321-338)
    def lazy_transform_slice(self, poly):
        self.compute_forward(poly, 1, 4)

    def transform_slice(self, poly):
        self.compute_forward(poly, 1, 1)

    def lazy_inverse_transform_slice(self, values):
        self.compute_inverse(values, 1, 2)

    def inverse_transform_slice(self, values):
        self.compute_inverse(values, 1, 1)


class UintCrtNttTable:
    """DcrtTable: one UintNttTable per modulus over limb-major [L][N] data."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def new(cls, log_n, moduli, integer_only=False):
        arr = np.ascontiguousarray(np.array(moduli, dtype=np.uint64))
        h = ctypes.c_void_p()
        _check(lib().ph_dcrt_table_new_ex(log_n, arr.ctypes.data_as(_p64), arr.size, int(bool(integer_only)),
                                          ctypes.byref(h)))
        return cls(h)

    def arith(self):
        return "f64" if lib().ph_dcrt_table_arith(self._h) == 1 else "int"

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ph_dcrt_table_free(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def poly_length(self):
        return lib().ph_dcrt_poly_length(self._h)

    def moduli_count(self):
        return lib().ph_dcrt_moduli_count(self._h)

    def crt_poly_length(self):
        return lib().ph_dcrt_crt_poly_length(self._h)

    def ntt_tables(self):
        return [UintNttTable(lib().ph_dcrt_ntt_table(self._h, i), owner=self) for i in range(self.moduli_count())]

    def moduli(self):
        return [t.modulus() for t in self.ntt_tables()]

    def lazy_transform_slice(self, poly):
        _check(lib().ph_dcrt_lazy_transform_slice(self._h, _slice_ptr(poly), poly.size))

    def transform_slice(self, poly):
        _check(lib().ph_dcrt_transform_slice(self._h, _slice_ptr(poly), poly.size))

    def lazy_inverse_transform_slice(self, poly):
        _check(lib().ph_dcrt_lazy_inverse_transform_slice(self._h, _slice_ptr(poly), poly.size))

    def inverse_transform_slice(self, poly):
        _check(lib().ph_dcrt_inverse_transform_slice(self._h, _slice_ptr(poly), poly.size))

    def transform_inplace(self, crt_poly):
        self.transform_slice(crt_poly)
        return crt_poly

    def inverse_transform_inplace(self, dcrt_poly):
        self.inverse_transform_slice(dcrt_poly)
        return dcrt_poly


class DeviceBuffer:
    """A device allocation of uint64 words owned by libprimus_hip (hipMalloc)."""

    def __init__(self, words):
        self.words = int(words)
        p = ctypes.c_void_p()
        _check(lib().ph_device_alloc(max(self.words, 1) * 8, ctypes.byref(p)))
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = cls(a.size)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        if a.size > self.words:
            raise ValueError("upload of %d words into a buffer of %d" % (a.size, self.words))
        _check(lib().ph_memcpy_htod(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.size * 8))

    def download(self, shape=None):
        out = np.empty(self.words, dtype=np.uint64)
        _check(lib().ph_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.words * 8))
        return out.reshape(shape) if shape is not None else out

    def free(self):
        if self.ptr:
            lib().ph_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Operand:
    """Device-resident Shoup-prepared NTT-domain operand ([L][N] canonical values)."""

    def __init__(self, table, values):
        vals = np.ascontiguousarray(values, dtype=np.uint64)
        if vals.size != table.crt_poly_length():
            raise ValueError("operand must hold L x N = %d values, got %d" % (table.crt_poly_length(), vals.size))
        h = ctypes.c_void_p()
        _check(lib().ph_operand_new(table.handle, vals.ctypes.data_as(_p64), ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ph_operand_free(self._h)
            except Exception:
                pass
            self._h = None


def _stream(stream):
    return ctypes.c_void_p(stream) if isinstance(stream, int) else stream


def transform_device(table, dptr, count, lazy=False, stream=None):
    _check(lib().ph_dcrt_transform_device(table.handle, dptr, count, int(lazy), _stream(stream)))


def inverse_transform_device(table, dptr, count, stream=None):
    _check(lib().ph_dcrt_inverse_transform_device(table.handle, dptr, count, _stream(stream)))


def fwd_mul_inv_device(table, dptr, count, operand, stream=None):
    _check(lib().ph_dcrt_fwd_mul_inv_device(table.handle, dptr, count, operand.handle, _stream(stream)))


def negacyclic_mul_device(table, da, db, count, stream=None):
    _check(lib().ph_dcrt_negacyclic_mul_device(table.handle, da, db, count, _stream(stream)))


def negacyclic_mul_rows_device(table, da, db, d_limb_of, count, stream=None):
    """negacyclic_mul_device over rows of uneven limbs: row i uses the table's
    modulus d_limb_of[i] (device bytes) instead of i % L."""
    _check(lib().ph_dcrt_negacyclic_mul_rows_device(table.handle, da, db, d_limb_of, count, _stream(stream)))


def mul_factor_device(table, dptr, count, factors, bit_shift=64, lazy=False, stream=None):
    """x <- factors[i] * x mod q_i per limb-polynomial (MultiplyFactor at bit_shift
    32, 52 or 64; lazy: [0, 2q))."""
    f = np.ascontiguousarray(factors, dtype=np.uint64)
    if f.size != table.moduli_count():
        raise ValueError("one factor per limb")
    _check(lib().ph_dcrt_mul_factor_device(table.handle, dptr, count, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                           bit_shift, int(bool(lazy)), _stream(stream)))


def pointwise_mul_device(table, da, db, dout, count, stream=None):
    _check(lib().ph_dcrt_pointwise_mul_device(table.handle, da, db, dout, count, _stream(stream)))


def pointwise_mul_add_device(table, dacc, da, db, count, stream=None):
    _check(lib().ph_dcrt_pointwise_mul_add_device(table.handle, dacc, da, db, count, _stream(stream)))


def automorphism_device(table, d_in, d_out, count, degree, stream=None):
    """x -> x^degree (odd) on `count` limb-polynomials in NTT form, out of place."""
    _check(lib().ph_dcrt_automorphism_device(table.handle, d_in, d_out, count, degree, _stream(stream)))


def synchronize(stream=None):
    _check(lib().ph_stream_synchronize(_stream(stream)))
