"""Host-side mirror of the batched callers of the NTT: the RLWE x RGSW external
product (primus_lattice/src/glwe/crt.rs:202-229 -> primus_lattice/src/glwe/dcrt.rs:170-247)
and RLWE encryption (primus_fhe_core/src/secret_key/rlwe.rs:227-268), over
libprimus_hip.so."""
import ctypes

import numpy as np

from ._lib import lib
from .ntt import _check, _stream

_p64 = ctypes.POINTER(ctypes.c_uint64)


DIGITS_APPROX_SIGNED = 0  # PH_DIGITS_APPROX_SIGNED: the reference's digits
DIGITS_UNSIGNED = 1  # PH_DIGITS_UNSIGNED: the literal unsigned digits


class ExtKey:
    """Device-resident GGSW key [2][l][2][N] (NTT form) with its gadget basis;
    unsigned_digits selects the literal unsigned digits (ph_ext_key_new_digits)
    instead of the reference's approximate signed ones."""

    def __init__(self, table, log_basis, decompose_length, key_ntt, unsigned_digits=False):
        k = np.ascontiguousarray(key_ntt, dtype=np.uint64)
        n = table.poly_length()
        if k.size != 4 * decompose_length * n:
            raise ValueError("key must hold 2 x l x 2 x N values (%d), got %d" % (4 * decompose_length * n, k.size))
        h = ctypes.c_void_p()
        digits = DIGITS_UNSIGNED if unsigned_digits else DIGITS_APPROX_SIGNED
        _check(lib().ph_ext_key_new_digits(table.handle, log_basis, decompose_length, digits,
                                           k.ctypes.data_as(_p64), ctypes.byref(h)))
        self._h = h
        self.table = table  # keeps the table alive
        self.n = n
        self.l = decompose_length

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ph_ext_key_free(self._h)
            except Exception:
                pass
            self._h = None

    def product(self, rlwe, ntt_output=False):
        """External products of a [count][2][N] batch of RLWE ciphertexts (host arrays)."""
        x = np.ascontiguousarray(rlwe, dtype=np.uint64)
        count = x.size // (2 * self.n)
        out = np.empty_like(x)
        _check(lib().ph_ext_product(self._h, x.ctypes.data_as(_p64), out.ctypes.data_as(_p64), count,
                                    int(bool(ntt_output))))
        return out

    def product_device(self, d_in, d_out, count, ntt_output=False, stream=None):
        _check(lib().ph_ext_product_device(self._h, d_in, d_out, count, int(bool(ntt_output)), _stream(stream)))


class ExtKeyRns:
    """Device-resident DCRT GGSW key [2][l][2][L][N] (NTT form) over a
    UintCrtNttTable's L limbs, with the gadget basis over their product Q
    (CrtGlwe::mul_dcrt_ggsw_inplace, primus_lattice/src/glwe/crt.rs:201-229)."""

    def __init__(self, table, log_basis, decompose_length, key_ntt):
        k = np.ascontiguousarray(key_ntt, dtype=np.uint64)
        self.L, self.n = table.moduli_count(), table.poly_length()
        want = 4 * decompose_length * self.L * self.n
        if k.size != want:
            raise ValueError("key must hold 2 x l x 2 x L x N values (%d), got %d" % (want, k.size))
        h = ctypes.c_void_p()
        _check(lib().ph_ext_key_rns_new(table.handle, log_basis, decompose_length, k.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(h)))
        self._h = h
        self.table = table  # keeps the table alive
        self.l = decompose_length

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ph_ext_key_rns_free(self._h)
            except Exception:
                pass
            self._h = None

    def product_device(self, d_in, d_out, count, ntt_output=False, stream=None):
        """[count][2][L][N] residues in, [count][2][L][N] out (device pointers)."""
        _check(lib().ph_ext_product_rns_device(self._h, d_in, d_out, count, int(bool(ntt_output)), _stream(stream)))


class RlweRequest(ctypes.Structure):
    """ph_rlwe_request: one request of an uneven encryption batch."""

    _fields_ = [("log_n", ctypes.c_uint32), ("limbs", ctypes.c_uint32), ("a_offset", ctypes.c_uint64),
                ("small_offset", ctypes.c_uint64)]


class RlweContext:
    """Device-resident tables for RLWE encryption over a grid of (ring degree, limb).

    moduli[log_n] lists the limb moduli of ring degree 2^log_n (at least max_limbs each).
    """

    def __init__(self, min_log_n, max_log_n, max_limbs, moduli):
        grid = np.zeros(((max_log_n - min_log_n + 1), max_limbs), dtype=np.uint64)
        for ln in range(min_log_n, max_log_n + 1):
            grid[ln - min_log_n] = moduli[ln][:max_limbs]
        h = ctypes.c_void_p()
        _check(lib().ph_rlwe_ctx_new(min_log_n, max_log_n, max_limbs, grid.ctypes.data_as(_p64), ctypes.byref(h)))
        self._h = h
        self.min_log_n, self.max_log_n, self.max_limbs = min_log_n, max_log_n, max_limbs

    def modulus(self, log_n, limb):
        return lib().ph_rlwe_ctx_modulus(self._h, log_n, limb)

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ph_rlwe_ctx_free(self._h)
            except Exception:
                pass
            self._h = None

    def encrypt_device(self, requests, d_a, d_s, d_e, d_m, d_out, stream=None):
        """requests: sequence of (log_n, limbs, a_offset, small_offset)."""
        arr = (RlweRequest * len(requests))(*[RlweRequest(*r) for r in requests])
        _check(lib().ph_rlwe_encrypt_device(self._h, ctypes.cast(arr, ctypes.c_void_p), len(requests), d_a, d_s,
                                            d_e, d_m, d_out, _stream(stream)))

    def sample_device(self, requests, seed, d_a, d_s, d_e, d_m, stream=None):
        """Draws s, e, m and a_hat for the requests on the device (ph_rlwe_sample_device)."""
        arr = (RlweRequest * len(requests))(*[RlweRequest(*r) for r in requests])
        _check(lib().ph_rlwe_sample_device(self._h, ctypes.cast(arr, ctypes.c_void_p), len(requests), seed, d_a,
                                           d_s, d_e, d_m, _stream(stream)))

    def plan(self, requests):
        """A prepared launch for this batch layout (RlwePlan)."""
        return RlwePlan(self, requests)


class RlwePlan:
    """ph_rlwe_plan: the batch's work items and scratch resident on the device;
    encrypt_device only enqueues the persistent kernel (capturable in a HIP graph)."""

    def __init__(self, ctx, requests):
        arr = (RlweRequest * len(requests))(*[RlweRequest(*r) for r in requests])
        h = ctypes.c_void_p()
        _check(lib().ph_rlwe_plan_new(ctx._h, ctypes.cast(arr, ctypes.c_void_p), len(requests), ctypes.byref(h)))
        self._h, self._ctx = h, ctx

    @property
    def grid(self):
        return lib().ph_rlwe_plan_grid(self._h)

    def encrypt_device(self, d_a, d_s, d_e, d_m, d_out, stream=None):
        _check(lib().ph_rlwe_encrypt_plan_device(self._h, d_a, d_s, d_e, d_m, d_out, _stream(stream)))

    def sample_device(self, seed, d_a, d_s, d_e, d_m, stream=None):
        """Draws the plan's s, e, m and a_hat on the device (ph_rlwe_sample_plan_device)."""
        _check(lib().ph_rlwe_sample_plan_device(self._h, seed, d_a, d_s, d_e, d_m, _stream(stream)))

    def autotune(self, d_a, d_s, d_e, d_m, d_out, reps=5):
        """Times persistent grids of 1..4 workgroups per CU; returns (grid, ms)."""
        g, ms = ctypes.c_uint32(), ctypes.c_float()
        _check(lib().ph_rlwe_plan_autotune(self._h, d_a, d_s, d_e, d_m, d_out, reps, ctypes.byref(g),
                                           ctypes.byref(ms)))
        return g.value, ms.value

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ph_rlwe_plan_free(self._h)
            except Exception:
                pass
            self._h = None


def rlwe_layout(shapes):
    """Offsets for a batch of (log_n, limbs) requests: (requests, a_words, small_elems)."""
    reqs, a_off, s_off = [], 0, 0
    for log_n, limbs in shapes:
        reqs.append((log_n, limbs, a_off, s_off))
        a_off += limbs << log_n
        s_off += 1 << log_n
    return reqs, a_off, s_off
