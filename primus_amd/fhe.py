"""Host-side mirror of the batched callers of the NTT: the RLWE x RGSW external
product (primus_lattice/src/glwe/crt.rs:202-229 -> primus_lattice/src/glwe/dcrt.rs:170-247)
and RLWE encryption (primus_fhe_core/src/secret_key/rlwe.rs:227-268), over
libprimus_hip.so."""
import ctypes

import numpy as np

from ._lib import lib
from .ntt import _check, _stream

_p64 = ctypes.POINTER(ctypes.c_uint64)


class ExtKey:
    """Device-resident GGSW key [2][l][2][N] (NTT form) with its gadget basis."""

    def __init__(self, table, log_basis, decompose_length, key_ntt):
        k = np.ascontiguousarray(key_ntt, dtype=np.uint64)
        n = table.poly_length()
        assert k.size == 4 * decompose_length * n
        h = ctypes.c_void_p()
        _check(lib().ph_ext_key_new(table.handle, log_basis, decompose_length, k.ctypes.data_as(_p64),
                                    ctypes.byref(h)))
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
