"""RNS base conversion over libprimus_hip.so: the reference's BaseConverter
(primus_rns/src/converter.rs) with its fast and exact conversions on
limb-major CRT polynomials resident on the device."""
import ctypes

import numpy as np

from ._lib import lib
from .ntt import _check, _stream

_p64 = ctypes.POINTER(ctypes.c_uint64)


class BaseConverter:
    """BaseConverter::new(ibase, obase): input moduli pairwise coprime, all in (1, 2^62)."""

    def __init__(self, ibase, obase):
        self.ibase = np.ascontiguousarray(ibase, dtype=np.uint64)
        self.obase = np.ascontiguousarray(obase, dtype=np.uint64)
        h = ctypes.c_void_p()
        _check(lib().ph_base_converter_new(self.ibase.ctypes.data_as(_p64), len(self.ibase),
                                           self.obase.ctypes.data_as(_p64), len(self.obase), ctypes.byref(h)))
        self._h = h

    def fast_convert_device(self, d_in, d_out, log_n, count, stream=None):
        """fast_convert_array: [count][len(ibase)][N] -> [count][len(obase)][N]."""
        _check(lib().ph_fast_convert_array_device(self._h, d_in, d_out, log_n, count, _stream(stream)))

    def exact_convert_device(self, d_in, d_out, log_n, count, stream=None):
        """exact_convert_array (one output modulus): [count][len(ibase)][N] -> [count][1][N]."""
        _check(lib().ph_exact_convert_array_device(self._h, d_in, d_out, log_n, count, _stream(stream)))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ph_base_converter_free(self._h)
            except Exception:
                pass
            self._h = None
