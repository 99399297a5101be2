"""Launch-shape autotuning over libprimus_hip.so's process-wide tuning table.

For each (ring degree, operation, arithmetic) the single-pass kernels
(N <= 2^13) run in one of these shapes:

* 8 or 16 coefficients per thread (radix-8 rounds with twice the waves, or
  radix-16 rounds with fewer LDS exchanges), one workgroup per polynomial;
* the same as a persistent streaming kernel (STREAM added; forward-first
  operations, N >= 2^10): one resident generation of workgroups walks the
  batch, each loading its next polynomial while it transforms the current one.

`autotune` times every shape on the caller's batch shape and keeps the
fastest; `save`/`load` persist the table.  Outputs never depend on the choice.
"""
import ctypes

from ._lib import lib
from .ntt import _check

OPS = {"forward_lazy": 0, "forward": 1, "inverse": 2, "fwd_mul_inv": 3, "negacyclic_mul": 4}
STREAM = 16  # PH_SHAPE_STREAM
SHAPES = (3, 4, 3 | STREAM, 4 | STREAM)


def describe(shape):
    """(coefficients per thread, streaming) of a tuning-table shape (0 = the default)."""
    return 1 << ((shape & 15) or 4), bool(shape & STREAM)


def set_shape(log_n, op, arith, shape):
    _check(lib().ph_tuning_set(log_n, OPS[op], 1 if arith == "f64" else 0, shape))


def get_shape(log_n, op, arith):
    return lib().ph_tuning_get(log_n, OPS[op], 1 if arith == "f64" else 0)


def autotune(table, op, count, reps=20):
    """Time every launch shape of `op` on `count` limb-polynomials; returns (shape, ms)."""
    chosen, best = ctypes.c_int(), ctypes.c_float()
    _check(lib().ph_autotune(table.handle, OPS[op], count, reps, ctypes.byref(chosen), ctypes.byref(best)))
    return chosen.value, best.value


def save(path):
    _check(lib().ph_tuning_save(path.encode()))


def load(path):
    _check(lib().ph_tuning_load(path.encode()))
