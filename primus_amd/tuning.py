"""Launch-shape autotuning over libprimus_hip.so's process-wide tuning table.

For each (ring degree, operation, arithmetic) the single-pass kernels can hold
8 or 16 coefficients per thread (radix-8 rounds with twice the waves, or
radix-16 rounds with fewer LDS exchanges).  `autotune` times both on the
caller's batch shape and keeps the faster; `save`/`load` persist the table.
Outputs never depend on the choice.
"""
import ctypes

from ._lib import lib
from .ntt import _check

OPS = {"forward_lazy": 0, "forward": 1, "inverse": 2, "fwd_mul_inv": 3, "negacyclic_mul": 4}


def set_shape(log_n, op, arith, loge):
    _check(lib().ph_tuning_set(log_n, OPS[op], 1 if arith == "f64" else 0, loge))


def get_shape(log_n, op, arith):
    return lib().ph_tuning_get(log_n, OPS[op], 1 if arith == "f64" else 0)


def autotune(table, op, count, reps=20):
    """Time every launch shape of `op` on `count` limb-polynomials; returns (loge, ms)."""
    chosen, best = ctypes.c_int(), ctypes.c_float()
    _check(lib().ph_autotune(table.handle, OPS[op], count, reps, ctypes.byref(chosen), ctypes.byref(best)))
    return chosen.value, best.value


def save(path):
    _check(lib().ph_tuning_save(path.encode()))


def load(path):
    _check(lib().ph_tuning_load(path.encode()))
