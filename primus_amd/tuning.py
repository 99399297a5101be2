"""Launch-shape autotuning over libprimus_hip.so's process-wide tuning table.

A shape is a set of bit fields (0 is the default shape):

* radix: 8 or 16 coefficients per thread (radix-8 rounds with twice the waves,
  or radix-16 rounds with fewer LDS exchanges), `3` or `4`;
* STREAM, for forward-first operations at 2^10 <= N <= 2^13: the persistent
  streaming kernel, in which each workgroup loads its next limb-polynomial
  while it transforms the current one; with lpb(k) each workgroup walks about
  2^k limb-polynomials (k = 0: the grid is one resident generation);
* tile(c), for N >= 2^14: the two-pass schedule's LDS tile, chunks of 2^c
  coefficients (c = 11..13) after a column pass over N / 2^c.

`autotune` times every shape the library lists for the operation and degree
(`shapes`) on the caller's batch and keeps the fastest; `save`/`load` persist
the table, and the file the environment variable PH_TUNING_FILE names is
loaded when the process creates its first table.  Outputs never depend on the
shape.
"""
import ctypes

from ._lib import lib
from .ntt import _check

OPS = {"forward_lazy": 0, "forward": 1, "inverse": 2, "fwd_mul_inv": 3, "negacyclic_mul": 4}
STREAM = 16  # PH_SHAPE_STREAM


def lpb(k):
    """PH_SHAPE_LPB(k): with STREAM, about 2^k limb-polynomials per workgroup."""
    return k << 5


def tile(c):
    """PH_SHAPE_TILE(c): the two-pass LDS tile of 2^c coefficients."""
    return c << 8


def describe(shape):
    """(coefficients per thread, streaming) of a tuning-table shape."""
    return 1 << ((shape & 15) or 4), bool(shape & STREAM)


def fields(shape):
    """Every field of a shape, as a dict."""
    return {"coefficients_per_thread": 1 << ((shape & 15) or 4), "streaming": bool(shape & STREAM),
            "limb_polys_per_workgroup_log2": (shape >> 5) & 3 if shape & STREAM else None,
            "two_pass_tile": 1 << (((shape >> 8) & 15) or 12)}


def shapes(log_n, op):
    """The shapes the autotuner tries for (log_n, op)."""
    L = lib()
    n = L.ph_tuning_shapes(log_n, OPS[op], None, 0)
    if n < 0:
        _check(-n)
    buf = (ctypes.c_int * max(n, 1))()
    L.ph_tuning_shapes(log_n, OPS[op], buf, n)
    return list(buf)[:n]


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
