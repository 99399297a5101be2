"""Continuous-batching request server over libprimus_hip.so (ph_server_*).

Requests are negacyclic products, RLWE encryptions and external products of
mixed ring degrees and limb counts.  Host arrays passed in a request are kept
referenced here until the server has finished with them.  Requests whose
arrays live in page-locked memory (`Request(..., pinned=True)`, allocated with
ph_host_alloc) are copied to and from the device without staging copies.
"""
import ctypes

import numpy as np

from ._lib import lib
from .ntt import _check

POLY_MUL, RLWE_ENCRYPT, EXT_PRODUCT = 0, 1, 2

_vp = ctypes.c_void_p


class ServeRequest(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("log_n", ctypes.c_uint32), ("limbs", ctypes.c_uint32),
                ("in0", _vp), ("in1", _vp), ("s", _vp), ("e", _vp), ("m", _vp), ("out", _vp)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_batch", ctypes.c_uint32), ("streams", ctypes.c_uint32),
                ("min_log_n", ctypes.c_uint32), ("max_log_n", ctypes.c_uint32), ("max_limbs", ctypes.c_uint32),
                ("moduli", ctypes.POINTER(ctypes.c_uint64)), ("ext_key", _vp), ("per_poly_launches", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("peak_inflight_bytes", ctypes.c_uint64), ("cancelled", ctypes.c_uint64), ("submit_s", ctypes.c_double),
                ("dispatch_s", ctypes.c_double), ("slot_wait_s", ctypes.c_double), ("device_wait_s", ctypes.c_double)]


def _declare():
    L = lib()
    if getattr(L, "_serving_declared", False):
        return L
    L.ph_server_new.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
    L.ph_server_new.restype = ctypes.c_int
    L.ph_server_free.argtypes = [_vp]
    L.ph_server_free.restype = None
    L.ph_server_submit.argtypes = [_vp, ctypes.POINTER(ServeRequest), ctypes.POINTER(ctypes.c_uint64)]
    L.ph_server_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
    L.ph_server_latency.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    L.ph_server_drain.argtypes = [_vp]
    L.ph_server_cancel.argtypes = [_vp, ctypes.c_uint64]
    L.ph_server_inject_fault.argtypes = [_vp, ctypes.c_uint32]
    L.ph_server_pause.argtypes = [_vp, ctypes.c_int]
    L.ph_server_stats_get.argtypes = [_vp, ctypes.POINTER(ServerStats)]
    L.ph_server_run_poisson.argtypes = [_vp, ctypes.POINTER(ServeRequest), ctypes.c_size_t, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
    L.ph_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(_vp)]
    L.ph_host_alloc.restype = ctypes.c_int
    L.ph_host_free.argtypes = [_vp]
    L.ph_host_free.restype = ctypes.c_int
    for f in ("ph_server_submit", "ph_server_wait", "ph_server_latency", "ph_server_drain", "ph_server_stats_get",
              "ph_server_run_poisson", "ph_server_cancel", "ph_server_inject_fault", "ph_server_pause"):
        getattr(L, f).restype = ctypes.c_int
    L._serving_declared = True
    return L


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


class PinnedBuffer:
    """Page-locked host memory (ph_host_alloc) with a numpy view; freed with the object."""

    def __init__(self, shape, dtype):
        L = _declare()
        dtype = np.dtype(dtype)
        nbytes = max(int(np.prod(shape)) * dtype.itemsize, 1)
        p = _vp()
        _check(L.ph_host_alloc(nbytes, ctypes.byref(p)))
        self._p = p
        raw = (ctypes.c_uint8 * nbytes).from_address(p.value)
        self.array = np.frombuffer(raw, dtype=np.uint8, count=int(np.prod(shape)) * dtype.itemsize).view(
            dtype).reshape(shape)

    def __del__(self):
        if getattr(self, "_p", None):
            try:
                lib().ph_host_free(self._p)
            except Exception:
                pass
            self._p = None


def _pinned_copy(a, keep):
    if a is None:
        return None
    buf = PinnedBuffer(a.shape, a.dtype)
    buf.array[...] = a
    keep.append(buf)
    return buf.array


class Request:
    """A request with its host arrays (kept alive by this object)."""

    def __init__(self, kind, log_n, limbs=1, in0=None, in1=None, s=None, e=None, m=None, pinned=False):
        n = 1 << log_n
        self.kind, self.log_n, self.limbs = kind, log_n, limbs
        self.in0 = np.ascontiguousarray(in0, dtype=np.uint64)
        self.in1 = None if in1 is None else np.ascontiguousarray(in1, dtype=np.uint64)
        self.s = None if s is None else np.ascontiguousarray(s, dtype=np.int8)
        self.e = None if e is None else np.ascontiguousarray(e, dtype=np.int8)
        self.m = None if m is None else np.ascontiguousarray(m, dtype=np.uint8)
        words = 2 * n if kind == EXT_PRODUCT else limbs * n
        self.out = np.zeros(words, dtype=np.uint64)
        self._pinned = []
        if pinned:
            for name in ("in0", "in1", "s", "e", "m", "out"):
                setattr(self, name, _pinned_copy(getattr(self, name), self._pinned))

    def pinned_copy(self):
        """The same request with every array in page-locked memory."""
        return Request(self.kind, self.log_n, self.limbs, self.in0, self.in1, self.s, self.e, self.m, pinned=True)

    def c(self):
        return ServeRequest(self.kind, self.log_n, self.limbs, _ptr(self.in0), _ptr(self.in1), _ptr(self.s),
                            _ptr(self.e), _ptr(self.m), _ptr(self.out))


class Server:
    def __init__(self, moduli, min_log_n, max_log_n, max_limbs, ext_key=None, hbm_budget_bytes=64 << 30,
                 max_batch=256, streams=4, per_poly_launches=False):
        L = _declare()
        grid = np.zeros((max_log_n - min_log_n + 1, max_limbs), dtype=np.uint64)
        for ln in range(min_log_n, max_log_n + 1):
            grid[ln - min_log_n] = moduli[ln][:max_limbs]
        self._grid = grid
        self._ext_key = ext_key
        cfg = ServerConfig(hbm_budget_bytes, max_batch, streams, min_log_n, max_log_n, max_limbs,
                           grid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                           ext_key.handle if ext_key is not None else None, int(per_poly_launches))
        h = _vp()
        _check(L.ph_server_new(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._live = {}

    def submit(self, req):
        t = ctypes.c_uint64()
        cr = req.c()
        _check(lib().ph_server_submit(self._h, ctypes.byref(cr), ctypes.byref(t)))
        self._live[t.value] = req
        return t.value

    def wait(self, ticket):
        st = ctypes.c_int()
        _check(lib().ph_server_wait(self._h, ticket, ctypes.byref(st)))
        self._live.pop(ticket, None)
        return st.value

    def latency(self, ticket):
        v = ctypes.c_double()
        _check(lib().ph_server_latency(self._h, ticket, ctypes.byref(v)))
        return v.value

    def drain(self):
        _check(lib().ph_server_drain(self._h))

    def cancel(self, ticket):
        """Cancels a still-queued request; True if it was cancelled (its status
        is then PH_ERR_CANCELLED), False if it was already in a batch."""
        rc = lib().ph_server_cancel(self._h, ticket)
        if rc == 0:
            self._live.pop(ticket, None)
        return rc == 0

    def pause(self, paused=True):
        """While paused, submissions queue and no new batch is formed."""
        _check(lib().ph_server_pause(self._h, int(bool(paused))))

    def inject_fault(self, batches=1):
        """The next `batches` batches complete with PH_ERR_DEVICE without running."""
        _check(lib().ph_server_inject_fault(self._h, batches))

    def stats(self):
        st = ServerStats()
        _check(lib().ph_server_stats_get(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def run_poisson(self, requests, rate_per_s, seed=0):
        """Open-loop Poisson load; returns (latencies in seconds, elapsed seconds)."""
        arr = (ServeRequest * len(requests))(*[r.c() for r in requests])
        lat = np.zeros(len(requests), dtype=np.float64)
        el = ctypes.c_double()
        _check(lib().ph_server_run_poisson(self._h, arr, len(requests), rate_per_s, seed,
                                           lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(el)))
        return lat, el.value

    def close(self):
        if getattr(self, "_h", None):
            lib().ph_server_free(self._h)
            self._h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
