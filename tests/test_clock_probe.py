"""The clock-probe measurement variant of the library (libprimus_hip_probe.so,
ntt_kernels.hip built with PH_CLOCK_PROBE=1): every wave of the single-pass
chunk kernels logs, as it ends, the shader-clock and 100 MHz counters, its
hardware slot and its workgroup.  It is measurement tooling (the engine clock
the fused batch kernel runs at), so its outputs must be the default library's
and its log must hold one entry per wave."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import primus_amd as pa
from primus_amd import _lib

UNSUPPORTED = 7  # PH_ERR_UNSUPPORTED


def test_default_build_has_no_clock_probe():
    assert _lib.lib().ph_debug_clock_report(None, None, 0) == UNSUPPORTED


def test_probe_variant_exports_the_header():
    if not os.path.exists(_lib.PROBE_LIB_PATH):
        pytest.skip("libprimus_hip_probe.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.PROBE_LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert exported == set(_lib.header_functions())


def _fused_batch(L, log_n, moduli, polys, operand):
    """forward -> Shoup multiply -> inverse on the device through library L."""
    saved = _lib._lib
    _lib._lib = L
    try:
        tab = pa.UintCrtNttTable.new(log_n, moduli)
        op = pa.Operand(tab, operand)
        buf = pa.DeviceBuffer.from_array(polys)
        _lib.check(L.ph_dcrt_fwd_mul_inv_device(tab.handle, buf.ptr, polys.shape[0], op.handle, None))
        pa.synchronize()
        out = buf.download(polys.shape)
        buf.free()
        del op, tab
        return out
    finally:
        _lib._lib = saved


@pytest.mark.gpu
@pytest.mark.parametrize("log_n,per_thread", [(12, 8), (12, 16), (10, 8)])
def test_probe_outputs_and_log(log_n, per_thread):
    from primus_amd import tuning

    n = 1 << log_n
    moduli = pa.ntt_primes_below(50, 2 * n, 4)
    rng = np.random.default_rng(log_n * 100 + per_thread)
    count = 64
    polys = np.stack([rng.integers(0, moduli[i % 4], size=n, dtype=np.uint64) for i in range(count)])
    operand = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in moduli])
    shape = {8: 3, 16: 4}[per_thread]
    probe = ctypes.CDLL(_lib.PROBE_LIB_PATH)
    _lib._declare(probe)
    saved = _lib._lib
    try:
        for L in (_lib.lib(), probe):
            _lib._lib = L
            tuning.set_shape(log_n, "fwd_mul_inv", "f64", shape)
    finally:
        _lib._lib = saved
    want = _fused_batch(_lib.lib(), log_n, moduli, polys, operand)
    assert probe.ph_debug_clock_report(None, None, 1) == 0
    got = _fused_batch(probe, log_n, moduli, polys, operand)
    np.testing.assert_array_equal(got, want)
    slots, depth = 65536, 32
    raw = (ctypes.c_uint64 * (slots * depth * 4))()
    counts = (ctypes.c_uint32 * slots)()
    assert probe.ph_debug_clock_report(raw, counts, 1) == 0
    n = np.frombuffer(counts, dtype=np.uint32)
    waves = count * n_waves_per_poly(n_coeffs=1 << log_n, per_thread=per_thread)
    assert int(n.sum()) == waves and n.max() <= depth
    log = np.frombuffer(raw, dtype=np.uint64).reshape(slots, depth, 4)
    e = np.concatenate([log[s, :n[s]] for s in np.nonzero(n)[0]])
    # every workgroup logged each of its waves once, with both counters running
    assert sorted(np.bincount(e[:, 3].astype(np.int64), minlength=count).tolist()) == [waves // count] * count
    assert (e[:, 0] > 0).all() and (e[:, 1] > 0).all()


def n_waves_per_poly(n_coeffs, per_thread):
    return max(n_coeffs // per_thread // 64, 1)
