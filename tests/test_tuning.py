"""The launch-shape tuning table on the host: the shapes the autotuner tries
per degree and operation, validation of every shape field, the text file
format, and the tuning file PH_TUNING_FILE names being loaded when the
process creates its first table.  GPU cases (every shape's outputs) are in
test_gpu_ntt.py."""
import os
import subprocess
import sys
import textwrap

import pytest

from primus_amd import _lib, tuning

INVALID = 5  # PH_ERR_INVALID
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_shapes_per_degree():
    s = tuning.STREAM
    assert tuning.shapes(12, "fwd_mul_inv") == [4] + [4 | s | tuning.lpb(k) for k in range(4)] + \
        [3] + [3 | s | tuning.lpb(k) for k in range(4)]
    # streaming applies to forward-first operations only
    assert tuning.shapes(12, "inverse") == [4, 3]
    assert tuning.shapes(12, "negacyclic_mul") == [4, 3]
    # below 2^10 one shape; two-pass degrees: LDS tile x radix, at most 32 columns per chunk
    assert tuning.shapes(9, "forward") == [4]
    assert tuning.shapes(14, "forward") == [e | tuning.tile(c) for c in (11, 12, 13) for e in (4, 3)]
    assert tuning.shapes(16, "inverse") == [e | tuning.tile(c) for c in (11, 12, 13) for e in (4, 3)]
    assert tuning.shapes(17, "fwd_mul_inv") == [e | tuning.tile(c) for c in (12, 13) for e in (4, 3)]


def test_every_listed_shape_is_accepted():
    try:
        for log_n in range(1, 18):
            for op in tuning.OPS:
                for shape in tuning.shapes(log_n, op):
                    tuning.set_shape(log_n, op, "int", shape)
                    assert tuning.get_shape(log_n, op, "int") == shape
    finally:
        for log_n in range(1, 18):
            for op in tuning.OPS:
                tuning.set_shape(log_n, op, "int", 0)


@pytest.mark.parametrize("log_n,shape", [
    (12, tuning.tile(12)),                 # a two-pass tile at a single-pass degree
    (17, 4 | tuning.tile(11)),             # 64 columns per chunk
    (14, tuning.tile(10)),                 # tile outside 11..13
    (12, 3 | tuning.lpb(2)),               # limb-polynomials per workgroup without streaming
    (12, 5),                               # 32 coefficients per thread
    (12, 1 << 7),                          # an unused bit
    (12, -1),
])
def test_invalid_shapes_rejected(log_n, shape):
    assert _lib.lib().ph_tuning_set(log_n, 3, 0, shape) == INVALID
    assert tuning.get_shape(log_n, "fwd_mul_inv", "int") == 0


def test_save_load_round_trip(tmp_path):
    path = str(tmp_path / "tuning.txt")
    entries = {(12, "fwd_mul_inv", "f64"): 3 | tuning.STREAM | tuning.lpb(1),
               (15, "negacyclic_mul", "int"): 3 | tuning.tile(13),
               (10, "forward", "int"): 4}
    try:
        for (log_n, op, ar), shape in entries.items():
            tuning.set_shape(log_n, op, ar, shape)
        tuning.save(path)
        for (log_n, op, ar) in entries:
            tuning.set_shape(log_n, op, ar, 0)
        tuning.load(path)
        for (log_n, op, ar), shape in entries.items():
            assert tuning.get_shape(log_n, op, ar) == shape
    finally:
        for (log_n, op, ar) in entries:
            tuning.set_shape(log_n, op, ar, 0)


def _fresh_process(code, env):
    return subprocess.run([sys.executable, "-c", textwrap.dedent(code)], cwd=ROOT, env=env,
                          capture_output=True, text=True, timeout=120)


def test_tuning_file_loaded_at_first_table(tmp_path):
    """The file is read when the first table is created: its entries are in
    the table afterwards (the creation itself then needs a device), and one
    invalid entry makes creation fail before anything touches the device."""
    good = tmp_path / "good.txt"
    good.write_text("# primus_hip tuning: log_n op arith shape\n12 3 1 51\n15 4 0 %d\n" % (3 | (13 << 8)))
    bad = tmp_path / "bad.txt"
    bad.write_text("12 3 1 51\n12 3 1 %d\n" % (1 << 8 | 12 << 8))
    code = """
        import ctypes, sys
        import primus_amd as pa
        from primus_amd import _lib
        L = _lib.lib()
        before = (L.ph_tuning_get(12, 3, 1), L.ph_tuning_get(15, 4, 0))
        out = ctypes.c_void_p()
        q = (ctypes.c_uint64 * 1)(576460752303421441)
        rc = L.ph_dcrt_table_new_ex(12, q, 1, 0, ctypes.byref(out))
        print(before, (L.ph_tuning_get(12, 3, 1), L.ph_tuning_get(15, 4, 0)), rc)
    """
    env = dict(os.environ, PH_TUNING_FILE=str(good))
    res = _fresh_process(code, env)
    assert res.returncode == 0, res.stderr[-2000:]
    assert res.stdout.startswith("(0, 0) (51, %d)" % (3 | (13 << 8)))
    assert int(res.stdout.split()[-1]) != INVALID
    env["PH_TUNING_FILE"] = str(bad)
    res = _fresh_process(code, env)
    assert res.returncode == 0, res.stderr[-2000:]
    assert res.stdout.split()[-1] == str(INVALID)
    # a file that does not exist yet is no error
    env["PH_TUNING_FILE"] = str(tmp_path / "missing.txt")
    res = _fresh_process(code, env)
    assert res.returncode == 0 and res.stdout.startswith("(0, 0) (0, 0)")
    assert int(res.stdout.split()[-1]) != INVALID
