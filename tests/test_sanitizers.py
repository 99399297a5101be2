"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU).

Two programs, each built with -fsanitize=address,undefined and
-fno-sanitize-recover=all, so that any out-of-bounds access, leak, signed
overflow, misaligned access or shift past the width ends the run:

* tests/host/table_check.cpp: the table construction the library runs on the
  host (primus_amd/csrc/host_tables.hpp: primitive roots, Shoup twiddle tables,
  binary64 and register-order forms, Barrett constants), checked against the
  definitions of what it computes, N = 2^1 .. 2^17, moduli of 17 to 62 bits;
* tests/host/oracle_check.c: every entry point of the CPU oracle
  (oracle/primus_oracle.c) over edge sizes and moduli, checked by identities
  (round trips, schoolbook products, lazy ranges, CRT).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def _build_and_run(tmp_path, compiler, std, sources, name):
    cc = shutil.which(compiler)
    if cc is None:
        pytest.skip("%s not found" % compiler)
    exe = str(tmp_path / name)
    build = subprocess.run([cc, std, "-Wall", "-Wextra", "-pthread"] + SAN + ["-o", exe] + sources + ["-lm"],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-3000:]
    # verify_asan_link_order=0: a library preloaded into every process may come
    # before the sanitizer runtime without changing what it checks here
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, (run.stdout[-1000:], run.stderr[-4000:])
    assert run.stdout.startswith("ok"), run.stdout[-1000:]
    assert "runtime error" not in run.stderr and "ERROR: AddressSanitizer" not in run.stderr, run.stderr[-4000:]
    return int(run.stdout.split()[1])


def test_table_construction_sanitized(tmp_path):
    n = _build_and_run(tmp_path, "g++", "-std=c++17", [os.path.join(HERE, "host", "table_check.cpp")],
                       "table_check")
    assert n > 1_000_000


def test_oracle_sanitized(tmp_path):
    n = _build_and_run(tmp_path, "gcc", "-std=c11",
                       [os.path.join(HERE, "host", "oracle_check.c"), os.path.join(REPO, "oracle", "primus_oracle.c")],
                       "oracle_check")
    assert n > 1_000_000
