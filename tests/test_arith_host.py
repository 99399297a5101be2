"""The integer helpers of primus_amd/csrc/arith.hpp that every integer
butterfly, Barrett reduction and RLWE multiply-accumulate is built from,
checked on the host against plain 64- and 128-bit arithmetic: the chained
low product w*b - k*q (mod 2^64) through 2^64 - q, the borrow-selected
conditional subtraction, the approximate high product's error bound, and the
approximate-quotient Shoup product's contract (congruent to w*b, below 4q).
The check is a host-only build of tests/host/arith_check.hip (no GPU)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_integer_helpers_host_check(tmp_path):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "arith_check")
    src = os.path.join(HERE, "host", "arith_check.hip")
    build = subprocess.run([hipcc, "--offload-host-only", "-std=c++17", "-O2", "-o", exe, src],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and run.stdout.startswith("ok"), run.stdout[-2000:]
    assert int(run.stdout.split()[1]) > 10_000_000
