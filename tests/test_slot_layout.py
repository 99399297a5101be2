"""The register-round slot maps (primus_amd/csrc/ntt_rounds.hpp), checked on
the host for every chunk geometry: each round's map is a bijection onto the
chunk, and a slot and its LDS position are a thread part plus a constant
register part, which the kernels' global and LDS addressing relies on.  The
check is a host-only build of tests/host/slot_layout_check.hip (no GPU)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_slot_layout_host_check(tmp_path):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "slot_layout_check")
    src = os.path.join(HERE, "host", "slot_layout_check.hip")
    build = subprocess.run([hipcc, "--offload-host-only", "-std=c++17", "-O1", "-o", exe, src],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and run.stdout.startswith("ok"), run.stdout[-2000:]
    assert int(run.stdout.split()[1]) > 1_000_000
