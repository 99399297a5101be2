"""CPU-side checks of the C ABI: the library loads, exports every function the
header declares, and rejects bad arguments before touching the GPU."""
import ctypes
import os

import pytest

import primus_amd as pa
from primus_amd import _lib


def test_library_built_and_loads():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"
    pa.lib()


def test_every_header_function_is_exported():
    names = pa.header_functions()
    assert len(names) >= 40
    L = pa.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def _exported_functions(path):
    """Defined function symbols of the library's dynamic symbol table."""
    import shutil
    import subprocess

    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm is not installed")
    out = subprocess.run([nm, "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    names = set()
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[1] in "TtWi":
            names.add(parts[2].split("@")[0])
    return names


def test_exported_set_equals_the_header():
    """Only the header's entry points leave the library: kernel stubs,
    launchers and other internals are local (linker version script)."""
    exported = _exported_functions(_lib.LIB_PATH)
    declared = set(pa.header_functions())
    assert exported == declared, {"not declared": sorted(exported - declared)[:20],
                                  "not exported": sorted(declared - exported)[:20]}


def test_status_strings():
    L = pa.lib()
    assert L.ph_status_string(0) == b"ok"
    assert L.ph_status_string(1) == b"no primitive root"
    assert L.ph_status_string(3) == b"degree too large"
    assert L.ph_status_string(99) == b"unknown status"


@pytest.mark.parametrize("q", [0, 1, 1 << 62, (1 << 64) - 1])
def test_bad_modulus_rejected_on_host(q):
    with pytest.raises(ValueError):
        pa.UintNttTable.new(10, q)


def test_no_primitive_root_on_host():
    # 2N = 8192 does not divide q - 1 = 12288: NttError::NoPrimitiveRoot
    with pytest.raises(pa.NoPrimitiveRoot):
        pa.UintNttTable.new(12, 12289)
    with pytest.raises(pa.NoPrimitiveRoot):
        pa.UintCrtNttTable.new(12, [12289, 1125899906826241])


@pytest.mark.parametrize("log_n", [0, 18, 30])
def test_unsupported_degree(log_n):
    with pytest.raises(pa.PrimusHipError) as e:
        pa.UintNttTable.new(log_n, 576460752303421441)
    assert e.value.code == 7


def test_null_arguments():
    L = pa.lib()
    out = ctypes.c_void_p()
    assert L.ph_ntt_table_new(10, 2147473409, None) == 5
    assert L.ph_transform_slice(None, None, 0) == 5
    assert L.ph_dcrt_table_new(10, None, 0, ctypes.byref(out)) == 5
    assert L.ph_ntt_table_poly_length(None) == 0


def test_null_arguments_batched_forms():
    """Batched entry points reject a null table or buffer before touching the
    device (status 5, PH_ERR_INVALID); a zero count with null buffers is a no-op
    only once the table is there, so these all fail on the table."""
    L = pa.lib()
    dummy = ctypes.c_void_p(0x1000)  # never dereferenced: the table check fails first
    assert L.ph_dcrt_negacyclic_mul_rows_device(None, dummy, dummy, dummy, 1, None) == 5
    assert L.ph_dcrt_negacyclic_mul_device(None, dummy, dummy, 1, None) == 5
    assert L.ph_dcrt_fwd_mul_inv_device(None, dummy, 1, None, None) == 5
    assert L.ph_rlwe_encrypt_device(None, None, 0, None, None, None, None, None, None) == 5


def test_hexl_factor_checks_on_host():
    """HEXL's range factors are checked before the table (no device needed)."""
    L = pa.lib()
    assert L.ph_hexl_compute_forward(None, None, 0, 3, 1) == 5
    assert L.ph_hexl_compute_forward(None, None, 0, 1, 2) == 5
    assert L.ph_hexl_compute_inverse(None, None, 0, 4, 1) == 5
    assert L.ph_hexl_compute_inverse(None, None, 0, 1, 4) == 5
    # valid factors reach the table check
    assert L.ph_hexl_compute_forward(None, None, 0, 4, 4) == 5
