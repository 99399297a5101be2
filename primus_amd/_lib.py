"""Loader and build step of libprimus_hip.so (the C ABI in include/primus_hip.h).

There is no CPU fallback: if the library is missing, loading raises.
"""
import ctypes
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libprimus_hip.so")
# debug variant: every transform kernel checks its values' ranges at round
# boundaries (PH_RANGE_CHECK=1, read back with ph_debug_range_report)
RC_LIB_PATH = os.path.join(PKG_DIR, "libprimus_hip_rc.so")
# measurement variant: the single-pass chunk kernels log each wave's end with
# the shader and 100 MHz clocks (PH_CLOCK_PROBE=1, read back with
# ph_debug_clock_report)
PROBE_LIB_PATH = os.path.join(PKG_DIR, "libprimus_hip_probe.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "primus_hip.h")
CSRC = os.path.join(PKG_DIR, "csrc")
SOURCES = ["capi.hip", "ntt_kernels.hip", "fhe_kernels.hip", "rns_kernels.hip", "copy_kernels.hip", "serving.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep packed-FP32 instructions out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the binary64 modular arithmetic relies on every product and FMA being
    # exactly the one written: no contraction of separate multiplies and adds
    "-ffp-contract=off",
]


def build(verbose=False):
    """Compile every HIP source into primus_amd/libprimus_hip.so for gfx950,
    the same sources with PH_RANGE_CHECK=1 into libprimus_hip_rc.so, and
    ntt_kernels.hip with PH_CLOCK_PROBE=1 (linked with the default build's
    other objects) into libprimus_hip_probe.so.

    Each (variant, source) is compiled to an object in parallel (one hipcc
    each), then each variant's objects are linked; every call recompiles
    everything."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    objdir0 = os.path.join(PKG_DIR, "_build")
    # (library, object folder, defines, sources compiled for it; the rest
    # come from the default build)
    variants = [(LIB_PATH, objdir0, [], srcs),
                (RC_LIB_PATH, os.path.join(objdir0, "rc"), ["-DPH_RANGE_CHECK=1"], srcs),
                (PROBE_LIB_PATH, os.path.join(objdir0, "probe"), ["-DPH_CLOCK_PROBE=1"],
                 [os.path.join(CSRC, "ntt_kernels.hip")])]
    procs = []
    for lib_path, objdir, defs, own in variants:
        os.makedirs(objdir, exist_ok=True)
        for src in own:
            obj = os.path.join(objdir, os.path.basename(src) + ".o")
            cmd = [hipcc] + compile_flags + defs + ["-pthread", "-c", "-o", obj, src]
            procs.append((lib_path, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                                          text=True)))
    objs, errors = {}, []
    for lib_path, obj, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errors.append(out)
        elif verbose:
            print(out)
        objs.setdefault(lib_path, []).append(obj)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    for lib_path, _, _, _ in variants:
        names = {os.path.basename(o) for o in objs[lib_path]}
        shared = [os.path.join(objdir0, os.path.basename(s) + ".o") for s in srcs
                  if os.path.basename(s) + ".o" not in names]
        tmp = lib_path + ".tmp"
        # only the ph_* entry points of the header are exported (exports.map)
        link = ["-pthread", "-Wl,--version-script=" + os.path.join(CSRC, "exports.map"), "-o", tmp]
        res = subprocess.run([hipcc] + HIPCC_FLAGS + link + objs[lib_path] + shared, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("hipcc link failed:\n" + res.stdout + res.stderr)
        os.replace(tmp, lib_path)
    return LIB_PATH


_lib = None


class PrimusHipError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__("primus_hip status %d (%s)%s" % (code, _status_string(code), (": " + msg) if msg else ""))
        self.code = code


def _status_string(code):
    try:
        return lib().ph_status_string(code).decode()
    except Exception:  # pragma: no cover
        return "?"


def lib():
    """The loaded libprimus_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libprimus_hip.so is not built (%s); run __graft_entry__.build() first" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def use_probe():
    """Routes this process's entry points to the clock-probe variant
    (measurement runs only; call before any other use of the package)."""
    global _lib
    if _lib is not None:
        raise RuntimeError("the library is already loaded in this process")
    if not os.path.exists(PROBE_LIB_PATH):
        raise RuntimeError("libprimus_hip_probe.so is not built (%s); run __graft_entry__.build() first"
                           % PROBE_LIB_PATH)
    L = ctypes.CDLL(PROBE_LIB_PATH)
    _declare(L)
    _lib = L
    return L


_rc_lib = None


def lib_rc():
    """The range-checking debug variant, libprimus_hip_rc.so (same entry points)."""
    global _rc_lib
    if _rc_lib is None:
        if not os.path.exists(RC_LIB_PATH):
            raise RuntimeError("libprimus_hip_rc.so is not built (%s); run __graft_entry__.build() first"
                               % RC_LIB_PATH)
        L = ctypes.CDLL(RC_LIB_PATH)
        _declare(L)
        _rc_lib = L
    return _rc_lib


def check(code):
    if code != 0:
        msg = lib().ph_last_error().decode(errors="replace")
        raise PrimusHipError(code, msg)


def _declare(L):
    u64, u32, sz, vp, i = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int
    p64 = ctypes.POINTER(ctypes.c_uint64)
    pvp = ctypes.POINTER(ctypes.c_void_p)
    sigs = {
        "ph_status_string": (ctypes.c_char_p, [i]),
        "ph_last_error": (ctypes.c_char_p, []),
        "ph_debug_range_report": (i, [p64, i]),
        "ph_debug_clock_report": (i, [p64, ctypes.POINTER(ctypes.c_uint32), i]),
        "ph_tuning_shapes": (i, [u32, i, ctypes.POINTER(ctypes.c_int), i]),
        "ph_ntt_table_new": (i, [u32, u64, pvp]),
        "ph_ntt_table_new_ex": (i, [u32, u64, u32, pvp]),
        "ph_ntt_table_arith": (i, [vp]),
        "ph_dcrt_table_new_ex": (i, [u32, p64, u32, u32, pvp]),
        "ph_dcrt_table_arith": (i, [vp]),
        "ph_ntt_table_free": (None, [vp]),
        "ph_ntt_table_poly_length": (sz, [vp]),
        "ph_ntt_table_log_n": (u32, [vp]),
        "ph_ntt_table_modulus": (u64, [vp]),
        "ph_ntt_table_root": (u64, [vp]),
        "ph_ntt_table_inv_root": (u64, [vp]),
        "ph_ntt_table_inv_n": (i, [vp, p64, p64]),
        "ph_ntt_table_powers": (i, [vp, i, p64, p64]),
        "ph_lazy_transform_slice": (i, [vp, p64, sz]),
        "ph_transform_slice": (i, [vp, p64, sz]),
        "ph_lazy_inverse_transform_slice": (i, [vp, p64, sz]),
        "ph_inverse_transform_slice": (i, [vp, p64, sz]),
        "ph_transform_monomial": (i, [vp, u64, sz, p64, sz]),
        "ph_transform_coeff_one_monomial": (i, [vp, sz, p64, sz]),
        "ph_transform_coeff_minus_one_monomial": (i, [vp, sz, p64, sz]),
        "ph_dcrt_table_new": (i, [u32, p64, u32, pvp]),
        "ph_dcrt_table_free": (None, [vp]),
        "ph_dcrt_poly_length": (sz, [vp]),
        "ph_dcrt_moduli_count": (sz, [vp]),
        "ph_dcrt_crt_poly_length": (sz, [vp]),
        "ph_dcrt_ntt_table": (vp, [vp, u32]),
        "ph_dcrt_lazy_transform_slice": (i, [vp, p64, sz]),
        "ph_dcrt_transform_slice": (i, [vp, p64, sz]),
        "ph_dcrt_lazy_inverse_transform_slice": (i, [vp, p64, sz]),
        "ph_dcrt_inverse_transform_slice": (i, [vp, p64, sz]),
        "ph_dcrt_transform_device": (i, [vp, vp, sz, i, vp]),
        "ph_dcrt_inverse_transform_device": (i, [vp, vp, sz, vp]),
        "ph_ntt_transform_device": (i, [vp, vp, sz, i, vp]),
        "ph_ntt_inverse_transform_device": (i, [vp, vp, sz, vp]),
        "ph_operand_new": (i, [vp, p64, pvp]),
        "ph_operand_free": (None, [vp]),
        "ph_dcrt_fwd_mul_inv_device": (i, [vp, vp, sz, vp, vp]),
        "ph_dcrt_negacyclic_mul_device": (i, [vp, vp, vp, sz, vp]),
        "ph_dcrt_negacyclic_mul_device_noalloc": (i, [vp, vp, vp, sz, vp]),
        "ph_dcrt_negacyclic_mul_rows_device": (i, [vp, vp, vp, vp, sz, vp]),
        "ph_hexl_compute_forward": (i, [vp, vp, sz, ctypes.c_uint64, ctypes.c_uint64]),
        "ph_ext_key_rns_new": (i, [vp, ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.POINTER(vp)]),
        "ph_ext_key_rns_free": (None, [vp]),
        "ph_ext_product_rns_device": (i, [vp, vp, vp, sz, ctypes.c_int, vp]),
        "ph_hexl_compute_inverse": (i, [vp, vp, sz, ctypes.c_uint64, ctypes.c_uint64]),
        "ph_dcrt_pointwise_mul_device": (i, [vp, vp, vp, vp, sz, vp]),
        "ph_dcrt_mul_factor_device": (i, [vp, vp, sz, p64, u32, i, vp]),
        "ph_dcrt_pointwise_mul_add_device": (i, [vp, vp, vp, vp, sz, vp]),
        "ph_dcrt_automorphism_device": (i, [vp, vp, vp, sz, u64, vp]),
        "ph_base_converter_new": (i, [p64, u32, p64, u32, pvp]),
        "ph_base_converter_free": (None, [vp]),
        "ph_fast_convert_array_device": (i, [vp, vp, vp, u32, sz, vp]),
        "ph_exact_convert_array_device": (i, [vp, vp, vp, u32, sz, vp]),
        "ph_device_alloc": (i, [sz, pvp]),
        "ph_device_free": (i, [vp]),
        "ph_memcpy_htod": (i, [vp, vp, sz]),
        "ph_memcpy_dtoh": (i, [vp, vp, sz]),
        "ph_stream_synchronize": (i, [vp]),
        "ph_ext_key_new": (i, [vp, u32, u32, p64, pvp]),
        "ph_ext_key_new_digits": (i, [vp, u32, u32, i, p64, pvp]),
        "ph_ext_key_free": (None, [vp]),
        "ph_ext_product_device": (i, [vp, vp, vp, sz, i, vp]),
        "ph_ext_product": (i, [vp, p64, p64, sz, i]),
        "ph_rlwe_ctx_new": (i, [u32, u32, u32, p64, pvp]),
        "ph_rlwe_ctx_free": (None, [vp]),
        "ph_rlwe_ctx_modulus": (u64, [vp, u32, u32]),
        "ph_rlwe_encrypt_device": (i, [vp, vp, u32, vp, vp, vp, vp, vp, vp]),
        "ph_rlwe_plan_new": (i, [vp, vp, u32, pvp]),
        "ph_rlwe_plan_free": (None, [vp]),
        "ph_rlwe_encrypt_plan_device": (i, [vp, vp, vp, vp, vp, vp, vp]),
        "ph_rlwe_plan_autotune": (i, [vp, vp, vp, vp, vp, vp, i, ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_float)]),
        "ph_rlwe_plan_grid": (u32, [vp]),
        "ph_rlwe_sample_device": (i, [vp, vp, u32, u64, vp, vp, vp, vp, vp]),
        "ph_rlwe_sample_plan_device": (i, [vp, u64, vp, vp, vp, vp, vp]),
        "ph_host_alloc": (i, [sz, pvp]),
        "ph_host_free": (i, [vp]),
        "ph_tuning_set": (i, [u32, i, i, i]),
        "ph_tuning_get": (i, [u32, i, i]),
        "ph_tuning_save": (i, [ctypes.c_char_p]),
        "ph_tuning_load": (i, [ctypes.c_char_p]),
        "ph_autotune": (i, [vp, i, sz, i, ctypes.POINTER(i), ctypes.POINTER(ctypes.c_float)]),
    }
    for name, (res, args) in sigs.items():
        if not hasattr(L, name):
            continue
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def header_functions():
    """Names of the functions include/primus_hip.h declares."""
    import re

    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ph_[a-z0-9_]+)\s*\(", text)))
