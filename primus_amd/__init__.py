"""primus_amd: Primus's negacyclic NTT and the modular arithmetic under it on AMD MI355X (gfx950).

The product is libprimus_hip.so (C ABI in include/primus_hip.h); this package
is its host-side mirror of the reference's NttTable / DcrtTable interface.
"""
from ._lib import LIB_PATH, PrimusHipError, build, header_functions, lib  # noqa: F401
from .ntt import (  # noqa: F401
    DegreeConversionErr,
    DegreeTooLarge,
    DeviceBuffer,
    NoPrimitiveRoot,
    NttError,
    NttTableErr,
    Operand,
    HexlNttTable,
    UintCrtNttTable,
    UintNttTable,
    automorphism_device,
    fwd_mul_inv_device,
    inverse_transform_device,
    mul_factor_device,
    negacyclic_mul_device,
    negacyclic_mul_rows_device,
    pointwise_mul_add_device,
    pointwise_mul_device,
    synchronize,
    transform_device,
)
from .primes import ntt_primes_below  # noqa: F401
