"""ctypes binding of libspc5_hip.so (the C ABI declared in include/spc5_hip.h).

The library is required: importing a function that needs it raises if the
shared object is missing or fails to load.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SPC5_HIP_LIB selects another build of the library (e.g. a tuning variant).
HIP_SO = os.environ.get("SPC5_HIP_LIB") or os.path.join(HERE, "libspc5_hip.so")

SPC5_HIP_OK = 0
SPC5_HIP_F64 = 0
SPC5_HIP_F32 = 1
SPC5_HIP_BF16 = 2
SPC5_HIP_OPT_ROW_SORT = 1
SPC5_HIP_OPT_SPMM_COLUMNS = 2
SPC5_HIP_OPT_TILE_VALUES = 3
SPC5_HIP_OPT_SPMM_X16 = 4

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_L = C.c_long

# (name, restype, argtypes) of every C entry point declared in include/spc5_hip.h
SIGNATURES = {
    # reference ABI
    **{f"core_SPC5_{r}rVc_Spmv_{t}": (None, [_L, _P, _P, _P, _P, _P])
       for r in (1, 2, 4, 8) for t in ("double", "float")},
    "SPC5_opti_merge_double": (None, [_P, _P, _I]),
    "SPC5_opti_merge_float": (None, [_P, _P, _I]),
    # device API
    "spc5_hip_last_error": (C.c_char_p, []),
    "spc5_hip_version": (C.c_char_p, []),
    "spc5_hip_mat_create": (_I, [_I, _I, _I, _I64, _I64, _I64, _I64, _P, _P, _P, _P, C.POINTER(_P)]),
    "spc5_hip_mat_create_values_view": (_I, [_P, _I, _P, C.POINTER(_P)]),
    "spc5_hip_mat_destroy": (_I, [_P]),
    "spc5_hip_mat_info": (_I, [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64)]),
    "spc5_hip_mat_set_option": (_I, [_P, _I, _I64]),
    "spc5_hip_coo_to_csr": (_I, [_I, _I64, _I64, _I64, _I, _P, _P, _P, _P, _P, _P, _P]),
    "spc5_hip_mat_save": (_I, [_P, C.c_char_p]),
    "spc5_hip_mat_load": (_I, [C.c_char_p, C.POINTER(_P)]),
    "spc5_hip_mm_info": (_I, [C.c_char_p, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64)]),
    "spc5_hip_mm_read": (_I, [C.c_char_p, _I64, _P, _P, _P, C.POINTER(_I64)]),
    "spc5_hip_spmv": (_I, [_P, _P, _P, _P]),
    "spc5_hip_convert_count": (_I, [_I, _I, _I64, _P, _P, _P, C.POINTER(_I64), _P]),
    "spc5_hip_convert_count_noalloc": (_I, [_I, _I, _I64, _P, _P, _P, C.POINTER(_I64), _P]),
    "spc5_hip_convert_fill": (_I, [_I, _I, _I, _I, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "spc5_hip_host_block_count": (_I, [_I, _I, _I64, _I64, _P, _P, C.POINTER(_I64)]),
    "spc5_hip_host_convert": (_I, [_I, _I, _I, _I, _I64, _I64, _P, _P, _P, _P, _P, _I64, _P, C.POINTER(_I64)]),
    "spc5_hip_mat_create_from_host": (_I, [_I, _I, _I, _I64, _I64, _I64, _I64, _P, _P, _P, C.POINTER(_P)]),
    "spc5_hip_spmv_host": (_I, [_P, _P, _P]),
    "spc5_hip_spmv_ex": (_I, [_P, C.c_double, _P, C.c_double, _P, _P]),
    "spc5_hip_values_convert": (_I, [_I, _P, _P, _I64, _P]),
    "spc5_hip_cg": (_I, [_P, _P, _P, C.c_double, _I, _P, _P]),
    "spc5_hip_ir_cg": (_I, [_P, _P, _P, _P, C.c_double, C.c_double, _I, _I, _P, _P]),
    "spc5_hip_mixed_cg": (_I, [_P, _P, _P, _P, C.c_double, C.c_double, _I, _I, C.c_double, _P, _P]),
    "spc5_hip_spmm": (_I, [_P, _I, _I, C.c_double, _P, _I64, C.c_double, _P, _I64, _P]),
    "spc5_hip_synth_banded_count": (_I, [_I64, _I64, C.c_uint64, C.c_uint64, _P, C.POINTER(_I64), _P]),
    "spc5_hip_synth_banded_fill": (_I, [_I64, _I64, C.c_uint64, C.c_uint64, _P, _P, _P, _P]),
    "spc5_hip_stream_probe": (_I, [_P, _I64, _P, _I64, _I64, _P]),
    "spc5_hip_access_probe": (_I, [_P, _I, _P, _P, _P]),
    "spc5_hip_batch_create": (_I, [_I, _P, _P, _P, _I, C.POINTER(_P)]),
    "spc5_hip_batch_spmv": (_I, [_P, _P]),
    "spc5_hip_batch_spmv_direct": (_I, [_P, _P]),
    "spc5_hip_batch_info": (_I, [_P, C.POINTER(_I64), C.POINTER(_I)]),
    "spc5_hip_batch_destroy": (_I, [_P]),
}


class SolveStats(C.Structure):
    """spc5_hip_solve_stats"""
    _fields_ = [("iterations", C.c_int), ("outer", C.c_int), ("relres", C.c_double),
                ("tail_iterations", C.c_int), ("reserved", C.c_int)]

_lib = None


class Spc5Error(RuntimeError):
    pass


def lib():
    """The loaded HIP library.  Raises if it is not built or cannot load."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_SO):
            raise Spc5Error(f"{HIP_SO} is missing: build it with `python -m spc5_amd.build`")
        L = C.CDLL(HIP_SO)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != SPC5_HIP_OK:
        msg = lib().spc5_hip_last_error().decode(errors="replace")
        raise Spc5Error(f"{what} failed (status {rc}): {msg}")
