"""ctypes binding of libffmm_hip.so (the C ABI declared in include/ffmm.h).

The library is the only compute path: loading fails loudly when it is missing, and
nothing here falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libffmm_hip.so")

FFMM_OK, FFMM_ERR_ARGUMENT, FFMM_ERR_DIMENSION, FFMM_ERR_DEVICE, FFMM_ERR_STATE, FFMM_ERR_MEMORY = range(6)
PREC_F64, PREC_MIXED = 0, 1
M2L_DENSE, M2L_SPECTRAL = 0, 1
POTENTIAL, GRADIENT, POTENTIAL_GRADIENT, PRESSURE = 0, 1, 2, 3
KERNEL_LAPLACE, KERNEL_YUKAWA, KERNEL_STOKESLET = 0, 1, 2


class ArgumentError(ValueError):
    """The reference's ArgumentError (invalid geometry, parameters or state)."""


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch (array sizes that do not agree)."""


class DeviceError(RuntimeError):
    """A HIP or rocBLAS failure inside the library."""


_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double

# name -> (restype, argtypes); every entry point of include/ffmm.h
SIGNATURES = {
    "ffmm_last_error": (ctypes.c_char_p, []),
    "ffmm_abi_version": (_i32, []),
    "ffmm_context_create": (_i32, [_i32, _vp, ctypes.POINTER(_vp)]),
    "ffmm_context_destroy": (_i32, [_vp]),
    "ffmm_bounds": (_i32, [_vp, _vp, _i64, _vp, _vp, ctypes.POINTER(_i32)]),
    "ffmm_direct": (_i32, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i32]),
    "ffmm_direct_noalloc": (_i32, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i32]),
    "ffmm_direct_stokes": (_i32, [_vp, _vp, _vp, _i32, _i64, _vp, _i64, _vp, _vp]),
    "ffmm_direct_stokes_pressure": (_i32, [_vp, _vp, _vp, _i32, _i64, _vp, _i64, _vp, _vp]),
    "ffmm_opcache_create": (_i32, [_vp, _i32, _i32, _d, ctypes.POINTER(_vp)]),
    "ffmm_opcache_destroy": (_i32, [_vp]),
    "ffmm_opcache_put_cube": (_i32, [_vp, _d, _vp, _vp, _i32, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _i32,
                                     _vp, _vp, _vp, _vp, _vp]),
    "ffmm_opcache_create_harmonic": (_i32, [_vp, _i32, _i32, _d, _d, ctypes.POINTER(_vp)]),
    "ffmm_opcache_put_harmonic_classes": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    "ffmm_opcache_put_harmonic_level": (_i32, [_vp, _d, _vp, _vp, _vp]),
    "ffmm_opcache_has_width": (_i32, [_vp, _d, ctypes.POINTER(_i32)]),
    "ffmm_opcache_put_class_factors": (_i32, [_vp, _d, _vp, _vp, _vp]),
    "ffmm_plan_set_tol": (_i32, [_vp, _d]),
    "ffmm_plan_capture": (_i32, [_vp, _i32, _i32]),
    "ffmm_circle_create": (_i32, [_vp, _i32, _vp, _i64, _i32, _vp, _d, _i32, ctypes.POINTER(_vp)]),
    "ffmm_circle_sizes": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i64), _vp, _vp]),
    "ffmm_circle_put_level": (_i32, [_vp, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    "ffmm_circle_build_expansions": (_i32, [_vp, _vp]),
    "ffmm_circle_evaluate_at_sources": (_i32, [_vp, _vp]),
    "ffmm_circle_evaluate": (_i32, [_vp, _vp, _i64, _vp]),
    "ffmm_circle_download_tree": (_i32, [_vp] * 11),
    "ffmm_circle_download_lists": (_i32, [_vp, _i32, _vp, _vp, _vp]),
    "ffmm_circle_destroy": (_i32, [_vp]),
    "ffmm_plan_replay": (_i32, [_vp, _vp, _vp]),
    "ffmm_plan_create": (_i32, [_vp, _vp, _i64, _i32, _vp, _d, _i32, _i32, _i32, _i32, ctypes.POINTER(_vp)]),
    "ffmm_plan_destroy": (_i32, [_vp]),
    "ffmm_plan_sizes": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i64), ctypes.POINTER(_i64), _vp, _vp, _vp,
                               _vp, _vp]),
    "ffmm_plan_download_tree": (_i32, [_vp] + [_vp] * 14),
    "ffmm_plan_perm": (_i32, [_vp, _vp]),
    "ffmm_plan_save": (_i32, [_vp, _vp, _i64, ctypes.POINTER(_i64)]),
    "ffmm_plan_load": (_i32, [_vp, _vp, _i64, ctypes.POINTER(_vp)]),
    "ffmm_plan_data_check": (_i32, [_vp, _i64]),
    "ffmm_plan_from_tree": (_i32, [_vp, _i32, _vp, _i64] + [_vp] * 11 + [_i32, _vp, _d, _i32, _i32, _i32,
                                                                       ctypes.POINTER(_vp)]),
    "ffmm_plan_download_lists": (_i32, [_vp, _i32, _vp, _vp, _vp]),
    "ffmm_build_expansions": (_i32, [_vp, _vp, _i32]),
    "ffmm_build_expansions_nd": (_i32, [_vp, _vp, _i32, _i32]),
    "ffmm_evaluate_at_sources": (_i32, [_vp, _vp, _i32]),
    "ffmm_evaluate": (_i32, [_vp, _vp, _i64, _vp, _i32]),
    "ffmm_evaluate_nodes": (_i32, [_vp, _vp, _i64, _vp, _vp, _i32]),
    "ffmm_matvec": (_i32, [_vp, _vp, _i32, _vp, _i32]),
    "ffmm_plan_download_expansions": (_i32, [_vp, _vp, _vp, _vp]),
    "ffmm_plan_stage_times": (_i32, [_vp, _vp, _i32]),
    "ffmm_plan_spectral_tiles": (_i32, [_vp, _vp, _vp, _i32]),
    "ffmm_server_create": (_i32, [_i32, _i32, _i64, ctypes.POINTER(_vp)]),
    "ffmm_server_submit": (_i32, [_vp, _vp, ctypes.POINTER(_i64)]),
    "ffmm_server_wait": (_i32, [_vp, _i64, ctypes.POINTER(_d)]),
    "ffmm_server_stats": (_i32, [_vp, _vp, _i32]),
    "ffmm_server_export": (_i32, [_vp, _vp, _i64, _vp, _i64, ctypes.POINTER(_i64)]),
    "ffmm_server_import": (_i32, [_vp, _vp, _i64, _vp, _i64]),
    "ffmm_server_destroy": (_i32, [_vp]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"libffmm_hip.so is not built ({LIBPATH}); run `python flexiblefmm.jl_amd/build.py`")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int):
    if rc == FFMM_OK:
        return
    msg = lib().ffmm_last_error().decode()
    if rc == FFMM_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if rc == FFMM_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if rc == FFMM_ERR_MEMORY:
        raise MemoryError(msg)
    raise DeviceError(f"[{rc}] {msg}")


def call(name: str, *args):
    check(getattr(lib(), name)(*args))
