"""ctypes binding of ``libvi_mi355x.so`` (the C ABI in include/vi_mi355x.h).

The library is built in-tree by ``viscoelasticinversion_amd.build``. Importing this module
never compiles: a missing library raises ``ImportError`` so that no caller can silently fall
back to a CPU path.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libvi_mi355x.so")

VI_F32, VI_F64 = 0, 1
VI_OK, VI_EINVAL, VI_ESTATE, VI_EDEVICE, VI_ENOMEM = 0, -1, -2, -3, -4
(VI_VX, VI_VZ, VI_SXX, VI_SZZ, VI_SXZ, VI_MXX, VI_MZZ, VI_MXZ, VI_PSI_VX_X, VI_PSI_VZ_X,
 VI_PSI_SXX_X, VI_PSI_SXZ_X, VI_PSI_VX_Z, VI_PSI_VZ_Z, VI_PSI_SXZ_Z, VI_PSI_SZZ_Z) = range(16)
VI_SRC_FORCE_Z, VI_SRC_EXPLOSIVE, VI_SRC_FORCE_X = 0, 1, 2
VI_REC_VX, VI_REC_VZ, VI_REC_SXX, VI_REC_SZZ, VI_REC_SXZ = range(5)


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("nz", ctypes.c_int), ("dx", ctypes.c_double),
                ("dz", ctypes.c_double), ("dt", ctypes.c_double), ("precision", ctypes.c_int),
                ("nrelax", ctypes.c_int), ("fd_order_x", ctypes.c_int),
                ("fd_order_z", ctypes.c_int)]


class CpmlDesc(ctypes.Structure):
    _fields_ = [("thickness", ctypes.c_int), ("d0", ctypes.c_double),
                ("alpha_max", ctypes.c_double), ("kappa_max", ctypes.c_double),
                ("power", ctypes.c_double), ("left", ctypes.c_int), ("right", ctypes.c_int),
                ("top", ctypes.c_int), ("bottom", ctypes.c_int)]


class ViError(RuntimeError):
    """A non-zero status returned by the C ABI, with vi_last_error()'s message."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_V, _I, _D, _L = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
_IP, _DP = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
_H = ctypes.c_void_p

# name -> (restype, argtypes); mirrors include/vi_mi355x.h
SIGNATURES = {
    "vi_last_error": (ctypes.c_char_p, []),
    "vi_version": (ctypes.c_char_p, []),
    "vi_fd_coeffs": (None, [_D, _DP]),
    "vi_fd_weights": (_I, [_D, _I, _DP]),
    "vi_afda_weights": (_I, [_D, _IP, _DP]),
    "vi_afda_default_points": (_I, [_IP]),
    "vi_afda_hermitian_weights": (_I, [_D, _I, _DP]),
    "vi_cpml_auto_d0": (_D, [_I, _D, _D, _D, _D]),
    "vi_cpml_profile": (_I, [_I, ctypes.POINTER(CpmlDesc), _D, _I, _I, _DP, _DP, _DP]),
    "vi_device_count": (_I, []),
    "vi_mem_info": (_I, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "vi_cached_bytes": (_I, [ctypes.POINTER(ctypes.c_size_t)]),
    "vi_release_cached": (_I, []),
    "vi_model_create": (_I, [_I, _I, _I, _V, _V, _V, _V, ctypes.POINTER(_H)]),
    "vi_model_create_vti": (_I, [_I, _I, _I, _V, _V, _V, _V, _V, _V, ctypes.POINTER(_H)]),
    "vi_model_destroy": (_I, [_H]),
    "vi_set_model_window": (_I, [_H, _H, _I]),
    "vi_device_bytes": (_I, [_H, ctypes.POINTER(ctypes.c_size_t)]),
    "vi_batch_create": (_I, [ctypes.POINTER(_H)]),
    "vi_batch_destroy": (_I, [_H]),
    "vi_batch_set": (_I, [_H, ctypes.POINTER(_H), _I]),
    "vi_batch_set_tiling": (_I, [_H, _I]),
    "vi_batch_get_tiling": (_I, [_H]),
    "vi_batch_set_cache_block": (_I, [_H, ctypes.c_size_t]),
    "vi_batch_get_cache_block": (ctypes.c_size_t, [_H]),
    "vi_batch_set_streams": (_I, [_H, _I]),
    "vi_batch_autotune": (_I, [_H, _I, _IP, _DP]),
    "vi_batch_set_graph": (_I, [_H, _I]),
    "vi_batch_step": (_I, [_H, _I]),
    "vi_batch_sync": (_I, [_H]),
    "vi_batch_profile_enable": (_I, [_H, _I]),
    "vi_batch_profile_read": (_I, [_H, _I, _DP, ctypes.POINTER(_L)]),
    "vi_create": (_I, [ctypes.POINTER(GridDesc), ctypes.POINTER(_H)]),
    "vi_rates": (_I, [_H, _V, _V, _V, _V]),
    "vi_rates_noalloc": (_I, [_H, _V, _V, _V, _V]),
    "vi_trim": (_I, [_H]),
    "vi_check_finite": (_I, [_H, ctypes.POINTER(_I)]),
    "vi_destroy": (_I, [_H]),
    "vi_set_model": (_I, [_H, _V, _V, _V, _V]),
    "vi_set_model_vti": (_I, [_H, _V, _V, _V, _V, _V, _V]),
    "vi_set_relaxation": (_I, [_H, _DP]),
    "vi_set_cpml": (_I, [_H, ctypes.POINTER(CpmlDesc)]),
    "vi_set_free_surface": (_I, [_H, _I]),
    "vi_set_afda_points": (_I, [_H, _IP]),
    "vi_set_afda_hermitian": (_I, [_H, _I]),
    "vi_set_sources": (_I, [_H, _I, _IP, _IP, _IP, _I, _V]),
    "vi_set_receivers": (_I, [_H, _I, _IP, _IP, _IP, _I]),
    "vi_set_state": (_I, [_H, _I, _V]),
    "vi_get_state": (_I, [_H, _I, _V]),
    "vi_reset": (_I, [_H]),
    "vi_step": (_I, [_H, _I]),
    "vi_sync": (_I, [_H]),
    "vi_step_index": (_L, [_H]),
    "vi_get_traces": (_I, [_H, _V, _I]),
    "vi_gradient": (_I, [_H, _I, _V, _I, ctypes.POINTER(ctypes.c_double), _V, _V, _V, _V]),
    "vi_gradient_bytes": (_I, [_H, _I, _I, ctypes.POINTER(ctypes.c_size_t)]),
    "vi_gradient_begin": (_I, [_H, _I, _V, _I, ctypes.POINTER(_H)]),
    "vi_gradient_advance": (_I, [_H, _I, ctypes.POINTER(_I)]),
    "vi_gradient_idle": (_I, [_H]),
    "vi_gradient_end": (_I, [_H, ctypes.POINTER(ctypes.c_double), _V, _V, _V, _V, _V, _V]),
    "vi_gradient_cancel": (_I, [_H]),
    "vi_plan_bytes": (_I, [ctypes.POINTER(GridDesc), _I, _I, _I, _I, _I, _I, _I, _I,
                           ctypes.POINTER(ctypes.c_size_t)]),
    "vi_gradient_vti": (_I, [_H, _I, _V, _I, ctypes.POINTER(ctypes.c_double), _V, _V, _V, _V, _V,
                             _V]),
    "vi_set_tiling": (_I, [_H, _I]),
    "vi_get_tiling": (_I, [_H]),
    "vi_autotune": (_I, [_H, _I, _IP, _DP]),
    "vi_profile_enable": (_I, [_H, _I]),
    "vi_profile_read": (_I, [_H, _I, _DP, ctypes.POINTER(_L)]),
    "vi_algorithmic_bytes": (_I, [_H, _I, _DP]),
}

_lib = None


def lib():
    """Load the library once; raise ImportError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m viscoelasticinversion_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != VI_OK:
        msg = lib().vi_last_error()
        raise ViError(rc, msg.decode() if msg else "")
