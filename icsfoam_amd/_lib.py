"""ctypes binding of libicsf.so (the C ABI declared in include/icsf.h).

The shared library is the product: every solver entry point below runs the
HIP kernels.  There is no CPU fallback; if the library is missing the import
of this module's loader raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libicsf.so")

ICSF_OK = 0
ERR_ARG, ERR_KEYWORD, ERR_ZERO_DIAG, ERR_NONFINITE, ERR_HIP, ERR_STATE, ERR_NODEVICE = -1, -2, -3, -4, -5, -6, -7


class IcsfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class KeywordError(IcsfError):
    pass


class ZeroDiagonalError(IcsfError):
    pass


class NoDeviceError(IcsfError):
    pass


_EXC = {ERR_KEYWORD: KeywordError, ERR_ZERO_DIAG: ZeroDiagonalError, ERR_NODEVICE: NoDeviceError}

_vp = C.c_void_p
_i = C.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "icsf_last_error": (C.c_char_p, []),
    "icsf_version": (_i, []),
    "icsf_rcm_order": (_i, [_i, _i, _vp, _vp, _vp]),
    "icsf_rcm_order_coupled": (_i, [_i, _i, _vp, _vp, _i, _vp, _vp, _vp]),
    "icsf_mesh_geometry_patches": (_i, [_i, _vp, _i, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp,
                                        _vp, _vp, _vp, _vp, _vp]),
    "icsf_get_interface_blocks": (_i, [_vp, _vp]),
    "icsf_get_tuning": (_i, [_vp, _vp]),
    "icsf_set_tuning": (_i, [_vp, _vp]),
    "icsf_set_mrf": (_i, [_vp, _i, _vp, _vp, _vp, C.c_double, _i, _vp]),
    "icsf_hb_set_mrf": (_i, [_vp, _i, _vp, _vp, _vp, C.c_double, _i, _vp]),
    "icsf_set_translating_frame": (_i, [_vp, _i, _vp, _vp, _i, _vp]),
    "icsf_hb_set_translating_frame": (_i, [_vp, _i, _vp, _vp, _i, _vp]),
    "icsf_hb_phase_lag_operator": (_i, [_i, _vp, _vp, _i, _vp, C.c_double, _vp]),
    "icsf_hb_set_phase_lag": (_i, [_vp, _i, _vp, _i, _vp, _vp]),
    "icsf_colour_cells": (_i, [_i, _i, _vp, _vp, _vp, _vp]),
    "icsf_colour_major_order": (_i, [_i, _vp, _i, _vp]),
    "icsf_renumber_mesh": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_sweep_levels": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_block_mesh_hex": (_i, [_i, _i, _i, _vp, _i, _vp, _vp, _vp, _vp]),
    "icsf_mesh_geometry": (_i, [_i, _vp, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_mesh_create": (_i, [_i, _vp, _i, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_mesh_destroy": (_i, [_vp]),
    "icsf_mesh_info": (_i, [_vp, _vp]),
    "icsf_case_create": (_i, [_vp, C.c_char_p, C.c_char_p, C.c_char_p, _vp]),
    "icsf_case_destroy": (_i, [_vp]),
    "icsf_set_boundary": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp]),
    "icsf_set_boundary_values": (_i, [_vp, _i, _vp, _vp, _vp]),
    "icsf_set_state": (_i, [_vp, _vp]),
    "icsf_pseudo_time_loop": (_i, [_vp, _i, _i, _vp, _vp]),
    "icsf_get_boundary_values": (_i, [_vp, _vp, _vp, _vp]),
    "icsf_set_fields": (_i, [_vp, _vp, _vp, _vp]),
    "icsf_get_fields": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_calc_flux": (_i, [_vp, _vp, _vp, _vp]),
    "icsf_residual": (_i, [_vp, _vp, _vp]),
    "icsf_hb_operator": (_i, [_i, _vp, _vp, _i, C.c_double, _vp, _vp, _vp]),
    "icsf_hb_create": (_i, [_vp, C.c_char_p, C.c_char_p, C.c_char_p, _i, _vp, _vp]),
    "icsf_hb_destroy": (_i, [_vp]),
    "icsf_hb_set_boundary": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp]),
    "icsf_hb_set_boundary_values": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "icsf_hb_reconstruct_weights": (_i, [_i, _vp, _vp, _i, _vp, C.c_double, _vp]),
    "icsf_hb_reconstruct": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_hb_set_fields": (_i, [_vp, _i, _vp, _vp, _vp]),
    "icsf_hb_get_fields": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "icsf_hb_residual": (_i, [_vp, _vp]),
    "icsf_hb_assemble_jacobian": (_i, [_vp]),
    "icsf_hb_matrix_mul": (_i, [_vp, _vp, _vp]),
    "icsf_hb_precondition": (_i, [_vp, _vp]),
    "icsf_hb_solve_for_incr": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "icsf_hb_pseudo_steps": (_i, [_vp, _i, _vp]),
    "icsf_hb_synchronize": (_i, [_vp]),
    "icsf_residual_vjp": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_residual_vjp_device": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_residual_vjp_device_noalloc": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_residual_device": (_i, [_vp, _vp, _vp]),
    "icsf_residual_records_device": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "icsf_assemble_jacobian": (_i, [_vp]),
    "icsf_get_jacobian": (_i, [_vp, _vp, _vp, _vp]),
    "icsf_set_jacobian": (_i, [_vp, _vp, _vp, _vp]),
    "icsf_matrix_mul": (_i, [_vp, _vp, _vp]),
    "icsf_precondition": (_i, [_vp, _vp]),
    "icsf_solve_for_incr": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "icsf_pseudo_steps": (_i, [_vp, _i, _vp, _vp]),
    "icsf_get_pseudo_co": (_i, [_vp, _vp]),
    "icsf_synchronize": (_i, [_vp]),
    "icsf_foam_mesh_read": (_i, [C.c_char_p, _vp]),
    "icsf_foam_mesh_sizes": (_i, [_vp, _vp]),
    "icsf_foam_mesh_get": (_i, [_vp] + [_vp] * 10),
    "icsf_foam_mesh_patch_name": (_i, [_vp, _i, C.c_char_p, _i]),
    "icsf_foam_mesh_patch_phase_lag": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, C.c_char_p, _i]),
    "icsf_foam_mesh_free": (_i, [_vp]),
    "icsf_foam_mesh_write": (_i, [C.c_char_p, _i, _vp, _i, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp,
                                  _vp]),
    "icsf_foam_field_read": (_i, [C.c_char_p, _i, _vp]),
    "icsf_foam_field_internal": (_i, [_vp, _vp, _vp]),
    "icsf_foam_field_patch_type": (_i, [_vp, C.c_char_p, C.c_char_p, _i]),
    "icsf_foam_field_patch_entry": (_i, [_vp, C.c_char_p, C.c_char_p, _i, _vp]),
    "icsf_foam_field_free": (_i, [_vp]),
    "icsf_foam_field_write": (_i, [C.c_char_p, C.c_char_p, C.c_char_p, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "icsf_timing_enable": (_i, [_vp, _i]),
    "icsf_timing_read": (_i, [_vp, _i, _vp, _vp]),
}

_lib = None


def lib():
    """Load libicsf.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m icsfoam_amd.build` "
                              "(the solver has no CPU fallback)")
        # torch bundles its own HIP runtime under the same soname
        # (libamdhip64.so.7).  Importing torch first makes the dynamic loader
        # bind libicsf.so to that already-loaded runtime, so that one HIP
        # runtime serves both and torch device tensors can be handed to the
        # library; loading libicsf.so first would bring in a second runtime
        # and torch's device initialisation would then fail.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(rc: int) -> None:
    if rc != ICSF_OK:
        msg = lib().icsf_last_error().decode(errors="replace")
        raise _EXC.get(rc, IcsfError)(rc, msg)


def ptr(a):
    """Data pointer of a contiguous numpy array (None passes NULL).

    The returned ctypes pointer holds a reference to the array, so a
    temporary made by a dtype or contiguity conversion in the argument list
    (``ptr(as_i32(x))``) stays alive until the C call that receives the
    pointer has returned."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)


def as_i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)
