"""ctypes binding of libfsgm.so (the C ABI declared in include/sgm_capi.h).

The library is built in-tree by ``fasterstereocuda_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback anywhere in the
product path.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfsgm.so")

SGM_OK = 0
SGM_E_INVALID_ARG = -1
SGM_E_HIP = -2
SGM_E_NOT_INIT = -3
SGM_E_UNSUPPORTED = -4
SGM_E_NO_DEVICE = -5

INVALID_FLOAT = -999999.0
INVALID_SHORT = -32768
LEVEL_RANGE = 32


class SgmOption(ctypes.Structure):
    """Field-for-field mirror of sgm_option (CuSGMOption)."""
    _fields_ = [
        ("p1", ctypes.c_int32),
        ("p2_init", ctypes.c_int32),
        ("p2_max", ctypes.c_int32),
        ("cs_type", ctypes.c_int32),
        ("num_paths", ctypes.c_int32),
        ("do_median_filter", ctypes.c_int32),
        ("do_lr_check", ctypes.c_int32),
        ("lr_threshold", ctypes.c_float),
        ("unique_threshold", ctypes.c_float),
        ("using_constant_p2", ctypes.c_int32),
        ("do_remove_peaks", ctypes.c_int32),
        ("peaks_ratio_threshold", ctypes.c_float),
        ("post_filter_type", ctypes.c_int32),
        ("morphology_type", ctypes.c_int32),
    ]


class SgmCamera(ctypes.Structure):
    """Field-for-field mirror of sgm_camera (CamParam_T, sgm_types.h:69-84)."""
    _fields_ = [("x0_left", ctypes.c_double), ("y0_left", ctypes.c_double),
                ("x0_right", ctypes.c_double), ("y0_right", ctypes.c_double),
                ("focus", ctypes.c_double), ("baseline", ctypes.c_double)]


class SgmRoi(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32),
                ("w", ctypes.c_int32), ("h", ctypes.c_int32)]


class SgmPairDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("min_disparity", ctypes.c_int32), ("disp_range", ctypes.c_int32),
                ("d_left", ctypes.c_void_p), ("d_right", ctypes.c_void_p),
                ("d_disp_left", ctypes.c_void_p), ("d_disp_right", ctypes.c_void_p)]


class SgmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"SGM error {code}: {message}")
        self.code = code


_lib = None

P = ctypes.c_void_p
I32 = ctypes.c_int32
F32 = ctypes.c_float
PP4 = P * 4

# name -> (restype, argtypes); the list is the C ABI of include/sgm_capi.h
SIGNATURES = {
    "sgm_create": (P, []),
    "sgm_destroy": (None, [P]),
    "sgm_init": (ctypes.c_int, [P, I32, I32, I32, I32, ctypes.POINTER(SgmOption), I32]),
    "sgm_match": (ctypes.c_int, [P, P, P, P, P, P, P, P]),
    "sgm_match_device": (ctypes.c_int, [P, P, P, P, P, P]),
    "sgm_depth_device": (ctypes.c_int, [P, I32, I32, ctypes.POINTER(SgmCamera), P, P]),
    "sgm_init2": (ctypes.c_int, [P, I32, I32, I32, I32, ctypes.POINTER(SgmOption),
                                 ctypes.POINTER(SgmCamera), I32]),
    "sgm_match2": (ctypes.c_int, [P, P, P, P, P, P, P]),
    "sgm_set_min_disparity": (ctypes.c_int, [P, I32]),
    "sgm_set_peak_strips": (ctypes.c_int, [I32]),
    "sgm_get_peak_strips": (I32, []),
    "sgm_stage_remove_peaks": (ctypes.c_int, [P, I32, I32, I32, F32, P, P]),
    "sgm_stage_remove_peaks_noalloc": (ctypes.c_int, [P, I32, I32, I32, F32, P]),
    "sgm_get_disparity_range": (I32, [P]),
    "sgm_release": (ctypes.c_int, [P]),
    "sgm_last_error": (ctypes.c_char_p, [P]),
    "sgm_host_alloc": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_size_t]),
    "sgm_host_free": (ctypes.c_int, [P]),
    "sgm_roi_from_disparity": (None, [P, I32, I32, ctypes.POINTER(SgmRoi)]),
    "sgm_invalid_float": (F32, []),
    "sgm_invalid_short": (ctypes.c_int16, []),
    "sgm_level_range": (ctypes.c_int16, []),
    "sgm_batch_create": (P, [ctypes.POINTER(SgmPairDesc), I32, ctypes.POINTER(SgmOption)]),
    "sgm_batch_run": (ctypes.c_int, [P, P]),
    "sgm_batch_set_timing": (ctypes.c_int, [P, I32]),
    "sgm_batch_stage_times": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_float),
                                             ctypes.POINTER(I32)]),
    "sgm_batch_device_bytes": (ctypes.c_size_t, [P]),
    "sgm_batch_set_schedule": (ctypes.c_int, [P, I32, I32]),
    "sgm_batch_get_schedule": (ctypes.c_int, [P, ctypes.POINTER(I32), ctypes.POINTER(I32)]),
    "sgm_batch_autotune": (ctypes.c_int, [P, P, I32, I32, ctypes.POINTER(I32), ctypes.POINTER(I32),
                                          ctypes.POINTER(ctypes.c_float)]),
    "sgm_batch_use_graph": (ctypes.c_int, [P, I32]),
    "sgm_set_cost_mode": (ctypes.c_int, [I32]),
    "sgm_get_cost_mode": (I32, []),
    "sgm_set_window_form": (ctypes.c_int, [I32]),
    "sgm_get_window_form": (I32, []),
    "sgm_set_wta_form": (ctypes.c_int, [I32]),
    "sgm_get_wta_form": (I32, []),
    "sgm_batch_destroy": (None, [P]),
    "sgm_dsgm_checkpoint_floats": (ctypes.c_size_t, [I32, I32, I32, I32]),
    "sgm_dsgm_workspace_floats": (ctypes.c_size_t, [I32, I32, I32, I32]),
    "sgm_dsgm_forward": (ctypes.c_int, [P, P, P, I32, I32, I32, I32, F32, P, P, P, P, P]),
    "sgm_dsgm_backward": (ctypes.c_int, [P, P, P, P, P, P, P, I32, I32, I32, I32, F32, P, P, P,
                                         P, P]),
    "sgm_pyramid_create": (P, [I32, I32, I32, I32, I32, I32, ctypes.POINTER(SgmOption)]),
    "sgm_pyramid_match": (ctypes.c_int, [P, P, P, P]),
    "sgm_pyramid_match_device": (ctypes.c_int, [P, P, P, P, P]),
    "sgm_pyramid_create2": (P, [I32, I32, F32, F32, I32, I32, ctypes.POINTER(SgmOption),
                                ctypes.POINTER(SgmCamera)]),
    "sgm_pyramid_match2": (ctypes.c_int, [P, P, P, P]),
    "sgm_pyramid_match2_device": (ctypes.c_int, [P, P, P, P, P]),
    "sgm_pyramid_coarsest_range": (I32, [P, ctypes.POINTER(I32)]),
    "sgm_pyramid_device_bytes": (ctypes.c_size_t, [P]),
    "sgm_pyramid_destroy": (None, [P]),
    "sgm_stage_zoom": (ctypes.c_int, [P, I32, I32, P, P]),
    "sgm_stage_init_next": (ctypes.c_int, [P, I32, I32, I32, P, P]),
    "sgm_stage_cost_init": (ctypes.c_int, [P, P, P, I32, I32, I32, I32, P, P]),
    "sgm_stage_aggregate_init": (ctypes.c_int, [P, P, I32, I32, I32, I32, I32, PP4, P]),
    "sgm_stage_wta_init": (ctypes.c_int, [PP4, P, I32, I32, I32, F32, P, P]),
    "sgm_stage_census": (ctypes.c_int, [P, I32, I32, I32, I32, P, P]),
    "sgm_stage_cost": (ctypes.c_int, [P, P, I32, I32, I32, I32, I32, I32, P, P]),
    "sgm_stage_aggregate": (ctypes.c_int, [P, I32, I32, I32, I32, I32, I32, PP4, P]),
    "sgm_stage_aggregate_wta_otf": (ctypes.c_int, [P, P, I32, I32, I32, I32, I32, I32, I32, I32,
                                                   F32, PP4, P, P]),
    "sgm_stage_wta": (ctypes.c_int, [PP4, I32, I32, I32, I32, F32, P, P]),
    "sgm_stage_post": (ctypes.c_int, [P, P, I32, I32, I32, F32, I32, P, P, P]),
}


def lib() -> ctypes.CDLL:
    """Load libfsgm.so.  Raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m fasterstereocuda_amd.build` "
                "(the SGM engine has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error(handle=None) -> str:
    msg = lib().sgm_last_error(handle)
    return msg.decode() if msg else ""


def check(rc: int, handle=None) -> None:
    if rc != SGM_OK:
        raise SgmError(rc, last_error(handle))
