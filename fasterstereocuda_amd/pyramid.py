"""Coarse-to-fine pyramid matching (the FasterStereoCuda surface).

``SgmPyramid`` exposes the engine's pyramid directly (any census / P1 / P2
/ level range); ``FasterStereoCuda`` mirrors the reference class
(FasterStereoCuda/FasterStereoCuda.h:24-127): Init with StereoOption1 and
Match a pair into a float disparity map, with the reference's per-level
option schedule (FasterStereoCuda.cpp:169-199: 5x5 census, P1 = 8, P2 = 32,
uniqueness 0.95 at level 0 and 1.0 above, LR check at the coarsest level,
3x3 median, dilation of the coarsest level, window of 32 disparities).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._checks import device_tensor, host_input
from ._lib import last_error, lib
from .stereo import CamParam, CuSGMOption


class SgmPyramid:
    def __init__(self, width: int, height: int, min_disp: int, max_disp: int, num_layers: int,
                 level_range: int, option: CuSGMOption):
        opt = option.to_c()
        self._lib = lib()    # kept for __del__, which may run after module teardown
        self._p = self._lib.sgm_pyramid_create(width, height, min_disp, max_disp, num_layers,
                                               level_range, ctypes.byref(opt))
        if not self._p:
            raise RuntimeError(f"sgm_pyramid_create failed: {last_error()}")
        self.width, self.height = width, height

    @classmethod
    def from_depth(cls, width: int, height: int, min_depth: float, max_depth: float,
                   num_layers: int, level_range: int, option: CuSGMOption,
                   cam: CamParam) -> "SgmPyramid":
        """Pyramid whose coarsest interval comes from a depth interval
        (HierSgmCudaImpl::Init2); match2 then returns depth."""
        self = cls.__new__(cls)
        opt, c = option.to_c(), cam.to_c()
        self._lib = lib()
        self._p = self._lib.sgm_pyramid_create2(width, height, min_depth, max_depth, num_layers,
                                                level_range, ctypes.byref(opt), ctypes.byref(c))
        if not self._p:
            raise RuntimeError(f"sgm_pyramid_create2 failed: {last_error()}")
        self.width, self.height = width, height
        return self

    def __del__(self):
        p, L = getattr(self, "_p", None), getattr(self, "_lib", None)
        if p and L is not None:
            L.sgm_pyramid_destroy(p)
            self._p = None

    @property
    def coarsest(self):
        m = ctypes.c_int32(0)
        r = lib().sgm_pyramid_coarsest_range(self._p, ctypes.byref(m))
        return m.value, r

    @property
    def device_bytes(self) -> int:
        return lib().sgm_pyramid_device_bytes(self._p)

    def _inputs(self, left, right):
        shape = (self.height, self.width)
        return host_input(left, shape, np.uint8, "left"), host_input(right, shape, np.uint8, "right")

    def _device(self, d_left, d_right, d_out, name):
        import torch
        shape = (self.height, self.width)
        device_tensor(d_left, shape, torch.uint8, "d_left")
        device_tensor(d_right, shape, torch.uint8, "d_right")
        device_tensor(d_out, shape, torch.float32, name)

    def match(self, left: np.ndarray, right: np.ndarray) -> np.ndarray:
        left, right = self._inputs(left, right)
        out = np.empty((self.height, self.width), np.float32)
        _lib.check(lib().sgm_pyramid_match(self._p, left.ctypes.data, right.ctypes.data,
                                           out.ctypes.data))
        return out

    def match_device(self, d_left, d_right, d_disp, stream=None) -> None:
        self._device(d_left, d_right, d_disp, "d_disp")
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(lib().sgm_pyramid_match_device(self._p, d_left.data_ptr(), d_right.data_ptr(),
                                                  d_disp.data_ptr(), stream))

    def match2(self, left: np.ndarray, right: np.ndarray) -> np.ndarray:
        """Depth map of a pair (pyramids made by from_depth)."""
        left, right = self._inputs(left, right)
        out = np.empty((self.height, self.width), np.float32)
        _lib.check(lib().sgm_pyramid_match2(self._p, left.ctypes.data, right.ctypes.data,
                                            out.ctypes.data))
        return out

    def match2_device(self, d_left, d_right, d_depth, stream=None) -> None:
        self._device(d_left, d_right, d_depth, "d_depth")
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(lib().sgm_pyramid_match2_device(self._p, d_left.data_ptr(), d_right.data_ptr(),
                                                   d_depth.data_ptr(), stream))


@dataclass
class StereoOption1:
    """ste_opt1_t (FasterStereoCuda.h:39-51), same fields and defaults."""
    width: int = 0
    height: int = 0
    min_disp: int = 0
    max_disp: int = 0
    num_layers: int = 1
    do_lr_check: bool = True
    do_rm_peaks: bool = True
    do_smooth: bool = False


@dataclass
class EpipolarStereoParam:
    """epi_ste_t (FasterStereoCuda.h:26-35), same fields and defaults."""
    x0_left: float = 0.0
    y0_left: float = 0.0
    x0_right: float = 0.0
    y0_right: float = 0.0
    baseline: float = 0.0
    focus: float = 0.0

    def camera(self) -> CamParam:
        return CamParam(self.x0_left, self.y0_left, self.x0_right, self.y0_right, self.focus,
                        self.baseline)


@dataclass
class StereoOption2:
    """ste_opt2_t (FasterStereoCuda.h:53-66), same fields and defaults."""
    width: int = 0
    height: int = 0
    min_depth: float = 0.0
    max_depth: float = 0.0
    num_layers: int = 1
    do_lr_check: bool = True
    do_rm_peaks: bool = True
    do_smooth: bool = False
    epi: EpipolarStereoParam = field(default_factory=EpipolarStereoParam)


def reference_schedule(opt) -> CuSGMOption:
    """Options FasterStereoCuda.cpp:169-199 gives every level (the level
    dependent parts -- uniqueness, LR level, dilation, where peaks are
    removed -- are applied by the pyramid itself).  Peaks go at the coarsest
    level only, with ratio 0.0005 / 0.001 / 0.005 for 1 / 2 / more levels."""
    layers = max(1, opt.num_layers)
    ratio = 0.0005 if layers <= 1 else (0.001 if layers <= 2 else 0.005)
    return CuSGMOption(p1=8, p2_init=32, p2_max=64, cs_type=CuSGMOption.CS_5x5, num_paths=4,
                       do_median_filter=True, do_lr_check=opt.do_lr_check, lr_threshold=1.0,
                       unique_threshold=0.95, using_constant_p2=True,
                       do_remove_peaks=opt.do_rm_peaks, peaks_ratio_threshold=ratio,
                       post_filter_type=1 if opt.do_smooth else 0)


class FasterStereoCuda:
    LEVEL_RANGE = 32

    def __init__(self):
        self._pyr = None
        self.last_error = ""

    def Init(self, option: StereoOption1) -> bool:
        if option.width <= 0 or option.height <= 0:
            self.last_error = "image size error"
            return False
        try:
            self._pyr = SgmPyramid(option.width, option.height, option.min_disp, option.max_disp,
                                   max(1, option.num_layers), self.LEVEL_RANGE,
                                   reference_schedule(option))
        except RuntimeError as e:
            self.last_error = str(e)
            self._pyr = None
            return False
        return True

    def Init2(self, option: StereoOption2) -> bool:
        if option.width <= 0 or option.height <= 0:
            self.last_error = "image size error"
            return False
        try:
            self._pyr = SgmPyramid.from_depth(option.width, option.height, option.min_depth,
                                              option.max_depth, max(1, option.num_layers),
                                              self.LEVEL_RANGE, reference_schedule(option),
                                              option.epi.camera())
        except RuntimeError as e:
            self.last_error = str(e)
            self._pyr = None
            return False
        return True

    def Match(self, bytes_left: np.ndarray, bytes_right: np.ndarray, disparity: np.ndarray) -> bool:
        if self._pyr is None:
            self.last_error = "not initialised"
            return False
        try:
            disparity[...] = self._pyr.match(bytes_left, bytes_right)
        except (_lib.SgmError, ValueError) as e:
            self.last_error = str(e)
            return False
        return True

    def Match2(self, bytes_left: np.ndarray, bytes_right: np.ndarray, depth: np.ndarray) -> bool:
        if self._pyr is None:
            self.last_error = "not initialised"
            return False
        try:
            depth[...] = self._pyr.match2(bytes_left, bytes_right)
        except (_lib.SgmError, ValueError) as e:
            self.last_error = str(e)
            return False
        return True

    def Release(self) -> None:
        self._pyr = None
