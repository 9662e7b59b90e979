"""Python mirror of the StereoCuda drop-in class.

Same method names, argument meaning and bool-return convention as the
reference's SgmStereoCuda/StereoCuda.h:14-100; every call goes through the
C ABI of libfsgm.so.  Host buffers are numpy arrays; ``match_device`` takes
device tensors (anything with ``data_ptr()``) and the current stream.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._checks import device_tensor, host_array
from ._lib import SgmCamera, SgmOption, SgmRoi, check, lib


@dataclass
class CuSGMOption:
    """Same fields and defaults as CuSGMOption (sgm_types.h:21-65)."""
    p1: int = 8
    p2_init: int = 32
    p2_max: int = 0
    cs_type: int = 0            # 0 = CS_5x5, 1 = CS_9x7
    num_paths: int = 0
    do_median_filter: bool = False
    do_lr_check: bool = False
    lr_threshold: float = 0.0
    unique_threshold: float = 0.0
    using_constant_p2: bool = False
    do_remove_peaks: bool = False
    peaks_ratio_threshold: float = 0.0
    post_filter_type: int = 0
    morphology_type: int = 0

    CS_5x5 = 0
    CS_9x7 = 1

    def to_c(self) -> SgmOption:
        return SgmOption(int(self.p1), int(self.p2_init), int(self.p2_max), int(self.cs_type),
                         int(self.num_paths), int(bool(self.do_median_filter)),
                         int(bool(self.do_lr_check)), float(self.lr_threshold),
                         float(self.unique_threshold), int(bool(self.using_constant_p2)),
                         int(bool(self.do_remove_peaks)), float(self.peaks_ratio_threshold),
                         int(self.post_filter_type), int(self.morphology_type))

    @classmethod
    def benchmark(cls, **kw) -> "CuSGMOption":
        """The options of the benchmark workloads: 9x7 census, P1=10, P2=150,
        uniqueness 0.95, LR check (tolerance 1) and 3x3 median."""
        o = cls(p1=10, p2_init=150, cs_type=1, do_median_filter=True, do_lr_check=True,
                lr_threshold=1.0, unique_threshold=0.95, using_constant_p2=True)
        for k, v in kw.items():
            setattr(o, k, v)
        return o


COST_AUTO, COST_STORED, COST_ON_THE_FLY = 0, 1, 2


@dataclass
class CamParam:
    """CamParam_T (sgm_types.h:69-84): principal points and focal length in
    pixels, baseline in the caller's depth unit."""
    x0_left: float = 0.0
    y0_left: float = 0.0
    x0_right: float = 0.0
    y0_right: float = 0.0
    focus: float = 0.0
    baseline: float = 0.0

    def to_c(self) -> SgmCamera:
        return SgmCamera(self.x0_left, self.y0_left, self.x0_right, self.y0_right, self.focus,
                         self.baseline)


def depth_device(d_disp, d_depth, cam: CamParam, stream=None) -> None:
    """Depth of a device disparity map (float32 H x W tensors), enqueued on
    `stream` (default: torch's current stream)."""
    if stream is None:
        import torch
        stream = torch.cuda.current_stream().cuda_stream
    H, W = d_disp.shape
    c = cam.to_c()
    check(lib().sgm_depth_device(d_disp.data_ptr(), W, H, ctypes.byref(c), d_depth.data_ptr(),
                                 stream))


def set_peak_strips(strips: int) -> None:
    """Row strips of do_remove_peaks (process-wide, default 1); the
    reference uses its host's hardware_concurrency() * 3 // 4."""
    check(lib().sgm_set_peak_strips(strips))


def get_peak_strips() -> int:
    return lib().sgm_get_peak_strips()


def set_cost_mode(mode: int) -> None:
    """Where the aggregation takes its matching cost, for handles, batches
    and pyramids created afterwards: COST_ON_THE_FLY recomputes it from the
    census words wherever the views allow, COST_STORED always builds the u8
    cost volume, COST_AUTO picks on the fly only for plans large enough to
    fill the GPU.  Results are byte-identical."""
    check(lib().sgm_set_cost_mode(int(mode)))


def get_cost_mode() -> int:
    return int(lib().sgm_get_cost_mode())


def _roi_ref(r):
    """A region of interest as the C ABI takes it: None, or a pointer to an
    sgm_roi built from an SgmRoi or an (x, y, w, h) sequence."""
    if r is None:
        return None
    if not isinstance(r, SgmRoi):
        x, y, w, h = (int(v) for v in r)
        r = SgmRoi(x, y, w, h)
    return ctypes.byref(r)


class StereoCuda:
    """Single-level SGM matcher (StereoCuda.h)."""

    def __init__(self):
        self._lib = lib()    # kept for __del__, which may run after module teardown
        self._h = self._lib.sgm_create()
        if not self._h:
            raise MemoryError("sgm_create failed")
        self.width = self.height = 0
        self.last_error = ""

    def __del__(self):
        h, L = getattr(self, "_h", None), getattr(self, "_lib", None)
        if h and L is not None:
            L.sgm_destroy(h)
            self._h = None

    def _ok(self, rc: int) -> bool:
        if rc != _lib.SGM_OK:
            self.last_error = _lib.last_error(self._h)
            return False
        return True

    def Init(self, width: int, height: int, min_disparity: int, disp_range: int,
             sgm_option: CuSGMOption, print_log: bool = False) -> bool:
        """disp_range 64 * 2^n (n = 0..3), or 32 (LEVEL_RANGE) for a handle
        that matches around an initial disparity map only."""
        opt = sgm_option.to_c()
        ok = self._ok(lib().sgm_init(self._h, width, height, min_disparity, disp_range,
                                     ctypes.byref(opt), int(print_log)))
        if ok:
            self.width, self.height = width, height
        return ok

    def Match(self, img_left: np.ndarray, img_right: np.ndarray, disp_left: np.ndarray,
              disp_right: np.ndarray | None = None, init_disp_left=None,
              ste_roi_left=None, ste_roi_right=None) -> bool:
        """init_disp_left (int16, the image shape) makes a disp_range 32
        handle search 32 disparities around each pixel's value, as the
        reference's finer pyramid levels do.  ste_roi_left / ste_roi_right
        (SgmRoi or an (x, y, w, h) tuple) restrict the left view's and the
        LR check's right view's matching to a rectangle; outside it the map
        is INVALID."""
        shape = (self.height, self.width)
        host_array(img_left, shape, np.uint8, "img_left")
        host_array(img_right, shape, np.uint8, "img_right")
        host_array(disp_left, shape, np.float32, "disp_left")
        if disp_right is not None:
            host_array(disp_right, shape, np.float32, "disp_right")
        if init_disp_left is not None:
            host_array(init_disp_left, shape, np.int16, "init_disp_left")
        ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        init = ptr(init_disp_left)
        rl, rr = _roi_ref(ste_roi_left), _roi_ref(ste_roi_right)
        return self._ok(lib().sgm_match(self._h, ptr(img_left), ptr(img_right), ptr(disp_left),
                                        ptr(disp_right), init, rl, rr))

    def Init2(self, width: int, height: int, min_disparity: int, disp_range: int,
              sgm_option: CuSGMOption, cam_param: "CamParam", print_log: bool = False) -> bool:
        """Init plus the camera Match2 converts with (StereoCuda.h:57)."""
        opt = sgm_option.to_c()
        cam = cam_param.to_c()
        ok = self._ok(lib().sgm_init2(self._h, width, height, min_disparity, disp_range,
                                      ctypes.byref(opt), ctypes.byref(cam), int(print_log)))
        if ok:
            self.width, self.height = width, height
        return ok

    def Match2(self, img_left: np.ndarray, img_right: np.ndarray, depth_left: np.ndarray,
               init_disp_left=None, ste_roi_left=None, ste_roi_right=None) -> bool:
        """Match, then the float depth of the left map into depth_left
        (StereoCuda.h:69); invalid pixels stay -999999.0f."""
        shape = (self.height, self.width)
        host_array(img_left, shape, np.uint8, "img_left")
        host_array(img_right, shape, np.uint8, "img_right")
        host_array(depth_left, shape, np.float32, "depth_left")
        if init_disp_left is not None:
            host_array(init_disp_left, shape, np.int16, "init_disp_left")
        init = None if init_disp_left is None else init_disp_left.ctypes.data
        rl, rr = _roi_ref(ste_roi_left), _roi_ref(ste_roi_right)
        return self._ok(lib().sgm_match2(self._h, img_left.ctypes.data, img_right.ctypes.data,
                                         depth_left.ctypes.data, init, rl, rr))

    def match_device(self, d_left, d_right, d_disp_left, d_disp_right=None, stream=None) -> bool:
        """Pipeline on device tensors (uint8 / float32, C-contiguous), enqueued
        on `stream` (a hipStream_t as int, default: torch's current stream)."""
        import torch
        shape = (self.height, self.width)
        device_tensor(d_left, shape, torch.uint8, "d_left")
        device_tensor(d_right, shape, torch.uint8, "d_right")
        device_tensor(d_disp_left, shape, torch.float32, "d_disp_left")
        if d_disp_right is not None:
            device_tensor(d_disp_right, shape, torch.float32, "d_disp_right")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        return self._ok(lib().sgm_match_device(
            self._h, d_left.data_ptr(), d_right.data_ptr(), d_disp_left.data_ptr(),
            None if d_disp_right is None else d_disp_right.data_ptr(), stream))

    def SetMinDisparity(self, min_disparity: int) -> None:
        self._ok(lib().sgm_set_min_disparity(self._h, min_disparity))

    def GetDispartyRange(self) -> int:
        return lib().sgm_get_disparity_range(self._h)

    def Release(self) -> None:
        lib().sgm_release(self._h)

    @staticmethod
    def GetRoiFromDispMap(disp: np.ndarray) -> SgmRoi:
        roi = SgmRoi()
        h, w = disp.shape
        d = np.ascontiguousarray(disp, dtype=np.float32)
        lib().sgm_roi_from_disparity(d.ctypes.data, w, h, ctypes.byref(roi))
        return roi

    @staticmethod
    def get_invad_float() -> float:
        return lib().sgm_invalid_float()

    @staticmethod
    def get_invad_short() -> int:
        return lib().sgm_invalid_short()

    @staticmethod
    def get_level_range() -> int:
        return lib().sgm_level_range()
