"""ctypes loader for the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module.  It wraps ``oracle/libsgm_oracle.so`` (built
from ``oracle/sgm_oracle.cpp`` by ``oracle/Makefile``), a C++/OpenMP
restatement of the reference SGM arithmetic; see the header of that file for
the reference lines each function follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsgm_oracle.so")
_lib = None

INVALID = np.float32(-999999.0)
DIR_UD, DIR_LR, DIR_DU, DIR_RL = 0, 1, 2, 3


class Options(ctypes.Structure):
    _fields_ = [
        ("cs_type", ctypes.c_int32),
        ("p1", ctypes.c_int32),
        ("p2", ctypes.c_int32),
        ("uniqueness", ctypes.c_float),
        ("do_lr_check", ctypes.c_int32),
        ("lr_threshold", ctypes.c_float),
        ("do_median", ctypes.c_int32),
        ("do_remove_peaks", ctypes.c_int32),
        ("peaks_ratio", ctypes.c_float),
        ("peak_strips", ctypes.c_int32),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        F = ctypes.c_float
        L.orc_num_threads.restype = I
        L.orc_census.argtypes = [P, I, I, I, I, P]
        L.orc_cost.argtypes = [P, P, I, I, I, I, I, I, P, P]
        L.orc_aggregate.argtypes = [P, I, I, I, I, I, I, I, P]
        L.orc_wta.argtypes = [P, P, P, P, I, I, I, I, I, F, P]
        L.orc_lrcheck.argtypes = [P, P, P, I, I, F]
        L.orc_median3x3.argtypes = [P, P, I, I]
        L.orc_match.argtypes = [P, P, I, I, I, I, ctypes.POINTER(Options), P, P]
        L.orc_match.restype = I
        L.orc_match_roi.argtypes = [P, P, I, I, I, I, ctypes.POINTER(Options), P, P, P, P]
        L.orc_match_roi.restype = I
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def num_threads() -> int:
    return lib().orc_num_threads()


def census(img: np.ndarray, cs_type: int, is_right: bool) -> np.ndarray:
    img = np.ascontiguousarray(img, dtype=np.uint8)
    H, W = img.shape
    out = np.empty((H, W), np.uint64)
    lib().orc_census(_p(img), W, H, cs_type, int(is_right), _p(out))
    return out


def cost(cs_base, cs_match, min_d: int, D: int, left_to_right: bool, cs_type: int,
         l2r_cost: np.ndarray | None = None, out: np.ndarray | None = None) -> np.ndarray:
    H, W = cs_base.shape
    if out is None:
        out = np.zeros((H, W, D), np.uint8)
    lib().orc_cost(_p(cs_base), _p(cs_match), W, H, min_d, D, int(left_to_right), cs_type,
                   _p(l2r_cost), _p(out))
    return out


def aggregate(cost_vol: np.ndarray, roi_x: int, direction: int, p1: int, p2: int) -> np.ndarray:
    H, W, D = cost_vol.shape
    out = np.zeros((H, W, D), np.uint8)
    lib().orc_aggregate(_p(cost_vol), W, H, D, roi_x, direction, p1, p2, _p(out))
    return out


def wta(paths, min_d: int, roi_x: int, uniqueness: float) -> np.ndarray:
    H, W, D = paths[0].shape
    out = np.empty((H, W), np.float32)
    lib().orc_wta(*[_p(p) for p in paths], W, H, D, min_d, roi_x, uniqueness, _p(out))
    return out


def lrcheck(disp_l: np.ndarray, disp_r: np.ndarray, threshold: float):
    """Returns (checked left map, right map with invalidations applied)."""
    H, W = disp_l.shape
    left = disp_l.copy()
    right_out = disp_r.copy()
    lib().orc_lrcheck(_p(left), _p(np.ascontiguousarray(disp_r)), _p(right_out), W, H, threshold)
    return left, right_out


def median3x3(disp: np.ndarray) -> np.ndarray:
    H, W = disp.shape
    out = np.empty_like(disp)
    lib().orc_median3x3(_p(np.ascontiguousarray(disp)), _p(out), W, H)
    return out


def match(left: np.ndarray, right: np.ndarray, min_d: int, D: int, cs_type: int = 1,
          p1: int = 10, p2: int = 150, uniqueness: float = 0.95, do_lr_check: bool = True,
          lr_threshold: float = 1.0, do_median: bool = True, peaks_ratio: float | None = None,
          peak_strips: int = 1):
    """Whole plain pipeline.  Returns (disp_left, disp_right).  With
    peaks_ratio set, the left map loses its peaks (remove_peaks)."""
    left = np.ascontiguousarray(left, dtype=np.uint8)
    right = np.ascontiguousarray(right, dtype=np.uint8)
    H, W = left.shape
    opt = Options(cs_type, p1, p2, uniqueness, int(do_lr_check), lr_threshold, int(do_median),
                  int(peaks_ratio is not None), peaks_ratio or 0.0, peak_strips)
    dl = np.empty((H, W), np.float32)
    dr = np.empty((H, W), np.float32)
    rc = lib().orc_match(_p(left), _p(right), W, H, min_d, D, ctypes.byref(opt), _p(dl), _p(dr))
    if rc != 0:
        raise RuntimeError(f"orc_match failed with {rc}")
    return dl, dr


def match_roi(left: np.ndarray, right: np.ndarray, min_d: int, D: int, roi_left=None,
              roi_right=None, cs_type: int = 1, p1: int = 10, p2: int = 150,
              uniqueness: float = 0.95, do_lr_check: bool = True, lr_threshold: float = 1.0,
              do_median: bool = True, peaks_ratio: float | None = None, peak_strips: int = 1):
    """The plain pipeline with a region of interest (x, y, w, h) per view
    (None = the whole matched region).  Returns (disp_left, disp_right)."""
    left = np.ascontiguousarray(left, dtype=np.uint8)
    right = np.ascontiguousarray(right, dtype=np.uint8)
    H, W = left.shape
    opt = Options(cs_type, p1, p2, uniqueness, int(do_lr_check), lr_threshold, int(do_median),
                  int(peaks_ratio is not None), peaks_ratio or 0.0, peak_strips)
    rl = None if roi_left is None else np.ascontiguousarray(roi_left, np.int32)
    rr = None if roi_right is None else np.ascontiguousarray(roi_right, np.int32)
    dl = np.empty((H, W), np.float32)
    dr = np.empty((H, W), np.float32)
    rc = lib().orc_match_roi(_p(left), _p(right), W, H, min_d, D, ctypes.byref(opt), _p(rl), _p(rr),
                             _p(dl), _p(dr))
    if rc != 0:
        raise RuntimeError(f"orc_match_roi failed with {rc}")
    return dl, dr


def stages(left: np.ndarray, right: np.ndarray, min_d: int, D: int, cs_type: int = 1,
           p1: int = 10, p2: int = 150, uniqueness: float = 0.95, lr_threshold: float = 1.0):
    """Every intermediate of the plain pipeline with the LR check on, for
    stage-by-stage parity checks.  Returns a dict of numpy arrays."""
    roi_x = max(0, min_d)
    out = {}
    out["census_l"] = cl = census(left, cs_type, False)
    out["census_r"] = cr = census(right, cs_type, True)
    out["cost_l"] = c_l = cost(cl, cr, min_d, D, True, cs_type)
    out["paths_l"] = pl = [aggregate(c_l, roi_x, r, p1, p2) for r in range(4)]
    out["wta_l"] = wl = wta(pl, min_d, roi_x, uniqueness)
    out["cost_r"] = c_r = cost(cr, cl, min_d, D, False, cs_type,
                               l2r_cost=c_l if cs_type == 0 else None,
                               out=c_l.copy())
    out["paths_r"] = pr = [aggregate(c_r, roi_x, r, p1, p2) for r in range(4)]
    out["wta_r"] = wr = wta(pr, min_d, roi_x, uniqueness)
    out["lr_l"], out["lr_r"] = lrcheck(wl, wr, lr_threshold)
    out["median"] = median3x3(out["lr_l"])
    return out


# ---------------------------------------------------------------- diff-SGM
_dlib = None


def dlib():
    """FP64 reference of the differentiable SGM (oracle/dsgm_oracle.cpp)."""
    global _dlib
    if _dlib is None:
        path = os.path.join(_HERE, "libdsgm_oracle.so")
        if not os.path.exists(path):
            build(force=True)
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I = ctypes.c_int
        Dd = ctypes.c_double
        L.dsgm_forward.argtypes = [P, P, P, P, I, I, I, Dd, P, P]
        L.dsgm_forward.restype = Dd
        L.dsgm_backward.argtypes = [P, P, P, P, I, I, I, Dd, P, P, P]
        L.dsgm_backward.restype = Dd
        _dlib = L
    return _dlib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dsgm_forward(cost, p1, p2, gt=None, temperature=0.1):
    """Returns (S [H,W,D], disp [H,W], loss) in float64."""
    cost, p1, p2 = _f64(cost), _f64(p1), _f64(p2)
    H, W, D = cost.shape
    S = np.empty((H, W, D), np.float64)
    disp = np.empty((H, W), np.float64)
    g = None if gt is None else _f64(gt)
    loss = dlib().dsgm_forward(_p(cost), _p(p1), _p(p2), _p(g), H, W, D, temperature,
                               _p(S), _p(disp))
    return S, disp, loss


def dsgm_backward(cost, p1, p2, gt, temperature=0.1):
    """Gradient of the mean L1 loss.  Returns (gC, gP1, gP2, loss) in float64."""
    cost, p1, p2, gt = _f64(cost), _f64(p1), _f64(p2), _f64(gt)
    H, W, D = cost.shape
    gC = np.empty((H, W, D), np.float64)
    g1 = np.empty((H, W), np.float64)
    g2 = np.empty((H, W), np.float64)
    loss = dlib().dsgm_backward(_p(cost), _p(p1), _p(p2), _p(gt), H, W, D, temperature,
                                _p(gC), _p(g1), _p(g2))
    return gC, g1, g2, loss


# ---------------------------------------------------------------- pyramid
INVALID_SHORT = -32768


def _plib():
    L = lib()
    if not getattr(L, "_pyr_ready", False):
        P = ctypes.c_void_p
        I = ctypes.c_int
        F = ctypes.c_float
        L.orc_zoom_half.argtypes = [P, I, I, P]
        L.orc_init_next.argtypes = [P, I, I, I, P]
        L.orc_cost_init.argtypes = [P, P, P, I, I, I, I, P]
        L.orc_aggregate_init.argtypes = [P, P, I, I, I, I, I, I, P]
        L.orc_wta_init.argtypes = [P, P, P, P, P, I, I, I, F, P]
        L.orc_match_init.argtypes = [P, P, P, I, I, I, ctypes.POINTER(Options), P]
        L.orc_match_init.restype = I
        L.orc_pyramid.argtypes = [P, P, I, I, I, I, I, I, ctypes.POINTER(Options), P]
        L.orc_pyramid.restype = I
        L.orc_dilation.argtypes = [P, P, I, I, I]
        L.orc_dilation_window.argtypes = [I, I]
        L.orc_dilation_window.restype = I
        L.orc_pyramid_at.argtypes = [P, P, I, I, I, I, I, I, ctypes.POINTER(Options), P]
        L.orc_pyramid_at.restype = I
        D6 = ctypes.POINTER(ctypes.c_double)
        PI = ctypes.POINTER(I)
        L.orc_disp_interval.argtypes = [I, I, I, I, PI, PI]
        L.orc_depth_interval.argtypes = [ctypes.c_double, ctypes.c_double, I, I, D6, PI, PI]
        L.orc_depth_interval.restype = I
        L.orc_round_interval.argtypes = [ctypes.c_double, I, PI, PI]
        L.orc_depth.argtypes = [P, ctypes.c_size_t, D6, P]
        L._pyr_ready = True
    return L


def zoom_half(img):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    out = np.empty((h // 2, w // 2), np.uint8)
    _plib().orc_zoom_half(_p(img), w, h, _p(out))
    return out


def init_next(coarse, w, h, level_range):
    coarse = np.ascontiguousarray(coarse, np.float32)
    out = np.empty((h, w), np.int16)
    _plib().orc_init_next(_p(coarse), w, h, level_range, _p(out))
    return out


def cost_init(cl, cr, init, level_range, cs_type=1):
    H, W = cl.shape
    out = np.empty((H, W, level_range), np.uint8)
    _plib().orc_cost_init(_p(cl), _p(cr), _p(np.ascontiguousarray(init, np.int16)), W, H,
                          level_range, cs_type, _p(out))
    return out


def aggregate_init(cost_vol, init, direction, p1, p2, fill=0):
    H, W, D = cost_vol.shape
    out = np.full((H, W, D), fill, np.uint8)
    _plib().orc_aggregate_init(_p(cost_vol), _p(np.ascontiguousarray(init, np.int16)), W, H, D,
                               direction, p1, p2, _p(out))
    return out


def wta_init(paths, init, uniqueness):
    H, W, D = paths[0].shape
    out = np.empty((H, W), np.float32)
    _plib().orc_wta_init(*[_p(p) for p in paths], _p(np.ascontiguousarray(init, np.int16)), W, H,
                         D, uniqueness, _p(out))
    return out


def dilation(disp, wnd):
    disp = np.ascontiguousarray(disp, np.float32)
    H, W = disp.shape
    out = np.empty_like(disp)
    _plib().orc_dilation(_p(disp), _p(out), W, H, wnd)
    return out


def dilation_window(w, h):
    return _plib().orc_dilation_window(w, h)


def match_init(left, right, init, level_range, cs_type=1, p1=10, p2=150, uniqueness=0.95,
               do_median=True):
    left = np.ascontiguousarray(left, np.uint8)
    right = np.ascontiguousarray(right, np.uint8)
    H, W = left.shape
    opt = Options(cs_type, p1, p2, uniqueness, 0, 1.0, int(do_median))
    out = np.empty((H, W), np.float32)
    rc = _plib().orc_match_init(_p(left), _p(right), _p(np.ascontiguousarray(init, np.int16)), W,
                                H, level_range, ctypes.byref(opt), _p(out))
    if rc:
        raise RuntimeError(f"orc_match_init failed with {rc}")
    return out


def pyramid(left, right, min_disp, max_disp, num_layers, level_range, cs_type=1, p1=10, p2=150,
            uniqueness=0.95, do_lr_check=True, lr_threshold=1.0, do_median=True,
            peaks_ratio=None, peak_strips=1):
    left = np.ascontiguousarray(left, np.uint8)
    right = np.ascontiguousarray(right, np.uint8)
    H, W = left.shape
    opt = Options(cs_type, p1, p2, uniqueness, int(do_lr_check), lr_threshold, int(do_median),
                  int(peaks_ratio is not None), peaks_ratio or 0.0, peak_strips)
    out = np.empty((H, W), np.float32)
    rc = _plib().orc_pyramid(_p(left), _p(right), W, H, min_disp, max_disp, num_layers,
                             level_range, ctypes.byref(opt), _p(out))
    if rc:
        raise RuntimeError(f"orc_pyramid failed with {rc}")
    return out


# ------------------------------------------------------------------ depth
def _cam6(cam):
    """cam: (x0_left, y0_left, x0_right, y0_right, focus, baseline)"""
    return (ctypes.c_double * 6)(*[float(v) for v in cam])


def depth(disp, cam):
    """Depth map of a disparity map (cusgm_util.cu:7-33)."""
    disp = np.ascontiguousarray(disp, np.float32)
    out = np.empty_like(disp)
    _plib().orc_depth(_p(disp), disp.size, _cam6(cam), _p(out))
    return out


def disp_interval(min_disp, max_disp, W, num_layers):
    """Coarsest (min, range) of an option-1 pyramid (FasterStereoCuda.cpp:265-281)."""
    a, b = ctypes.c_int(), ctypes.c_int()
    _plib().orc_disp_interval(min_disp, max_disp, W, num_layers, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def depth_interval(min_depth, max_depth, W, num_layers, cam):
    """Coarsest (min, range) of a depth pyramid (FasterStereoCuda.cpp:373-391).
    min_depth / max_depth are the float32 option values."""
    a, b = ctypes.c_int(), ctypes.c_int()
    rc = _plib().orc_depth_interval(float(np.float32(min_depth)), float(np.float32(max_depth)), W,
                                    num_layers, _cam6(cam), ctypes.byref(a), ctypes.byref(b))
    if rc:
        raise ValueError("bad depth interval")
    return a.value, b.value


def round_interval(span, W, dmin):
    a, b = ctypes.c_int(dmin), ctypes.c_int()
    _plib().orc_round_interval(float(span), W, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def pyramid_at(left, right, dmin, drange, num_layers, level_range, cs_type=1, p1=10, p2=150,
               uniqueness=0.95, do_lr_check=True, lr_threshold=1.0, do_median=True,
               peaks_ratio=None, peak_strips=1):
    left = np.ascontiguousarray(left, np.uint8)
    right = np.ascontiguousarray(right, np.uint8)
    H, W = left.shape
    opt = Options(cs_type, p1, p2, uniqueness, int(do_lr_check), lr_threshold, int(do_median),
                  int(peaks_ratio is not None), peaks_ratio or 0.0, peak_strips)
    out = np.empty((H, W), np.float32)
    rc = _plib().orc_pyramid_at(_p(left), _p(right), W, H, dmin, drange, num_layers, level_range,
                                ctypes.byref(opt), _p(out))
    if rc:
        raise RuntimeError(f"orc_pyramid_at failed with {rc}")
    return out


def remove_peaks(disp, strips=1, ratio=0.0005):
    """Peak removal of a disparity map (StereoCudaImpl.cpp:32-97, 463-498);
    returns (new map, pixels removed)."""
    out = np.array(disp, np.float32, copy=True, order="C")
    H, W = out.shape
    L = lib()
    L.orc_remove_peaks.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float]
    L.orc_remove_peaks.restype = ctypes.c_int
    n = L.orc_remove_peaks(_p(out), W, H, strips, ratio)
    return out, n
