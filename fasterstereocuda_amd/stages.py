"""Stage-level entry points on device tensors (for parity checks).

Each function runs one stage of the plain pipeline through the C ABI
(sgm_stage_*), on torch tensors resident on the GPU, and returns new
tensors.  Layouts: census int64 [H, W] (bit pattern of the uint64 census),
cost / path volumes uint8 [H, W, D], disparity float32 [H, W].
"""
from __future__ import annotations

import torch

from ._lib import PP4, check, lib

DIRS = ("top->bottom", "left->right", "bottom->top", "right->left")


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _dev(t: torch.Tensor, dtype) -> torch.Tensor:
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"expected a contiguous {dtype} CUDA tensor")
    return t


def census(img: torch.Tensor, cs_type: int, is_right: bool) -> torch.Tensor:
    _dev(img, torch.uint8)
    H, W = img.shape
    out = torch.empty((H, W), dtype=torch.int64, device=img.device)
    check(lib().sgm_stage_census(img.data_ptr(), W, H, cs_type, int(is_right), out.data_ptr(),
                                 _stream()))
    return out


def cost(census_base: torch.Tensor, census_match: torch.Tensor, min_d: int, D: int,
         left_to_right: bool, cs_type: int) -> torch.Tensor:
    _dev(census_base, torch.int64)
    _dev(census_match, torch.int64)
    H, W = census_base.shape
    out = torch.zeros((H, W, D), dtype=torch.uint8, device=census_base.device)
    check(lib().sgm_stage_cost(census_base.data_ptr(), census_match.data_ptr(), W, H, min_d, D,
                               int(left_to_right), cs_type, out.data_ptr(), _stream()))
    return out


def aggregate(cost_vol: torch.Tensor, min_d: int, p1: int, p2: int):
    _dev(cost_vol, torch.uint8)
    H, W, D = cost_vol.shape
    paths = [torch.zeros_like(cost_vol) for _ in range(4)]
    ptrs = PP4(*[p.data_ptr() for p in paths])
    check(lib().sgm_stage_aggregate(cost_vol.data_ptr(), W, H, min_d, D, p1, p2, ptrs, _stream()))
    return paths


def aggregate_wta_otf(census_base: torch.Tensor, census_match: torch.Tensor, min_d: int, D: int,
                      is_right: bool, cs_type: int, p1: int, p2: int, uniqueness: float):
    """The on-the-fly cost aggregation of one view: returns the top->bottom,
    left->right and right->left volumes (index 2, bottom->top, is None: the
    fused pass does not store it) and the raw WTA map."""
    _dev(census_base, torch.int64)
    _dev(census_match, torch.int64)
    H, W = census_base.shape
    if census_match.shape != (H, W):
        raise ValueError("census maps differ in shape")
    paths = [torch.zeros((H, W, D), dtype=torch.uint8, device=census_base.device)
             if r != 2 else None for r in range(4)]
    out = torch.empty((H, W), dtype=torch.float32, device=census_base.device)
    ptrs = PP4(*[0 if p is None else p.data_ptr() for p in paths])
    check(lib().sgm_stage_aggregate_wta_otf(census_base.data_ptr(), census_match.data_ptr(), W, H,
                                            min_d, D, int(is_right), cs_type, p1, p2, uniqueness,
                                            ptrs, out.data_ptr(), _stream()))
    return paths, out


def wta(paths, min_d: int, uniqueness: float) -> torch.Tensor:
    for p in paths:
        _dev(p, torch.uint8)
    H, W, D = paths[0].shape
    out = torch.empty((H, W), dtype=torch.float32, device=paths[0].device)
    ptrs = PP4(*[p.data_ptr() for p in paths])
    check(lib().sgm_stage_wta(ptrs, W, H, min_d, D, uniqueness, out.data_ptr(), _stream()))
    return out


def post(raw_left: torch.Tensor, raw_right: torch.Tensor | None, do_lr: bool, lr_threshold: float,
         do_median: bool, want_right: bool = False):
    _dev(raw_left, torch.float32)
    H, W = raw_left.shape
    out_l = torch.empty_like(raw_left)
    out_r = torch.empty_like(raw_left) if want_right else None
    check(lib().sgm_stage_post(raw_left.data_ptr(),
                               None if raw_right is None else _dev(raw_right, torch.float32).data_ptr(),
                               W, H, int(do_lr), lr_threshold, int(do_median), out_l.data_ptr(),
                               None if out_r is None else out_r.data_ptr(), _stream()))
    return out_l, out_r


def remove_peaks(disp: torch.Tensor, strips: int = 1, ratio: float = 0.0005) -> torch.Tensor:
    """Peak removal of a device disparity map; returns a new map."""
    _dev(disp, torch.float32)
    H, W = disp.shape
    out = disp.clone()
    scratch = torch.empty(2 * H * W, dtype=torch.int32, device=disp.device)
    check(lib().sgm_stage_remove_peaks(out.data_ptr(), W, H, strips, ratio, scratch.data_ptr(),
                                       _stream()))
    return out


# ---- coarse-to-fine pyramid stages -------------------------------------
def zoom_half(img: torch.Tensor) -> torch.Tensor:
    _dev(img, torch.uint8)
    h, w = img.shape
    out = torch.empty((h // 2, w // 2), dtype=torch.uint8, device=img.device)
    check(lib().sgm_stage_zoom(img.data_ptr(), w, h, out.data_ptr(), _stream()))
    return out


def init_next(coarse: torch.Tensor, w: int, h: int, level_range: int) -> torch.Tensor:
    _dev(coarse, torch.float32)
    out = torch.empty((h, w), dtype=torch.int16, device=coarse.device)
    check(lib().sgm_stage_init_next(coarse.data_ptr(), w, h, level_range, out.data_ptr(), _stream()))
    return out


def cost_init(census_l, census_r, init, level_range: int, cs_type: int = 1) -> torch.Tensor:
    _dev(census_l, torch.int64)
    _dev(census_r, torch.int64)
    _dev(init, torch.int16)
    h, w = census_l.shape
    out = torch.empty((h, w, level_range), dtype=torch.uint8, device=census_l.device)
    check(lib().sgm_stage_cost_init(census_l.data_ptr(), census_r.data_ptr(), init.data_ptr(), w, h,
                                    level_range, cs_type, out.data_ptr(), _stream()))
    return out


WINDOW_AUTO, WINDOW_U32, WINDOW_PAIRS, WINDOW_QUADS = 0, 1, 2, 3


def set_window_form(form: int) -> None:
    """How the window aggregation holds a scanline's disparities (one per
    lane in 32-bit arithmetic, two or four per lane as packed u16, or AUTO);
    process-wide, read at every launch.  Results are byte-identical."""
    check(lib().sgm_set_window_form(int(form)))


def get_window_form() -> int:
    return int(lib().sgm_get_window_form())


WTA_AUTO, WTA_FUSED, WTA_SEPARATE = 0, 1, 2


def set_wta_form(form: int) -> None:
    """How the stored-cost aggregation reaches the WTA: bottom->top fused
    with it after the other three directions, or the four directions in one
    launch and then the WTA, or AUTO (separate for single pairs); process-wide,
    read when the launches are enqueued.  Results are byte-identical."""
    check(lib().sgm_set_wta_form(int(form)))


def get_wta_form() -> int:
    return int(lib().sgm_get_wta_form())


def aggregate_init(cost_vol, init, p1: int, p2: int):
    _dev(cost_vol, torch.uint8)
    _dev(init, torch.int16)
    h, w, lr = cost_vol.shape
    paths = [torch.zeros_like(cost_vol) for _ in range(4)]
    ptrs = PP4(*[p.data_ptr() for p in paths])
    check(lib().sgm_stage_aggregate_init(cost_vol.data_ptr(), init.data_ptr(), w, h, lr, p1, p2, ptrs,
                                         _stream()))
    return paths


def wta_init(paths, init, uniqueness: float) -> torch.Tensor:
    h, w, lr = paths[0].shape
    out = torch.empty((h, w), dtype=torch.float32, device=paths[0].device)
    ptrs = PP4(*[p.data_ptr() for p in paths])
    check(lib().sgm_stage_wta_init(ptrs, _dev(init, torch.int16).data_ptr(), w, h, lr, uniqueness,
                                   out.data_ptr(), _stream()))
    return out
