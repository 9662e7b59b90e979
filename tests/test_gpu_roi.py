"""Matching with regions of interest (StereoCuda::Match's ste_roi_left /
ste_roi_right) on the GPU against the oracle, bit for bit: rectangles inside
the image, past its edges, left of min_disparity and empty; both census
sizes, D = 64 ... 512, +-min_d, with and without the LR check, the median
and peak removal; Match2 with a region; and the argument checks."""
import numpy as np
import pytest

import oracle
from fasterstereocuda_amd import scenes
from fasterstereocuda_amd.stereo import CamParam, CuSGMOption, StereoCuda

pytestmark = pytest.mark.gpu


def _eq(a, b, what):
    a = np.asarray(a, np.float32).view(np.uint32)
    b = np.asarray(b, np.float32).view(np.uint32)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} of {a.size} differ, first at {tuple(bad[0])}"


CASES = [
    # W, H, D, min_d, cs, roi_left, roi_right
    (450, 375, 64, 0, 1, (40, 30, 300, 200), None),
    (450, 375, 64, 0, 1, (40, 31, 301, 201), (20, 10, 350, 250)),
    (333, 121, 128, 5, 1, (-10, -10, 200, 90), (100, 50, 500, 500)),
    (200, 90, 64, -7, 0, (3, 2, 150, 77), (0, 0, 120, 60)),
    (160, 100, 256, 0, 1, (10, 10, 140, 81), None),
    (300, 80, 512, 12, 1, (0, 5, 290, 70), None),
    (120, 70, 64, 20, 1, (5, 5, 60, 40), None),            # starts left of min_d
    (120, 70, 64, 0, 1, (50, 20, 0, 30), (10, 10, 50, 50)),  # empty left region
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}D{c[2]}m{c[3]}cs{c[4]}-{i}"
                                             for i, c in enumerate(CASES)])
@pytest.mark.parametrize("lr,median", [(True, True), (False, False)])
def test_match_with_regions_matches_oracle(cuda, case, lr, median):
    W, H, D, min_d, cs, rl, rr = case
    print(f"[roi] {case} lr={lr} median={median}", flush=True)
    p = scenes.cone_like(W + H, H=H, W=W, D=min(D, 128))
    exp_l, exp_r = oracle.match_roi(p.left, p.right, min_d, D, rl, rr, cs_type=cs,
                                    do_lr_check=lr, do_median=median)
    st = StereoCuda()
    assert st.Init(W, H, min_d, D, CuSGMOption.benchmark(cs_type=cs, do_lr_check=lr,
                                                         do_median_filter=median)), st.last_error
    dl = np.full((H, W), 3.0, np.float32)
    dr = np.full((H, W), 3.0, np.float32)
    assert st.Match(p.left, p.right, dl, dr, None, rl, rr), st.last_error
    _eq(dl, exp_l, "left map")
    _eq(dr, exp_r, "right map")
    # the same handle matches without regions afterwards as before
    full_l, full_r = oracle.match(p.left, p.right, min_d, D, cs_type=cs, do_lr_check=lr,
                                  do_median=median)
    assert st.Match(p.left, p.right, dl, dr), st.last_error
    _eq(dl, full_l, "left map without regions")
    st.Release()


def test_regions_with_peaks_and_depth(cuda):
    W, H, D = 400, 240, 64
    p = scenes.cone_like(7, H=H, W=W, D=D)
    rl, rr = (30, 20, 340, 190), (10, 10, 300, 200)
    print("[roi] peaks and Match2", flush=True)
    opt = CuSGMOption.benchmark(do_remove_peaks=True, peaks_ratio_threshold=0.002)
    exp_l, _ = oracle.match_roi(p.left, p.right, 0, D, rl, rr, peaks_ratio=0.002)
    st = StereoCuda()
    assert st.Init(W, H, 0, D, opt), st.last_error
    dl = np.empty((H, W), np.float32)
    assert st.Match(p.left, p.right, dl, None, None, rl, rr), st.last_error
    _eq(dl, exp_l, "left map with peaks removed in the region's rows")
    st.Release()
    cam = CamParam(200.0, 0.0, 190.0, 0.0, 400.0, 0.1)
    st2 = StereoCuda()
    assert st2.Init2(W, H, 0, D, opt, cam), st2.last_error
    z = np.empty((H, W), np.float32)
    assert st2.Match2(p.left, p.right, z, None, rl, rr), st2.last_error
    _eq(z, oracle.depth(exp_l, (200.0, 0.0, 190.0, 0.0, 400.0, 0.1)), "Match2 with regions")


def test_region_argument_checks(cuda):
    W, H = 96, 64
    p = scenes.cone_like(8, H=H, W=W, D=64)
    st = StereoCuda()
    assert st.Init(W, H, 0, 64, CuSGMOption.benchmark()), st.last_error
    dl = np.empty((H, W), np.float32)
    assert not st.Match(p.left, p.right, dl, None, None, (0, 0, -5, 10))
    assert "negative" in st.last_error
    win = StereoCuda()
    assert win.Init(W, H, 0, 32, CuSGMOption.benchmark(do_lr_check=False)), win.last_error
    init = np.zeros((H, W), np.int16)
    assert not win.Match(p.left, p.right, dl, None, init, (0, 0, 10, 10))
    assert "initial disparity" in win.last_error
