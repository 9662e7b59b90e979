"""Parity of the HIP plain path with the CPU oracle, stage by stage and end
to end.  Integer stages (census, cost, per-path aggregation) must match bit
for bit; the float disparity maps must match bit for bit as well (the HIP
kernels reproduce the reference's float steps exactly), which is stricter
than the 1e-5 px bound on subpixel output."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest
import torch

import oracle
from fasterstereocuda_amd import _lib, scenes, stages
from fasterstereocuda_amd.stereo import CuSGMOption, StereoCuda

pytestmark = pytest.mark.gpu

# (W, H, D, min_d, cs_type, p1, p2): the 450x375 workload, ragged sizes,
# every supported D, the 5x5 census, positive / negative min_d, P1 > 127,
# large penalties (path values past 2047 and past 8 bits of P2), and images
# narrower than D and smaller than one scanline group.
CASES = [
    (450, 375, 64, 0, 1, 10, 150),
    (97, 61, 64, 0, 1, 10, 150),
    (130, 45, 128, 0, 1, 10, 150),
    (300, 70, 256, 0, 1, 10, 150),
    (121, 50, 64, 0, 0, 8, 32),
    (200, 33, 64, 7, 1, 10, 150),
    (150, 40, 64, -5, 1, 10, 150),
    (90, 30, 64, 0, 1, 200, 300),
    (600, 21, 512, 0, 1, 10, 150),
    (77, 9, 128, 3, 0, 10, 150),
    (140, 40, 64, 0, 1, 83, 1900),
    (140, 40, 128, 0, 1, 84, 1900),
    (120, 36, 64, 0, 1, 10, 5000),
    (12, 8, 64, 0, 1, 10, 150),
    (20, 12, 256, 3, 0, 10, 150),
]
UNIQ, LR_THR = 0.95, 1.0


def _ids(c):
    return "W{}H{}D{}m{}cs{}p{}-{}".format(*c)


def _pair(W, H, D, seed=0):
    p = scenes.cone_like(seed, H=H, W=W, D=D)
    return p.left, p.right


def _eq_float(a: np.ndarray, b: np.ndarray, what: str):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert a.shape == b.shape, what
    same_bits = a.view(np.uint32) == b.view(np.uint32)
    if not same_bits.all():
        bad = np.argwhere(~same_bits)[:5]
        pytest.fail(f"{what}: {int((~same_bits).sum())} of {a.size} differ, e.g. "
                    + ", ".join(f"{tuple(i)}: gpu {a[tuple(i)]!r} oracle {b[tuple(i)]!r}" for i in bad))


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_stages_match_oracle(cuda, case):
    W, H, D, min_d, cs, p1, p2 = case
    print(f"[stages] {case}", flush=True)
    left, right = _pair(W, H, D)
    ref = oracle.stages(left, right, min_d, D, cs_type=cs, p1=p1, p2=p2, uniqueness=UNIQ,
                        lr_threshold=LR_THR)
    roi_x = max(0, min_d)
    L = torch.from_numpy(left).to(cuda)
    R = torch.from_numpy(right).to(cuda)
    cl = stages.census(L, cs, False)
    cr = stages.census(R, cs, True)
    np.testing.assert_array_equal(cl.cpu().numpy().view(np.uint64), ref["census_l"])
    np.testing.assert_array_equal(cr.cpu().numpy().view(np.uint64), ref["census_r"])

    for view, (cb, cm, l2r) in (("l", (cl, cr, True)), ("r", (cr, cl, False))):
        c = stages.cost(cb, cm, min_d, D, l2r, cs)
        np.testing.assert_array_equal(c.cpu().numpy()[:, roi_x:], ref[f"cost_{view}"][:, roi_x:],
                                      err_msg=f"cost {view}")
        paths = stages.aggregate(c, min_d, p1, p2)
        for r in range(4):
            np.testing.assert_array_equal(paths[r].cpu().numpy()[:, roi_x:],
                                          ref[f"paths_{view}"][r][:, roi_x:],
                                          err_msg=f"path {stages.DIRS[r]} view {view}")
        d = stages.wta(paths, min_d, UNIQ)
        _eq_float(d.cpu().numpy(), ref[f"wta_{view}"], f"wta {view}")

    wl = torch.from_numpy(ref["wta_l"]).to(cuda)
    wr = torch.from_numpy(ref["wta_r"]).to(cuda)
    out_l, out_r = stages.post(wl, wr, True, LR_THR, True, want_right=True)
    _eq_float(out_l.cpu().numpy(), ref["median"], "LR + median")
    _eq_float(out_r.cpu().numpy(), ref["lr_r"], "right map after LR")
    lr_only, _ = stages.post(wl, wr, True, LR_THR, False)
    _eq_float(lr_only.cpu().numpy(), ref["lr_l"], "LR only")
    med_only, _ = stages.post(wl, None, False, LR_THR, True)
    _eq_float(med_only.cpu().numpy(), oracle.median3x3(ref["wta_l"]), "median only")


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_on_the_fly_stages_match_oracle(cuda, case):
    """The kernels the batched benchmark runs (cost from the census words
    inside the aggregation, bottom->top fused with the WTA): their stored
    top->bottom, left->right and right->left volumes and the raw WTA map of
    both views, bit for bit against the oracle's stages (no median)."""
    W, H, D, min_d, cs, p1, p2 = case
    if D > 256:
        pytest.skip("the on-the-fly cost takes D <= 256")
    print(f"[otf stages] {case}", flush=True)
    left, right = _pair(W, H, D, seed=7)
    ref = oracle.stages(left, right, min_d, D, cs_type=cs, p1=p1, p2=p2, uniqueness=UNIQ,
                        lr_threshold=LR_THR)
    roi_x = max(0, min_d)
    cl = torch.from_numpy(ref["census_l"].view(np.int64)).to(cuda)
    cr = torch.from_numpy(ref["census_r"].view(np.int64)).to(cuda)
    views = [("l", cl, cr, False)] + ([] if cs == 0 else [("r", cr, cl, True)])
    for view, cb, cm, is_right in views:
        paths, raw = stages.aggregate_wta_otf(cb, cm, min_d, D, is_right, cs, p1, p2, UNIQ)
        for r in (0, 1, 3):
            np.testing.assert_array_equal(paths[r].cpu().numpy()[:, roi_x:],
                                          ref[f"paths_{view}"][r][:, roi_x:],
                                          err_msg=f"path {stages.DIRS[r]} view {view}")
        _eq_float(raw.cpu().numpy(), ref[f"wta_{view}"], f"raw wta {view}")


@pytest.mark.parametrize("case", CASES[:6], ids=_ids)
@pytest.mark.parametrize("lr,median", [(True, True), (False, True), (True, False)])
def test_stereocuda_match_matches_oracle(cuda, case, lr, median):
    W, H, D, min_d, cs, p1, p2 = case
    print(f"[match] {case} lr={lr} median={median}", flush=True)
    left, right = _pair(W, H, D, seed=3)
    exp_l, exp_r = oracle.match(left, right, min_d, D, cs_type=cs, p1=p1, p2=p2, uniqueness=UNIQ,
                                do_lr_check=lr, lr_threshold=LR_THR, do_median=median)
    st = StereoCuda()
    opt = CuSGMOption.benchmark(cs_type=cs, p1=p1, p2_init=p2, do_lr_check=lr,
                                do_median_filter=median)
    assert st.Init(W, H, min_d, D, opt), st.last_error
    dl = np.empty((H, W), np.float32)
    dr = np.empty((H, W), np.float32)
    for _ in range(2):   # run twice: results must not depend on buffer history
        assert st.Match(left, right, dl, dr), st.last_error
        _eq_float(dl, exp_l, "left map")
        _eq_float(dr, exp_r, "right map")
    st.Release()


@pytest.fixture
def cost_mode():
    from fasterstereocuda_amd import stereo
    before = stereo.get_cost_mode()
    yield stereo.set_cost_mode
    stereo.set_cost_mode(before)


@pytest.mark.parametrize("mode", [2, 1], ids=["on-the-fly", "stored"])
@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("lr,median", [(True, True), (False, True), (True, False), (False, False)])
def test_cost_modes_match_oracle(cuda, cost_mode, mode, case, lr, median):
    """The on-the-fly cost kernels (census words read inside the aggregation,
    no cost volume) and the stored-cost kernels give the oracle's maps, for
    every D, both census sizes, +-min_d, P1 > 127 and ragged sizes, with and
    without the median (which could hide an isolated wrong disparity)."""
    W, H, D, min_d, cs, p1, p2 = case
    print(f"[cost mode {mode}] {case} lr={lr} median={median}", flush=True)
    cost_mode(mode)
    left, right = _pair(W, H, D, seed=5)
    exp_l, exp_r = oracle.match(left, right, min_d, D, cs_type=cs, p1=p1, p2=p2, uniqueness=UNIQ,
                                do_lr_check=lr, lr_threshold=LR_THR, do_median=median)
    st = StereoCuda()
    opt = CuSGMOption.benchmark(cs_type=cs, p1=p1, p2_init=p2, do_lr_check=lr,
                                do_median_filter=median)
    assert st.Init(W, H, min_d, D, opt), st.last_error
    dl = np.empty((H, W), np.float32)
    dr = np.empty((H, W), np.float32)
    assert st.Match(left, right, dl, dr), st.last_error
    _eq_float(dl, exp_l, "left map")
    _eq_float(dr, exp_r, "right map")
    st.Release()


@pytest.fixture
def wta_form():
    before = stages.get_wta_form()
    yield stages.set_wta_form
    stages.set_wta_form(before)


@pytest.mark.parametrize("form", [stages.WTA_FUSED, stages.WTA_SEPARATE], ids=["fused", "separate"])
@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("lr,median", [(True, True), (False, False)])
def test_wta_forms_match_oracle(cuda, cost_mode, wta_form, form, case, lr, median):
    """The stored-cost path with bottom->top fused into the WTA and with the
    four directions followed by the standalone WTA kernel gives the oracle's
    maps."""
    W, H, D, min_d, cs, p1, p2 = case
    print(f"[wta form {form}] {case} lr={lr} median={median}", flush=True)
    cost_mode(1)
    wta_form(form)
    left, right = _pair(W, H, D, seed=7)
    exp_l, exp_r = oracle.match(left, right, min_d, D, cs_type=cs, p1=p1, p2=p2, uniqueness=UNIQ,
                                do_lr_check=lr, lr_threshold=LR_THR, do_median=median)
    st = StereoCuda()
    opt = CuSGMOption.benchmark(cs_type=cs, p1=p1, p2_init=p2, do_lr_check=lr,
                                do_median_filter=median)
    assert st.Init(W, H, min_d, D, opt), st.last_error
    dl = np.empty((H, W), np.float32)
    dr = np.empty((H, W), np.float32)
    assert st.Match(left, right, dl, dr), st.last_error
    _eq_float(dl, exp_l, "left map")
    _eq_float(dr, exp_r, "right map")
    st.Release()


def test_driving_full_size(cuda):
    pair = scenes.driving(0)
    print("[match] driving 1242x375", flush=True)
    exp_l, exp_r = oracle.match(pair.left, pair.right, 0, 64)
    st = StereoCuda()
    assert st.Init(1242, 375, 0, 64, CuSGMOption.benchmark()), st.last_error
    dl = np.empty((375, 1242), np.float32)
    dr = np.empty((375, 1242), np.float32)
    assert st.Match(pair.left, pair.right, dl, dr), st.last_error
    _eq_float(dl, exp_l, "left map")
    _eq_float(dr, exp_r, "right map")
    # device-resident entry point gives the same bytes
    Ld = torch.from_numpy(pair.left).to(cuda)
    Rd = torch.from_numpy(pair.right).to(cuda)
    od = torch.empty((375, 1242), dtype=torch.float32, device=cuda)
    assert st.match_device(Ld, Rd, od)
    torch.cuda.synchronize()
    _eq_float(od.cpu().numpy(), exp_l, "device entry point")


@pytest.mark.parametrize("cs,lr", [(1, True), (0, True), (1, False)])
def test_match_device_binds_caller_buffers(cuda, cs, lr):
    """match_device reads the caller's images and writes the caller's maps in
    place (no staging copies): alternating buffer sets, with and without the
    right map and peak removal, interleaved with host Match calls."""
    W, H, D = 220, 90, 64
    opt = CuSGMOption.benchmark(cs_type=cs, do_lr_check=lr, do_remove_peaks=True,
                                peaks_ratio_threshold=0.002)
    st = StereoCuda()
    assert st.Init(W, H, 0, D, opt), st.last_error
    pairs = [_pair(W, H, D, seed=s) for s in (21, 22)]
    exps = [oracle.match(l, r, 0, D, cs_type=cs, do_lr_check=lr, peaks_ratio=0.002) for l, r in pairs]
    bufs = [(torch.from_numpy(l).to(cuda), torch.from_numpy(r).to(cuda),
             torch.full((H, W), 7.0, device=cuda), torch.full((H, W), 7.0, device=cuda))
            for l, r in pairs]
    for rep in range(2):
        for k, ((L, R, ol, orr), (el, er)) in enumerate(zip(bufs, exps)):
            print(f"[match_device] cs={cs} lr={lr} rep={rep} pair={k}", flush=True)
            want_right = (rep + k) % 2 == 0
            assert st.match_device(L, R, ol, orr if want_right else None), st.last_error
            torch.cuda.synchronize()
            _eq_float(ol.cpu().numpy(), el, "device left")
            if want_right:
                _eq_float(orr.cpu().numpy(), er, "device right")
            dl = np.empty((H, W), np.float32)
            assert st.Match(pairs[k][0], pairs[k][1], dl), st.last_error
            _eq_float(dl, el, "host left after a device match")


def test_set_min_disparity_rebuilds_the_plan(cuda):
    """SetMinDisparity after Init (StereoCuda.h:71-72): each later Match uses
    the new minimum; positive, negative and back to 0 on one handle."""
    W, H, D = 200, 60, 64
    left, right = _pair(W, H, D, seed=11)
    st = StereoCuda()
    assert st.Init(W, H, 0, D, CuSGMOption.benchmark()), st.last_error
    for md in (9, -6, 0):
        print(f"[min disparity] {md}", flush=True)
        st.SetMinDisparity(md)
        exp_l, exp_r = oracle.match(left, right, md, D, uniqueness=UNIQ, lr_threshold=LR_THR)
        dl = np.empty((H, W), np.float32)
        dr = np.empty((H, W), np.float32)
        assert st.Match(left, right, dl, dr), st.last_error
        _eq_float(dl, exp_l, f"left map, min_d {md}")
        _eq_float(dr, exp_r, f"right map, min_d {md}")
    assert st.GetDispartyRange() == D


def test_match_from_page_locked_buffers(cuda):
    """The reference's calling pattern: images and maps in page-locked host
    memory from MallocPageLockedPtr (StereoCuda.h:94-97, sgm_host_alloc)."""
    W, H, D = 450, 375, 64
    left, right = _pair(W, H, D, seed=3)
    lib = _lib.lib()
    ptrs = []

    def pinned(dtype):
        nbytes = H * W * np.dtype(dtype).itemsize
        p = ctypes.c_void_p()
        assert lib.sgm_host_alloc(ctypes.byref(p), nbytes) == 0, _lib.last_error()
        ptrs.append(p)
        raw = (ctypes.c_uint8 * nbytes).from_address(p.value)
        return np.ctypeslib.as_array(raw).view(dtype).reshape(H, W)

    try:
        L, R = pinned(np.uint8), pinned(np.uint8)
        dl, dr = pinned(np.float32), pinned(np.float32)
        L[:] = left
        R[:] = right
        st = StereoCuda()
        assert st.Init(W, H, 0, D, CuSGMOption.benchmark()), st.last_error
        print("[match] page-locked 450x375", flush=True)
        assert st.Match(L, R, dl, dr), st.last_error
        got_l, got_r = dl.copy(), dr.copy()
        st.Release()
    finally:
        for p in ptrs:
            assert lib.sgm_host_free(p) == 0
    exp_l, exp_r = oracle.match(left, right, 0, D, uniqueness=UNIQ, lr_threshold=LR_THR)
    _eq_float(got_l, exp_l, "left map")
    _eq_float(got_r, exp_r, "right map")


def test_cpp_dropin_demo(cuda, tmp_path):
    """The C++ StereoCuda class (libStereoCuda.so) used the way a reference
    caller uses it, through the stereo_demo program."""
    demo = os.path.join(os.path.dirname(oracle.__file__), "..", "fasterstereocuda_amd", "stereo_demo")
    pair = scenes.cone_like(7)
    H, W = pair.left.shape
    lp, rp = tmp_path / "l.u8", tmp_path / "r.u8"
    ol, orr = tmp_path / "dl.f32", tmp_path / "dr.f32"
    pair.left.tofile(lp)
    pair.right.tofile(rp)
    print("[demo] cone 450x375", flush=True)
    res = subprocess.run([demo, str(lp), str(rp), str(W), str(H), "0", "64", str(ol), str(orr)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr + res.stdout
    exp_l, exp_r = oracle.match(pair.left, pair.right, 0, 64)
    _eq_float(np.fromfile(ol, np.float32).reshape(H, W), exp_l, "demo left")
    _eq_float(np.fromfile(orr, np.float32).reshape(H, W), exp_r, "demo right")


@pytest.mark.parametrize("layers", [1, 2, 3])
def test_cpp_faster_stereo_cuda_demo(cuda, tmp_path, layers):
    """The C++ FasterStereoCuda class (libFasterStereoCuda.so) through
    stereo_demo --layers: the reference's pyramid schedule, against the oracle."""
    demo = os.path.join(os.path.dirname(oracle.__file__), "..", "fasterstereocuda_amd", "stereo_demo")
    pair = scenes.cone_like(3, H=300, W=400, D=128, workload=2)
    H, W = pair.left.shape
    lp, rp, ol = tmp_path / "l.u8", tmp_path / "r.u8", tmp_path / "d.f32"
    pair.left.tofile(lp)
    pair.right.tofile(rp)
    print(f"[demo] FasterStereoCuda 400x300 D128 layers={layers}", flush=True)
    res = subprocess.run([demo, str(lp), str(rp), str(W), str(H), "0", "128", str(ol),
                          "--layers", str(layers)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr + res.stdout
    exp = oracle.pyramid(pair.left, pair.right, 0, 128, layers, 32, cs_type=0, p1=8, p2=32)
    _eq_float(np.fromfile(ol, np.float32).reshape(H, W), exp, "FasterStereoCuda demo")


def test_errors_are_reported(cuda):
    st = StereoCuda()
    assert not st.Init(100, 50, 0, 96, CuSGMOption.benchmark())
    assert "disp_range" in st.last_error
    assert not st.Init(100, 50, 0, 64, CuSGMOption.benchmark(do_remove_peaks=True,
                                                            peaks_ratio_threshold=1.5))
    assert "peaks_ratio_threshold" in st.last_error
    left = np.zeros((50, 100), np.uint8)
    out = np.empty((50, 100), np.float32)
    fresh = StereoCuda()
    fresh.width, fresh.height = 100, 50
    assert not fresh.Match(left, left, out)        # not initialised
    assert "not initialised" in fresh.last_error
    assert st.Init(100, 50, 0, 64, CuSGMOption.benchmark())
    assert st.Match(left, left, out)
    assert st.GetDispartyRange() == 64
    st.Release()
    st.Release()                                   # double release is safe
