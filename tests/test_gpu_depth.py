"""Depth output (Init2 / Match2) on the GPU against the oracle, bit for bit:
the depth kernel on raw disparities, the plain handle, the pyramid with its
coarsest interval taken from a depth interval, and the C++ classes through
stereo_demo --cam."""
import os
import subprocess

import numpy as np
import pytest
import torch

import oracle
from fasterstereocuda_amd import scenes
from fasterstereocuda_amd.pyramid import (EpipolarStereoParam, FasterStereoCuda, SgmPyramid,
                                          StereoOption2, reference_schedule)
from fasterstereocuda_amd.stereo import CamParam, CuSGMOption, StereoCuda, depth_device

pytestmark = pytest.mark.gpu

# KITTI-like: focus and principal points in pixels, baseline in metres
KITTI = CamParam(x0_left=612.5, y0_left=173.0, x0_right=598.25, y0_right=173.0, focus=721.5,
                 baseline=0.54)
DEMO = os.path.join(os.path.dirname(oracle.__file__), "..", "fasterstereocuda_amd", "stereo_demo")


def _t(c):
    return (c.x0_left, c.y0_left, c.x0_right, c.y0_right, c.focus, c.baseline)


def _eq(a, b, what):
    a = np.asarray(a, np.float32).view(np.uint32)
    b = np.asarray(b, np.float32).view(np.uint32)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} of {a.size} differ, first at {tuple(bad[0])}"


def test_depth_kernel_matches_oracle(cuda):
    print("[depth kernel] 97x131 random disparities", flush=True)
    rng = np.random.default_rng(3)
    d = rng.uniform(-5.0, 80.0, size=(97, 131)).astype(np.float32)
    d[rng.random(d.shape) < 0.2] = oracle.INVALID
    # zero denominator (d = x0l - x0r = 14.25), signed zeros, tiny and huge values
    d[0, :7] = [14.25, 0.0, -0.0, 1e-7, 64.0, 63.5, 1e30]
    g = torch.from_numpy(d).to(cuda)
    out = torch.empty_like(g)
    depth_device(g, out, KITTI)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    _eq(got, oracle.depth(d, _t(KITTI)), "depth kernel")
    assert np.isposinf(got[0, 0]) and got[1][d[1] == oracle.INVALID].tolist() == \
        [oracle.INVALID] * int((d[1] == oracle.INVALID).sum())


@pytest.mark.parametrize("min_d", [0, 8])
def test_stereocuda_match2_matches_oracle(cuda, min_d):
    print(f"[match2] 200x120 D64 min_d={min_d}", flush=True)
    pair = scenes.cone_like(5, H=120, W=200, D=64)
    H, W = pair.left.shape
    st = StereoCuda()
    assert st.Init2(W, H, min_d, 64, CuSGMOption.benchmark(), KITTI), st.last_error
    depth = np.empty((H, W), np.float32)
    assert st.Match2(pair.left, pair.right, depth), st.last_error
    disp, _ = oracle.match(pair.left, pair.right, min_d, 64)
    _eq(depth, oracle.depth(disp, _t(KITTI)), "Match2 depth")
    # the same handle still matches disparities, and Init (no camera) refuses Match2
    dl = np.empty((H, W), np.float32)
    assert st.Match(pair.left, pair.right, dl)
    _eq(dl, disp, "Match after Init2")
    assert st.Init(W, H, min_d, 64, CuSGMOption.benchmark())
    assert not st.Match2(pair.left, pair.right, depth)
    assert "Init2" in st.last_error


# (layers, camera, depth interval): f2 = 0 and f2 = 10 at full resolution
PYR_CASES = [
    (2, CamParam(200.0, 0.0, 200.0, 0.0, 400.0, 10.0), (4000.0 / 128.0, 4000.0)),
    (2, CamParam(190.0, 0.0, 200.0, 0.0, 400.0, 10.0), (4000.0 / 128.0, 4000.0)),
    (1, CamParam(200.0, 0.0, 200.0, 0.0, 400.0, 10.0), (4000.0 / 100.0, 400.0)),
    (3, CamParam(201.5, 0.0, 200.0, 0.0, 400.0, 10.0), (4000.0 / 120.0, 1000.0)),
]


@pytest.mark.parametrize("layers,cam,zr", PYR_CASES)
def test_pyramid_match2_matches_oracle(cuda, layers, cam, zr):
    print(f"[pyramid match2] 400x300 layers={layers} cam={_t(cam)} z={zr}", flush=True)
    pair = scenes.cone_like(3, H=300, W=400, D=128, workload=2)
    H, W = pair.left.shape
    o = StereoOption2(width=W, height=H, min_depth=zr[0], max_depth=zr[1], num_layers=layers,
                      do_lr_check=True, do_rm_peaks=False, do_smooth=False,
                      epi=EpipolarStereoParam(x0_left=cam.x0_left, x0_right=cam.x0_right,
                                              focus=cam.focus, baseline=cam.baseline))
    dmin, drange = oracle.depth_interval(zr[0], zr[1], W, layers, _t(cam))
    pyr = SgmPyramid.from_depth(W, H, zr[0], zr[1], layers, 32, reference_schedule(o), cam)
    assert pyr.coarsest == (dmin, drange)
    fs = FasterStereoCuda()
    assert fs.Init2(o), fs.last_error
    depth = np.empty((H, W), np.float32)
    assert fs.Match2(pair.left, pair.right, depth), fs.last_error
    disp = oracle.pyramid_at(pair.left, pair.right, dmin, drange, layers, 32, cs_type=0, p1=8, p2=32)
    exp = oracle.depth(disp, _t(cam))
    _eq(depth, exp, "FasterStereoCuda.Match2")
    # device form, on torch's stream
    g = torch.empty((H, W), dtype=torch.float32, device=cuda)
    pyr.match2_device(torch.from_numpy(pair.left).to(cuda), torch.from_numpy(pair.right).to(cuda), g)
    torch.cuda.synchronize()
    _eq(g.cpu().numpy(), exp, "match2_device")
    assert (exp != oracle.INVALID).mean() > 0.5


def test_pyramid_depth_interval_errors(cuda):
    opt = reference_schedule(StereoOption2(do_rm_peaks=False))
    cam = CamParam(200.0, 0.0, 200.0, 0.0, 400.0, 10.0)
    for zr in ((0.0, 100.0), (50.0, 50.0), (-1.0, 10.0)):
        with pytest.raises(RuntimeError, match="depth interval"):
            SgmPyramid.from_depth(400, 300, zr[0], zr[1], 2, 32, opt, cam)
    # a disparity pyramid has no camera
    pyr = SgmPyramid(400, 300, 0, 128, 2, 32, opt)
    with pytest.raises(Exception, match="Init2"):
        pyr.match2(np.zeros((300, 400), np.uint8), np.zeros((300, 400), np.uint8))


@pytest.mark.parametrize("layers", [0, 2])
def test_cpp_depth_demo(cuda, tmp_path, layers):
    """StereoCuda::Init2/Match2 (layers 0) and FasterStereoCuda::Init2/Match2
    through stereo_demo --cam, against the oracle."""
    pair = scenes.cone_like(3, H=300, W=400, D=128, workload=2)
    H, W = pair.left.shape
    lp, rp, out = tmp_path / "l.u8", tmp_path / "r.u8", tmp_path / "z.f32"
    pair.left.tofile(lp)
    pair.right.tofile(rp)
    cam = CamParam(200.0, 0.0, 200.0, 0.0, 400.0, 10.0)
    args = [DEMO, str(lp), str(rp), str(W), str(H), "0", "128", str(out),
            "--cam", "200,200,400,10"]
    if layers:
        args += ["--layers", str(layers), "--zrange", f"{4000.0 / 128.0},4000"]
    print(f"[demo depth] layers={layers}", flush=True)
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr + res.stdout
    got = np.fromfile(out, np.float32).reshape(H, W)
    if layers:
        dmin, drange = oracle.depth_interval(np.float32(4000.0 / 128.0), 4000.0, W, layers, _t(cam))
        disp = oracle.pyramid_at(pair.left, pair.right, dmin, drange, layers, 32, cs_type=0, p1=8,
                                 p2=32)
    else:
        disp, _ = oracle.match(pair.left, pair.right, 0, 128)   # the demo defaults: 9x7, 10/150
    _eq(got, oracle.depth(disp, _t(cam)), "demo depth")
