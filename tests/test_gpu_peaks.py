"""Peak removal on the GPU against the oracle, bit for bit: the union-find
kernel on adversarial maps (one and several strips), the plain handle with
do_remove_peaks, Match2 after it, the pyramid's coarsest level, and the C++
classes through stereo_demo --peaks."""
import os
import subprocess

import numpy as np
import pytest
import torch

import oracle
from fasterstereocuda_amd import scenes, stages
from fasterstereocuda_amd.pyramid import FasterStereoCuda, StereoOption1
from fasterstereocuda_amd.stereo import (CamParam, CuSGMOption, StereoCuda, get_peak_strips,
                                         set_peak_strips)

pytestmark = pytest.mark.gpu
DEMO = os.path.join(os.path.dirname(oracle.__file__), "..", "fasterstereocuda_amd", "stereo_demo")


def _eq(a, b, what):
    a = np.asarray(a, np.float32).view(np.uint32)
    b = np.asarray(b, np.float32).view(np.uint32)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} of {a.size} differ, first at {tuple(bad[0])}"


@pytest.fixture
def strips():
    """Sets the process-wide strip count for one test and restores 1."""
    def set_(n):
        set_peak_strips(n)
        assert get_peak_strips() == n
    yield set_
    set_peak_strips(1)


def _map(rng, H, W, levels, invalid):
    d = rng.integers(0, levels, size=(H, W)).astype(np.float32) * 0.75
    # smooth patches so that segments of every size appear
    d = np.repeat(np.repeat(d[::3, ::3], 3, 0), 3, 1)[:H, :W].copy()
    d += rng.choice([0.0, 0.5, 1.25], size=(H, W), p=[0.8, 0.15, 0.05]).astype(np.float32)
    d[rng.random((H, W)) < invalid] = oracle.INVALID
    return d


@pytest.mark.parametrize("H,W,nstrips,ratio", [
    (61, 97, 1, 0.002), (61, 97, 4, 0.01), (375, 1242, 1, 0.0005), (375, 1242, 18, 0.005),
    (852, 1200, 7, 0.001), (33, 17, 40, 0.5),
])
def test_remove_peaks_kernel_matches_oracle(cuda, H, W, nstrips, ratio):
    print(f"[peaks kernel] {W}x{H} strips={nstrips} ratio={ratio}", flush=True)
    rng = np.random.default_rng(H * W + nstrips)
    d = _map(rng, H, W, 12, 0.1)
    exp, n = oracle.remove_peaks(d, nstrips, ratio)
    got = stages.remove_peaks(torch.from_numpy(d).to(cuda), nstrips, ratio).cpu().numpy()
    _eq(got, exp, "remove_peaks")
    if nstrips <= H:
        assert n > 0


@pytest.mark.parametrize("nstrips", [1, 6])
def test_stereocuda_remove_peaks_matches_oracle(cuda, strips, nstrips):
    print(f"[match + peaks] 450x375 D64 strips={nstrips}", flush=True)
    strips(nstrips)
    pair = scenes.cone_like(1, D=64)
    H, W = pair.left.shape
    ratio = 0.002
    opt = CuSGMOption.benchmark(do_remove_peaks=True, peaks_ratio_threshold=ratio)
    st = StereoCuda()
    assert st.Init(W, H, 0, 64, opt), st.last_error
    dl, dr = np.empty((H, W), np.float32), np.empty((H, W), np.float32)
    assert st.Match(pair.left, pair.right, dl, dr), st.last_error
    exp_l, exp_r = oracle.match(pair.left, pair.right, 0, 64, peaks_ratio=ratio, peak_strips=nstrips)
    _eq(dl, exp_l, "left map")
    _eq(dr, exp_r, "right map (no peak removal)")
    plain, _ = oracle.match(pair.left, pair.right, 0, 64)
    assert (exp_l != plain).any(), "the case removes no peak"
    # Match2 converts the map after the removal
    cam = CamParam(300.0, 0.0, 290.0, 0.0, 500.0, 0.2)
    assert st.Init2(W, H, 0, 64, opt, cam), st.last_error
    z = np.empty((H, W), np.float32)
    assert st.Match2(pair.left, pair.right, z), st.last_error
    _eq(z, oracle.depth(exp_l, (300.0, 0.0, 290.0, 0.0, 500.0, 0.2)), "Match2 after peaks")


@pytest.mark.parametrize("layers", [1, 2, 3])
def test_pyramid_remove_peaks_matches_oracle(cuda, layers):
    print(f"[pyramid + peaks] 400x300 D128 layers={layers}", flush=True)
    # scene 6 leaves coarse segments under the 3-level threshold of 37 pixels
    pair = scenes.cone_like(6, H=300, W=400, D=128, workload=2)
    H, W = pair.left.shape
    fs = FasterStereoCuda()
    o = StereoOption1(width=W, height=H, min_disp=0, max_disp=128, num_layers=layers,
                      do_lr_check=True, do_rm_peaks=True, do_smooth=False)
    assert fs.Init(o), fs.last_error
    out = np.empty((H, W), np.float32)
    assert fs.Match(pair.left, pair.right, out), fs.last_error
    ratio = {1: 0.0005, 2: 0.001, 3: 0.005}[layers]
    exp = oracle.pyramid(pair.left, pair.right, 0, 128, layers, 32, cs_type=0, p1=8, p2=32,
                         peaks_ratio=ratio)
    _eq(out, exp, f"pyramid layers={layers}")
    base = oracle.pyramid(pair.left, pair.right, 0, 128, layers, 32, cs_type=0, p1=8, p2=32)
    assert (exp != base).any(), "the case removes no peak"


@pytest.mark.parametrize("layers", [0, 2])
def test_cpp_remove_peaks_demo(cuda, tmp_path, layers):
    pair = scenes.cone_like(3, H=300, W=400, D=128, workload=2)
    H, W = pair.left.shape
    lp, rp, ol = tmp_path / "l.u8", tmp_path / "r.u8", tmp_path / "d.f32"
    pair.left.tofile(lp)
    pair.right.tofile(rp)
    args = [DEMO, str(lp), str(rp), str(W), str(H), "0", "128", str(ol), "--peaks", "0.003"]
    if layers:
        args += ["--layers", str(layers)]
    print(f"[demo peaks] layers={layers}", flush=True)
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr + res.stdout
    got = np.fromfile(ol, np.float32).reshape(H, W)
    if layers:
        exp = oracle.pyramid(pair.left, pair.right, 0, 128, layers, 32, cs_type=0, p1=8, p2=32,
                             peaks_ratio=0.001)
    else:
        exp, _ = oracle.match(pair.left, pair.right, 0, 128, peaks_ratio=0.003)
    _eq(got, exp, "demo")
