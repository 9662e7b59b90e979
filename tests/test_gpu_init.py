"""StereoCuda with an initial disparity map, and with the reference's
default options, on the GPU against the oracle, bit for bit.

The reference's pyramid drives its single-level class exactly this way:
HierSgmCudaImpl::Init gives the finer levels disp_range 32 (LEVEL_RANGE) and
Match passes the propagated map (FasterStereoCuda.cpp:282-289, 455-469), which
reaches the window cost, aggregation and WTA (StereoCudaImpl.cpp:363-446,
cusgm_ca.cu:4052-4077).  A default-constructed CuSGMOption has
using_constant_p2 = false, which the plain aggregation ignores
(cusgm_ca.cu:4080-4135)."""
import os
import subprocess

import numpy as np
import pytest
import torch

import oracle
from fasterstereocuda_amd import scenes
from fasterstereocuda_amd.batch import SgmBatch
from fasterstereocuda_amd.stereo import CamParam, CuSGMOption, StereoCuda

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(oracle.__file__), "..", "fasterstereocuda_amd", "stereo_demo")
INV = oracle.INVALID


def _eq(a, b, what):
    a = np.asarray(a, np.float32).view(np.uint32)
    b = np.asarray(b, np.float32).view(np.uint32)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} of {a.size} differ, first at {tuple(bad[0])}"


def _level(seed, W, H, D=64, holes=0.1):
    """A pair and the initial map the pyramid would give it: the oracle's
    match of the 2x decimated pair, dilated, propagated with window 32, plus
    random INVALID holes."""
    pair = scenes.cone_like(seed, H=H, W=W, D=D)
    zl, zr = oracle.zoom_half(pair.left), oracle.zoom_half(pair.right)
    coarse, _ = oracle.match(zl, zr, 0, 64, uniqueness=1.0)
    coarse = oracle.dilation(coarse, oracle.dilation_window(*coarse.shape[::-1]))
    init = oracle.init_next(coarse, W, H, 32)
    rng = np.random.default_rng(seed)
    init[rng.random(init.shape) < holes] = -32768
    return pair.left, pair.right, init


def _opt(cs, p1, p2, uniq, median, **kw):
    return CuSGMOption(p1=p1, p2_init=p2, cs_type=cs, unique_threshold=uniq,
                       do_median_filter=median, using_constant_p2=True, **kw)


CASES = [  # (W, H, cs, p1, p2, uniqueness, median)
    (240, 160, 1, 10, 150, 0.95, True),
    (240, 160, 1, 10, 150, 1.0, False),
    (181, 97, 0, 8, 32, 1.0, True),
    (64, 40, 1, 200, 300, 0.95, True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "W{}H{}cs{}p{}-{}u{}m{}".format(*c))
def test_match_with_initial_disparity(cuda, case):
    W, H, cs, p1, p2, u, median = case
    print(f"[init] {case}", flush=True)
    left, right, init = _level(11, W, H)
    exp = oracle.match_init(left, right, init, 32, cs_type=cs, p1=p1, p2=p2, uniqueness=u,
                            do_median=median)
    st = StereoCuda()
    assert st.Init(W, H, 0, 32, _opt(cs, p1, p2, u, median)), st.last_error
    dl = np.empty((H, W), np.float32)
    dr = np.empty((H, W), np.float32)
    for _ in range(2):
        assert st.Match(left, right, dl, dr, init), st.last_error
        _eq(dl, exp, "left map")
        assert (dr == INV).all()           # the reference's right map is never written here
    # a negative min_disparity changes nothing: the window ignores it
    st2 = StereoCuda()
    assert st2.Init(W, H, -7, 32, _opt(cs, p1, p2, u, median)), st2.last_error
    assert st2.Match(left, right, dl, None, init), st2.last_error
    _eq(dl, exp, "left map, min_d -7")


def test_match_with_initial_disparity_and_peaks(cuda):
    W, H = 300, 200
    print(f"[init+peaks] {W}x{H}", flush=True)
    left, right, init = _level(4, W, H, holes=0.02)
    exp = oracle.match_init(left, right, init, 32, uniqueness=0.95)
    exp, removed = oracle.remove_peaks(exp, 1, 0.002)
    assert removed > 0
    st = StereoCuda()
    assert st.Init(W, H, 0, 32, _opt(1, 10, 150, 0.95, True, do_remove_peaks=True,
                                     peaks_ratio_threshold=0.002)), st.last_error
    dl = np.empty((H, W), np.float32)
    assert st.Match(left, right, dl, None, init), st.last_error
    _eq(dl, exp, "left map after peak removal")


def test_match2_with_initial_disparity(cuda):
    W, H = 256, 144
    print(f"[init match2] {W}x{H}", flush=True)
    left, right, init = _level(9, W, H)
    cam = CamParam(x0_left=130.0, y0_left=70.0, x0_right=121.5, y0_right=70.0, focus=300.0,
                   baseline=0.5)
    st = StereoCuda()
    assert st.Init2(W, H, 0, 32, _opt(1, 10, 150, 0.95, True), cam), st.last_error
    depth = np.empty((H, W), np.float32)
    assert st.Match2(left, right, depth, init), st.last_error
    disp = oracle.match_init(left, right, init, 32, uniqueness=0.95)
    _eq(depth, oracle.depth(disp, (130.0, 70.0, 121.5, 70.0, 300.0, 0.5)), "depth")


def test_initial_disparity_refusals(cuda):
    """What the reference cannot do, or this build does not reach, is refused
    with a message instead of returning stale or wrong maps."""
    W, H = 96, 64
    left, right, init = _level(2, W, H)
    dl = np.empty((H, W), np.float32)
    st = StereoCuda()
    assert st.Init(W, H, 0, 64, _opt(1, 10, 150, 0.95, True))
    assert not st.Match(left, right, dl, None, init)
    assert "disp_range 32" in st.last_error
    st = StereoCuda()
    assert st.Init(W, H, 0, 32, _opt(1, 10, 150, 0.95, True, do_lr_check=True, lr_threshold=1.0))
    assert not st.Match(left, right, dl, None, init)
    assert "LR check" in st.last_error
    st = StereoCuda()
    assert st.Init(W, H, 0, 32, _opt(1, 10, 150, 0.95, True))
    assert not st.Match(left, right, dl)                    # no map: plain D = 32
    assert "initial disparity" in st.last_error
    d_l = torch.from_numpy(left).to(cuda)
    d_o = torch.empty((H, W), dtype=torch.float32, device=cuda)
    assert not st.match_device(d_l, d_l, d_o)
    st = StereoCuda()
    o = _opt(1, 10, 150, 0.95, True)
    o.using_constant_p2 = False
    assert st.Init(W, H, 0, 32, o)
    assert not st.Match(left, right, dl, None, init)
    assert "constant P2" in st.last_error


def test_cpp_match_with_initial_disparity(cuda, tmp_path):
    """StereoCuda::Match(L, R, disp, nullptr, init) from C++, as
    HierSgmCudaImpl::Match calls it (FasterStereoCuda.cpp:455-469)."""
    W, H = 320, 180
    left, right, init = _level(21, W, H)
    lp, rp, ip, op = (tmp_path / n for n in ("l.u8", "r.u8", "i.i16", "d.f32"))
    left.tofile(lp)
    right.tofile(rp)
    init.tofile(ip)
    print(f"[demo init] {W}x{H}", flush=True)
    res = subprocess.run([DEMO, str(lp), str(rp), str(W), str(H), "0", "32", str(op), "--lr", "0",
                          "--uniq", "1.0", "--init", str(ip)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr + res.stdout
    exp = oracle.match_init(left, right, init, 32, uniqueness=1.0)
    _eq(np.fromfile(op, np.float32).reshape(H, W), exp, "demo init")


@pytest.mark.parametrize("cs", [0, 1])
def test_cpp_default_option(cuda, tmp_path, cs):
    """A C++ caller that default-constructs CuSGMOption and sets only P1, P2
    and the census type (using_constant_p2 stays false, uniqueness 0, no LR
    check, no median)."""
    pair = scenes.cone_like(8, H=150, W=260, D=64)
    H, W = pair.left.shape
    lp, rp, op = tmp_path / "l.u8", tmp_path / "r.u8", tmp_path / "d.f32"
    pair.left.tofile(lp)
    pair.right.tofile(rp)
    print(f"[demo default option] cs={cs}", flush=True)
    res = subprocess.run([DEMO, str(lp), str(rp), str(W), str(H), "0", "64", str(op),
                          "--default-opt", "1", "--cs", str(cs), "--p1", "10", "--p2", "150"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr + res.stdout
    exp, _ = oracle.match(pair.left, pair.right, 0, 64, cs_type=cs, p1=10, p2=150,
                          uniqueness=0.0, do_lr_check=False, do_median=False)
    _eq(np.fromfile(op, np.float32).reshape(H, W), exp, "default option")
    assert (exp != INV).any()


def test_default_constant_p2_flag_everywhere(cuda):
    """using_constant_p2 = false with the benchmark's other options: handle,
    device-resident handle and batch all give the oracle's maps."""
    pair = scenes.cone_like(2, H=120, W=200, D=64)
    H, W = pair.left.shape
    opt = CuSGMOption.benchmark(using_constant_p2=False)
    exp_l, exp_r = oracle.match(pair.left, pair.right, 0, 64)
    st = StereoCuda()
    assert st.Init(W, H, 0, 64, opt), st.last_error
    dl, dr = np.empty((H, W), np.float32), np.empty((H, W), np.float32)
    assert st.Match(pair.left, pair.right, dl, dr), st.last_error
    _eq(dl, exp_l, "handle left")
    _eq(dr, exp_r, "handle right")
    L = torch.from_numpy(pair.left).to(cuda)
    R = torch.from_numpy(pair.right).to(cuda)
    out = torch.empty((H, W), dtype=torch.float32, device=cuda)
    b = SgmBatch([dict(left=L, right=R, disp_left=out, disp_range=64)], opt)
    b.run()
    torch.cuda.synchronize()
    _eq(out.cpu().numpy(), exp_l, "batch left")
