"""Coarse-to-fine pyramid on the GPU against the CPU oracle: every window
stage bit-exact, then whole pyramids (window 16 and 32) including the
4800x3409, D = 256, 3-level workload."""
import numpy as np
import pytest
import torch

import oracle
from fasterstereocuda_amd import scenes, stages
from fasterstereocuda_amd.pyramid import FasterStereoCuda, SgmPyramid, StereoOption1
from fasterstereocuda_amd.stereo import CuSGMOption

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.fixture(params=[stages.WINDOW_AUTO, stages.WINDOW_U32, stages.WINDOW_PAIRS, stages.WINDOW_QUADS],
                ids=["auto", "u32", "pairs", "quads"])
def window_form(request):
    """Runs a test under each form of the window aggregation (one disparity
    per lane in u32, two or four per lane as packed u16, the automatic
    choice); all must give the same bytes."""
    before = stages.get_window_form()
    stages.set_window_form(request.param)
    yield request.param
    stages.set_window_form(before)


def _random_init(rng, h, w, lr, invalid_frac=0.15):
    init = rng.integers(-3, 50, size=(h, w)).astype(np.int16)
    # piecewise-constant runs with jumps, like propagated disparities
    init = np.repeat(init[:, ::4], 4, axis=1)[:, :w].copy()
    init[rng.random((h, w)) < invalid_frac] = oracle.INVALID_SHORT
    return init


@pytest.mark.parametrize("lr", [16, 32])
@pytest.mark.parametrize("cs", [1, 0])
def test_window_stages_match_oracle(cuda, window_form, lr, cs):
    print(f"[window stages] lr={lr} cs={cs} form={window_form}", flush=True)
    rng = np.random.default_rng(lr + cs)
    pair = scenes.cone_like(2, H=57, W=131, D=64)
    h, w = pair.left.shape
    init = _random_init(rng, h, w, lr)
    cl = oracle.census(pair.left, cs, False)
    cr = oracle.census(pair.right, cs, True)
    ref_cost = oracle.cost_init(cl, cr, init, lr, cs)
    g_init = torch.from_numpy(init).to(cuda)
    g_cost = stages.cost_init(torch.from_numpy(cl.view(np.int64)).to(cuda),
                              torch.from_numpy(cr.view(np.int64)).to(cuda), g_init, lr, cs)
    np.testing.assert_array_equal(g_cost.cpu().numpy(), ref_cost)
    p1, p2 = (10, 150) if cs else (8, 32)
    paths = stages.aggregate_init(g_cost, g_init, p1, p2)
    ref_paths = [oracle.aggregate_init(ref_cost, init, r, p1, p2) for r in range(4)]
    for r in range(4):
        np.testing.assert_array_equal(paths[r].cpu().numpy(), ref_paths[r], err_msg=f"direction {r}")
    for uniq in (0.95, 1.0):
        d = stages.wta_init(paths, g_init, uniq)
        np.testing.assert_array_equal(_bits(d.cpu().numpy()), _bits(oracle.wta_init(ref_paths, init, uniq)))


@pytest.mark.parametrize("lr", [16, 32])
@pytest.mark.parametrize("p1,p2", [(10, 254), (20, 400), (0, 60000)])
def test_window_aggregation_penalty_forms(cuda, window_form, lr, p1, p2):
    """P2 / max(1, |b' - b + 1|) comes from a table up to P2 = 254 and from a
    division above; both against the oracle, with base jumps of every size."""
    print(f"[window aggregation] lr={lr} p1={p1} p2={p2} form={window_form}", flush=True)
    rng = np.random.default_rng(lr * 7 + p2)
    h, w = 41, 203
    init = rng.integers(-40, 300, size=(h, w)).astype(np.int16)   # large jumps
    init[:, ::3] = init[:, 1::3][:, : init[:, ::3].shape[1]]          # and runs of equal bases
    init[rng.random((h, w)) < 0.1] = oracle.INVALID_SHORT
    cost = rng.integers(0, 64, size=(h, w, lr), dtype=np.uint8)
    paths = stages.aggregate_init(torch.from_numpy(cost).to(cuda), torch.from_numpy(init).to(cuda),
                                  p1, p2)
    for r in range(4):
        np.testing.assert_array_equal(paths[r].cpu().numpy(),
                                      oracle.aggregate_init(cost, init, r, p1, p2), err_msg=f"dir {r}")


@pytest.mark.parametrize("lr", [16, 32])
@pytest.mark.parametrize("h,w", [(1, 1), (1, 7), (3, 1), (2, 17), (5, 9), (17, 3), (9, 33), (1, 40), (40, 1)])
def test_window_aggregation_tiny_shapes(cuda, window_form, lr, h, w):
    """Scanlines shorter than one block of steps and fewer scanlines than a
    wave holds: the blocks past a line's end and the pixels past the last
    scanline are computed but their stores are dropped; every pixel of the
    four direction volumes must match the oracle."""
    print(f"[window aggregation tiny] lr={lr} {h}x{w} form={window_form}", flush=True)
    rng = np.random.default_rng(h * 100 + w + lr)
    init = (rng.integers(0, 3, size=(h, w)).cumsum(axis=1) - 1).astype(np.int16)   # steps of -1..1
    init[rng.random((h, w)) < 0.15] = oracle.INVALID_SHORT
    cost = rng.integers(0, 64, size=(h, w, lr), dtype=np.uint8)
    paths = stages.aggregate_init(torch.from_numpy(cost).to(cuda), torch.from_numpy(init).to(cuda),
                                  10, 150)
    for r in range(4):
        np.testing.assert_array_equal(paths[r].cpu().numpy(),
                                      oracle.aggregate_init(cost, init, r, 10, 150), err_msg=f"dir {r}")


@pytest.mark.parametrize("lr", [16, 32])
@pytest.mark.parametrize("p1,p2", [(127, 254), (127, 127), (0, 0), (1, 254), (128, 254), (127, 255)])
def test_window_aggregation_packed_limits(cuda, window_form, lr, p1, p2):
    """The packed-u16 window aggregation (P1 <= 127, P2 <= 254) at the edges
    of that range, and the first penalties past it (the u32 kernel): costs up
    to 255 with many at 255, INVALID bases in runs and at the first pixel of
    scanlines, base jumps far beyond the window, each direction bit for bit
    against the oracle."""
    print(f"[window aggregation limits] lr={lr} p1={p1} p2={p2} form={window_form}", flush=True)
    rng = np.random.default_rng(1000 + lr + p1 * 3 + p2)
    h, w = 37, 149
    init = rng.integers(-300, 300, size=(h, w)).astype(np.int16)
    init = np.where(rng.random((h, w)) < 0.5, np.roll(init, 1, axis=1), init).astype(np.int16)
    init[rng.random((h, w)) < 0.12] = oracle.INVALID_SHORT
    init[:, 0] = oracle.INVALID_SHORT            # horizontal scanlines start invalid
    init[0, ::2] = oracle.INVALID_SHORT          # vertical ones too, every other column
    init[5:9, 20:60] = oracle.INVALID_SHORT      # a block of invalid bases
    cost = rng.integers(0, 256, size=(h, w, lr), dtype=np.uint8)
    cost[rng.random((h, w, lr)) < 0.3] = 255
    paths = stages.aggregate_init(torch.from_numpy(cost).to(cuda), torch.from_numpy(init).to(cuda),
                                  p1, p2)
    for r in range(4):
        np.testing.assert_array_equal(paths[r].cpu().numpy(),
                                      oracle.aggregate_init(cost, init, r, p1, p2), err_msg=f"dir {r}")


@pytest.mark.parametrize("lr", [16, 32])
@pytest.mark.parametrize("jump_rate", [0.0, 0.02])
def test_window_aggregation_small_base_steps(cuda, window_form, lr, jump_rate):
    """Bases that move by -1, 0 or +1 from pixel to pixel along both axes (a
    random walk, as propagated disparities of smooth surfaces do), with rare
    large jumps and INVALID runs: steps where every scanline of a wave
    shifts by at most one disparity and steps where one does not, mixed
    within the same scanlines.  Each direction bit for bit against the
    oracle."""
    print(f"[window aggregation small steps] lr={lr} jump_rate={jump_rate} form={window_form}", flush=True)
    rng = np.random.default_rng(77 + lr + int(jump_rate * 100))
    h, w = 53, 211
    walk_x = np.cumsum(rng.integers(-1, 2, size=(1, w)), axis=1)
    walk_y = np.cumsum(rng.integers(-1, 2, size=(h, 1)), axis=0)
    init = (40 + walk_x + walk_y).astype(np.int64)
    init += (rng.random((h, w)) < 0.05).astype(np.int64)        # isolated +1 bumps
    jumps = rng.random((h, w)) < jump_rate
    init[jumps] += rng.integers(-60, 60, size=int(jumps.sum()))
    init = init.astype(np.int16)
    init[rng.random((h, w)) < 0.03] = oracle.INVALID_SHORT
    init[10:12, 50:70] = oracle.INVALID_SHORT
    cost = rng.integers(0, 256, size=(h, w, lr), dtype=np.uint8)
    for p1, p2 in ((10, 150), (127, 254)):
        paths = stages.aggregate_init(torch.from_numpy(cost).to(cuda), torch.from_numpy(init).to(cuda),
                                      p1, p2)
        for r in range(4):
            np.testing.assert_array_equal(paths[r].cpu().numpy(),
                                          oracle.aggregate_init(cost, init, r, p1, p2),
                                          err_msg=f"p1={p1} p2={p2} dir {r}")


@pytest.mark.parametrize("w,h", [(130, 57), (131, 58), (64, 33)])
def test_zoom_and_init_propagation(cuda, w, h):
    rng = np.random.default_rng(w * h)
    img = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
    z = stages.zoom_half(torch.from_numpy(img).to(cuda))
    np.testing.assert_array_equal(z.cpu().numpy(), oracle.zoom_half(img))
    coarse = rng.uniform(0, 30, size=(h // 2, w // 2)).astype(np.float32)
    coarse[rng.random(coarse.shape) < 0.2] = oracle.INVALID
    for lr in (16, 32):
        nxt = stages.init_next(torch.from_numpy(coarse).to(cuda), w, h, lr)
        np.testing.assert_array_equal(nxt.cpu().numpy(), oracle.init_next(coarse, w, h, lr))


@pytest.mark.parametrize("W,H,D,layers,lr,cs", [(400, 300, 128, 3, 16, 1), (400, 300, 128, 3, 32, 1),
                                                (333, 250, 64, 2, 32, 0), (256, 200, 64, 1, 32, 1)])
def test_pyramid_matches_oracle(cuda, W, H, D, layers, lr, cs):
    print(f"[pyramid] {W}x{H} D{D} layers{layers} lr{lr} cs{cs}", flush=True)
    p = scenes.cone_like(5, H=H, W=W, D=D, workload=2)
    opt = CuSGMOption.benchmark(cs_type=cs)
    pyr = SgmPyramid(W, H, 0, D, layers, lr, opt)
    got = pyr.match(p.left, p.right)
    exp = oracle.pyramid(p.left, p.right, 0, D, layers, lr, cs_type=cs)
    np.testing.assert_array_equal(_bits(got), _bits(exp))
    again = pyr.match(p.left, p.right)
    np.testing.assert_array_equal(_bits(again), _bits(got))


@pytest.mark.parametrize("W,H,dmin,dmax,layers,lr,cs", [(517, 389, 12, 160, 4, 16, 1),
                                                        (450, 375, -16, 100, 3, 32, 0)])
def test_pyramid_disparity_interval_matches_oracle(cuda, W, H, dmin, dmax, layers, lr, cs):
    """A disparity interval that does not start at 0, divided down to the
    coarsest of 3 or 4 levels (FasterStereoCuda.cpp:265-281)."""
    print(f"[pyramid] {W}x{H} [{dmin}, {dmax}) layers{layers} lr{lr} cs{cs}", flush=True)
    p = scenes.cone_like(9, H=H, W=W, D=128, workload=2)
    opt = CuSGMOption.benchmark(cs_type=cs)
    pyr = SgmPyramid(W, H, dmin, dmax, layers, lr, opt)
    got = pyr.match(p.left, p.right)
    exp = oracle.pyramid(p.left, p.right, dmin, dmax, layers, lr, cs_type=cs)
    np.testing.assert_array_equal(_bits(got), _bits(exp))
    assert (got != oracle.INVALID).mean() > 0.2


@pytest.mark.parametrize("layers,peaks", [(3, False), (2, True), (1, False), (1, True)])
def test_pyramid_device_entry_binds_caller_buffers(cuda, layers, peaks):
    """match_device reads the caller's images and writes the caller's map in
    place (no staging copies): the same bytes as the host entry point and the
    oracle, including a single level whose map also goes through peak removal,
    and with the caller's buffers reused for a second pair."""
    W, H, D, lr = 320, 240, 128, 16
    print(f"[pyramid] device entry layers{layers} peaks={peaks}", flush=True)
    opt = CuSGMOption.benchmark(do_remove_peaks=peaks)
    pyr = SgmPyramid(W, H, 0, D, layers, lr, opt)
    L = torch.empty((H, W), dtype=torch.uint8, device=cuda)
    R = torch.empty_like(L)
    out = torch.full((H, W), 7.0, dtype=torch.float32, device=cuda)
    for seed in (3, 4):
        p = scenes.cone_like(seed, H=H, W=W, D=D, workload=2)
        L.copy_(torch.from_numpy(p.left))
        R.copy_(torch.from_numpy(p.right))
        pyr.match_device(L, R, out)
        torch.cuda.synchronize()
        host = pyr.match(p.left, p.right)
        np.testing.assert_array_equal(_bits(out.cpu().numpy()), _bits(host))
        if not peaks:
            exp = oracle.pyramid(p.left, p.right, 0, D, layers, lr)
            np.testing.assert_array_equal(_bits(host), _bits(exp))


def test_building_4800x3409_pyramid(cuda):
    """The large-image workload: 4800x3409, D = 256, 3 levels, window 16."""
    p = scenes.building(0)
    print("[pyramid] building 4800x3409", flush=True)
    pyr = SgmPyramid(4800, 3409, 0, 256, 3, 16, CuSGMOption.benchmark())
    assert pyr.coarsest == (0, 64)
    got = pyr.match(p.left, p.right)
    exp = oracle.pyramid(p.left, p.right, 0, 256, 3, 16)
    np.testing.assert_array_equal(_bits(got), _bits(exp))
    valid = got != oracle.INVALID
    assert valid.mean() > 0.5


def test_faster_stereo_cuda_mirror(cuda):
    p = scenes.cone_like(1, H=300, W=400, D=128, workload=2)
    fs = FasterStereoCuda()
    opt = StereoOption1(width=400, height=300, min_disp=0, max_disp=128, num_layers=2,
                        do_lr_check=True, do_rm_peaks=False)
    assert fs.Init(opt), fs.last_error
    out = np.empty((300, 400), np.float32)
    assert fs.Match(p.left, p.right, out)
    exp = oracle.pyramid(p.left, p.right, 0, 128, 2, 32, cs_type=0, p1=8, p2=32)
    np.testing.assert_array_equal(_bits(out), _bits(exp))
    # the defaults (peak removal on) are accepted; smoothing is outside the built scope
    assert FasterStereoCuda().Init(StereoOption1(width=400, height=300, max_disp=128))
    bad = FasterStereoCuda()
    assert not bad.Init(StereoOption1(width=400, height=300, max_disp=128, do_smooth=True))
    assert "post filters" in bad.last_error
