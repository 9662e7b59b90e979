"""Batched matching (sgm_batch_*): every pair of an uneven batch must be
byte-identical to the plain path (StereoCuda.Match) on the same pair, and
to the CPU oracle."""
import numpy as np
import pytest
import torch

import oracle
from fasterstereocuda_amd import _lib, scenes
from fasterstereocuda_amd.batch import SgmBatch
from fasterstereocuda_amd.stereo import CuSGMOption, StereoCuda

pytestmark = pytest.mark.gpu

# (W, H, D, min_d): uneven sizes, both D of the uneven workload, ragged edges
SPECS = [(450, 375, 64, 0), (160, 90, 128, 0), (333, 101, 64, 5), (97, 61, 128, -3),
         (640, 48, 64, 0), (121, 77, 256, 0)]


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("lr", [True, False])
def test_uneven_batch_matches_plain_and_oracle(cuda, lr):
    opt = CuSGMOption.benchmark(do_lr_check=lr)
    scenes_ = [scenes.cone_like(i, H=h, W=w, D=d, workload=3) for i, (w, h, d, _) in enumerate(SPECS)]
    pairs = []
    for sc, (w, h, d, m) in zip(scenes_, SPECS):
        pairs.append({"left": torch.from_numpy(sc.left).to(cuda),
                      "right": torch.from_numpy(sc.right).to(cuda),
                      "disp_left": torch.full((h, w), 7.0, device=cuda),
                      "disp_right": torch.full((h, w), 7.0, device=cuda),
                      "min_disparity": m, "disp_range": d})
    print(f"[batch] {len(pairs)} uneven pairs lr={lr}", flush=True)
    b = SgmBatch(pairs, opt)
    for _ in range(2):                       # reruns give the same bytes
        b.run()
        torch.cuda.synchronize()
        for i, (sc, (w, h, d, m)) in enumerate(zip(scenes_, SPECS)):
            got_l = pairs[i]["disp_left"].cpu().numpy()
            got_r = pairs[i]["disp_right"].cpu().numpy()
            st = StereoCuda()
            assert st.Init(w, h, m, d, opt), st.last_error
            pl = np.empty((h, w), np.float32)
            pr = np.empty((h, w), np.float32)
            assert st.Match(sc.left, sc.right, pl, pr), st.last_error
            np.testing.assert_array_equal(_bits(got_l), _bits(pl), err_msg=f"pair {i} left vs plain")
            np.testing.assert_array_equal(_bits(got_r), _bits(pr), err_msg=f"pair {i} right vs plain")
            el, er = oracle.match(sc.left, sc.right, m, d, do_lr_check=lr)
            np.testing.assert_array_equal(_bits(got_l), _bits(el), err_msg=f"pair {i} left vs oracle")
            np.testing.assert_array_equal(_bits(got_r), _bits(er), err_msg=f"pair {i} right vs oracle")


@pytest.mark.parametrize("mode", [2, 1], ids=["on-the-fly", "stored"])
@pytest.mark.parametrize("specs", [
    [(12, 8, 64, 0), (200, 40, 512, 0), (97, 33, 128, 2)],
    [(12, 8, 64, 0), (20, 12, 256, 3), (150, 50, 64, -4)],
], ids=["with-d512", "tiny"])
def test_extreme_batches_match_oracle(cuda, specs, mode):
    """Batches mixing a tiny image, D = 512 (no on-the-fly kernel: the batch
    keeps the stored cost) and +-min_d, pair by pair against the oracle."""
    lib = _lib.lib()
    before = lib.sgm_get_cost_mode()
    assert lib.sgm_set_cost_mode(mode) == 0
    try:
        opt = CuSGMOption.benchmark()
        scs = [scenes.cone_like(20 + i, H=h, W=w, D=d) for i, (w, h, d, _) in enumerate(specs)]
        pairs = [{"left": torch.from_numpy(sc.left).to(cuda), "right": torch.from_numpy(sc.right).to(cuda),
                  "disp_left": torch.full((h, w), 7.0, device=cuda),
                  "disp_right": torch.full((h, w), 7.0, device=cuda),
                  "min_disparity": m, "disp_range": d}
                 for sc, (w, h, d, m) in zip(scs, specs)]
        print(f"[batch] extreme {specs} mode={mode}", flush=True)
        b = SgmBatch(pairs, opt)
        b.run()
        torch.cuda.synchronize()
    finally:
        assert lib.sgm_set_cost_mode(before) == 0
    for i, (sc, (w, h, d, m)) in enumerate(zip(scs, specs)):
        el, er = oracle.match(sc.left, sc.right, m, d)
        np.testing.assert_array_equal(_bits(pairs[i]["disp_left"].cpu().numpy()), _bits(el),
                                      err_msg=f"pair {i} left vs oracle")
        np.testing.assert_array_equal(_bits(pairs[i]["disp_right"].cpu().numpy()), _bits(er),
                                      err_msg=f"pair {i} right vs oracle")


def test_driving_batch_full_size(cuda):
    """The benchmark's own configuration: 32 driving pairs of 1242x375,
    D = 64, LR check and median, batched with the default cost mode (the
    cost computed on the fly at this size), every map against the oracle."""
    pairs_np = scenes.driving_batch(32)
    opt = CuSGMOption.benchmark()
    pairs = [{"left": torch.from_numpy(p.left).to(cuda), "right": torch.from_numpy(p.right).to(cuda),
              "disp_left": torch.full((375, 1242), 7.0, device=cuda),
              "disp_right": torch.full((375, 1242), 7.0, device=cuda),
              "min_disparity": 0, "disp_range": 64} for p in pairs_np]
    print("[batch] 32 driving pairs 1242x375", flush=True)
    b = SgmBatch(pairs, opt)
    b.run()
    torch.cuda.synchronize()
    for i, p in enumerate(pairs_np):
        el, er = oracle.match(p.left, p.right, 0, 64)
        np.testing.assert_array_equal(_bits(pairs[i]["disp_left"].cpu().numpy()), _bits(el),
                                      err_msg=f"pair {i} left vs oracle")
        np.testing.assert_array_equal(_bits(pairs[i]["disp_right"].cpu().numpy()), _bits(er),
                                      err_msg=f"pair {i} right vs oracle")


def test_batch_stage_timing(cuda):
    sc = scenes.driving(0)
    h, w = sc.left.shape
    pairs = [{"left": torch.from_numpy(sc.left).to(cuda), "right": torch.from_numpy(sc.right).to(cuda),
              "disp_left": torch.empty((h, w), device=cuda), "disp_range": 64}]
    b = SgmBatch(pairs, CuSGMOption.benchmark())
    b.set_timing(True)
    for _ in range(3):
        b.run()
    ms, runs = b.stage_times()
    assert runs == 3 and all(v > 0 for v in ms.values())
    ms, runs = b.stage_times()
    assert runs == 0
    assert b.device_bytes > 10 * w * h * 64


def test_batch_rejects_bad_input(cuda):
    t = torch.zeros((10, 10), dtype=torch.uint8, device=cuda)
    o = torch.zeros((10, 10), dtype=torch.float32, device=cuda)
    with pytest.raises(RuntimeError, match="bad size or range"):
        SgmBatch([{"left": t, "right": t, "disp_left": o, "disp_range": 96}], CuSGMOption.benchmark())
    with pytest.raises(ValueError):
        SgmBatch([{"left": t.float(), "right": t, "disp_left": o, "disp_range": 64}],
                 CuSGMOption.benchmark())


def _uneven_pairs(cuda, lr):
    scenes_ = [scenes.cone_like(i, H=h, W=w, D=d, workload=3) for i, (w, h, d, _) in enumerate(SPECS)]
    pairs = []
    for sc, (w, h, d, m) in zip(scenes_, SPECS):
        pairs.append({"left": torch.from_numpy(sc.left).to(cuda),
                      "right": torch.from_numpy(sc.right).to(cuda),
                      "disp_left": torch.full((h, w), 7.0, device=cuda),
                      "disp_right": torch.full((h, w), 7.0, device=cuda) if lr else None,
                      "min_disparity": m, "disp_range": d})
    return pairs


def _outputs(pairs):
    torch.cuda.synchronize()
    out = []
    for p in pairs:
        out.append(_bits(p["disp_left"].cpu().numpy()))
        if p["disp_right"] is not None:
            out.append(_bits(p["disp_right"].cpu().numpy()))
    return out


def _clobber(pairs):
    for p in pairs:
        p["disp_left"].fill_(7.0)
        if p["disp_right"] is not None:
            p["disp_right"].fill_(7.0)


@pytest.fixture
def cost_mode():
    from fasterstereocuda_amd import stereo
    before = stereo.get_cost_mode()
    yield stereo.set_cost_mode
    stereo.set_cost_mode(before)


@pytest.mark.parametrize("mode", [2, 1], ids=["on-the-fly", "stored"])
@pytest.mark.parametrize("lr", [True, False])
def test_schedules_give_identical_bytes(cuda, cost_mode, mode, lr):
    """Grid, job-major and longest-first schedules, with full and tiny
    persistent grids (each wave then drains many queue entries), produce the
    bytes of the grid schedule, run after run (the queue counters restart),
    for both the on-the-fly and the stored-cost kernels; and match the
    oracle."""
    from fasterstereocuda_amd.batch import SCHED_GRID, SCHED_JOB_MAJOR, SCHED_LONGEST_FIRST
    cost_mode(mode)
    pairs = _uneven_pairs(cuda, lr)
    b = SgmBatch(pairs, CuSGMOption.benchmark(do_lr_check=lr))
    b.set_schedule(SCHED_GRID)
    b.run()
    ref = _outputs(pairs)
    k = 0
    for i, (w, h, d, m) in enumerate(SPECS):
        sc = scenes.cone_like(i, H=h, W=w, D=d, workload=3)
        el, er = oracle.match(sc.left, sc.right, m, d, do_lr_check=lr)
        np.testing.assert_array_equal(ref[k], _bits(el), err_msg=f"pair {i} left vs oracle")
        k += 1
        if lr:
            np.testing.assert_array_equal(ref[k], _bits(er), err_msg=f"pair {i} right vs oracle")
            k += 1
    for policy in (SCHED_JOB_MAJOR, SCHED_LONGEST_FIRST):
        for waves in (0, 7, 3000):
            print(f"[schedule] lr={lr} policy={policy} waves={waves}", flush=True)
            b.set_schedule(policy, waves)
            got_policy, got_waves = b.schedule
            assert got_policy == policy and got_waves > 0
            for _ in range(2):
                _clobber(pairs)
                b.run()
                for k, (g, r) in enumerate(zip(_outputs(pairs), ref)):
                    np.testing.assert_array_equal(g, r, err_msg=f"policy {policy} waves {waves} map {k}")
    with pytest.raises(RuntimeError):
        b.set_schedule(9)


@pytest.mark.parametrize("lr", [True, False])
def test_wta_forms_give_identical_bytes(cuda, cost_mode, lr):
    """Stored-cost batches on the grid schedule with bottom->top fused into
    the WTA and with the four directions followed by the standalone WTA give
    the same bytes, eagerly and captured in a graph."""
    from fasterstereocuda_amd import stages
    from fasterstereocuda_amd.batch import SCHED_GRID
    cost_mode(1)
    before = stages.get_wta_form()
    try:
        pairs = _uneven_pairs(cuda, lr)
        b = SgmBatch(pairs, CuSGMOption.benchmark(do_lr_check=lr))
        b.set_schedule(SCHED_GRID)
        stages.set_wta_form(stages.WTA_FUSED)
        b.run()
        ref = _outputs(pairs)
        stages.set_wta_form(stages.WTA_SEPARATE)
        for graph in (False, True):
            print(f"[wta form] lr={lr} separate, graph={graph}", flush=True)
            b.use_graph(graph)
            _clobber(pairs)
            b.run()
            for k, (g, r) in enumerate(zip(_outputs(pairs), ref)):
                np.testing.assert_array_equal(g, r, err_msg=f"graph={graph} map {k}")
    finally:
        stages.set_wta_form(before)


@pytest.mark.parametrize("mode", [2, 1], ids=["on-the-fly", "stored"])
@pytest.mark.parametrize("policy", [0, 2])
def test_graph_replays_match_eager(cuda, cost_mode, mode, policy):
    """The captured pipeline (kernel nodes only; the queue counters are reset
    by a kernel) replayed several times, with eager torch work and
    synchronisation between replays, gives the eager bytes every time."""
    cost_mode(mode)
    pairs = _uneven_pairs(cuda, True)
    b = SgmBatch(pairs, CuSGMOption.benchmark())
    b.set_schedule(policy)
    b.run()
    ref = _outputs(pairs)
    b.use_graph(True)
    with pytest.raises(RuntimeError, match="graph mode"):
        b.set_timing(True)
    scratch = torch.randn(1 << 20, device=cuda)
    for rep in range(4):
        print(f"[graph] policy={policy} replay {rep}", flush=True)
        _clobber(pairs)
        scratch.mul_(0.5).add_(1.0)           # eager work between replays
        torch.cuda.synchronize()
        b.run()
        for k, (g, r) in enumerate(zip(_outputs(pairs), ref)):
            np.testing.assert_array_equal(g, r, err_msg=f"replay {rep} map {k}")
    b.use_graph(False)
    _clobber(pairs)
    b.run()
    for g, r in zip(_outputs(pairs), ref):
        np.testing.assert_array_equal(g, r)


def test_autotune_picks_and_remembers(cuda):
    pairs = _uneven_pairs(cuda, True)
    b = SgmBatch(pairs, CuSGMOption.benchmark())
    b.run()
    ref = _outputs(pairs)
    policy, waves, ms = b.autotune(reps=2, force=True)
    print(f"[autotune] policy={policy} waves={waves} ms={ms:.3f}", flush=True)
    assert policy in (0, 1, 2) and ms > 0
    assert b.schedule == (policy, waves)
    _clobber(pairs)
    b.run()
    for g, r in zip(_outputs(pairs), ref):
        np.testing.assert_array_equal(g, r)
    # a second batch of the same signature starts on the remembered schedule
    pairs2 = _uneven_pairs(cuda, True)
    b2 = SgmBatch(pairs2, CuSGMOption.benchmark())
    assert b2.schedule == (policy, waves)
    assert b2.autotune(reps=2)[2] == -1.0     # cached: applied without timing
    b2.run()
    for g, r in zip(_outputs(pairs2), ref):
        np.testing.assert_array_equal(g, r)
