"""Batched matching (sgm_batch_*): every pair of an uneven batch must be
byte-identical to the plain path (StereoCuda.Match) on the same pair, and
to the CPU oracle."""
import numpy as np
import pytest
import torch

import oracle
from fasterstereocuda_amd import scenes
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
