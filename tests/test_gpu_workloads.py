"""The large benchmark workloads run at their real sizes on the GPU and
checked against the oracle:
  - the uneven batch: 64 pairs of five formats, D in {64, 128}, autotuned and
    run as a HIP graph, every map of every pair bit for bit;
  - the differentiable SGM at batch 8 x 1242 x 375 x 128: four images of a
    cost volume whose aggregation is exact in FP32 against the FP64
    reference within 1e-4, and the census-cost workload itself (finite,
    repeatable);
  - a dense single level at D = 256 whose per-direction volumes exceed 2^31
    bytes, the case the 64-bit volume offsets exist for, in both cost modes."""
import numpy as np
import pytest
import torch

import oracle
from fasterstereocuda_amd import scenes
from fasterstereocuda_amd.batch import SgmBatch
from fasterstereocuda_amd.diff_sgm import census_cost_volume, diff_sgm
from fasterstereocuda_amd.stereo import CuSGMOption, StereoCuda

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _eq(a, b, what):
    a, b = _bits(a), _bits(b)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} of {a.size} differ, first at {tuple(bad[0])}"


def test_uneven_batch_workload(cuda):
    specs = scenes.uneven_specs(64)
    print(f"[uneven workload] 64 pairs, {sum(w * h * d for w, h, d in specs) / 1e9:.2f} G cells",
          flush=True)
    pairs_np = scenes.uneven_batch(64)
    opt = CuSGMOption.benchmark()
    pairs = [{"left": torch.from_numpy(p.left).to(cuda), "right": torch.from_numpy(p.right).to(cuda),
              "disp_left": torch.full(p.left.shape, 7.0, device=cuda),
              "disp_right": torch.full(p.left.shape, 7.0, device=cuda),
              "min_disparity": 0, "disp_range": d} for p, (_, _, d) in zip(pairs_np, specs)]
    b = SgmBatch(pairs, opt)
    policy, waves, ms = b.autotune(reps=2, force=True)
    print(f"[uneven workload] autotuned policy={policy} waves={waves} {ms:.2f} ms", flush=True)
    b.use_graph(True)
    for p in pairs:
        p["disp_left"].fill_(7.0)
        p["disp_right"].fill_(7.0)
    b.run()
    torch.cuda.synchronize()
    for i, (p, (w, h, d)) in enumerate(zip(pairs_np, specs)):
        el, er = oracle.match(p.left, p.right, 0, d)
        _eq(pairs[i]["disp_left"].cpu().numpy(), el, f"pair {i} ({w}x{h} D{d}) left")
        _eq(pairs[i]["disp_right"].cpu().numpy(), er, f"pair {i} ({w}x{h} D{d}) right")
        if i % 16 == 15:
            print(f"[uneven workload] {i + 1} pairs checked", flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _worst_line(a, b, axis):
    """Largest relative error over the rows (axis 0) or columns (axis 1)."""
    d = np.moveaxis(a - b, axis, 0).reshape(a.shape[axis], -1)
    r = np.moveaxis(b, axis, 0).reshape(b.shape[axis], -1)
    return float((np.linalg.norm(d, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-30)).max())


def test_diffsgm_full_size_against_fp64(cuda):
    """Batch 8 x 375 x 1242 x 128 at the workload's temperature 0.1 with an
    L1 loss; images 0, 3, 5 and 7 against the FP64 reference, gradient
    relative error (norm of the difference over the norm of the reference)
    <= 1e-5, and every row and column within 2e-5 (measured: about 8e-7 and
    at most 1.4e-6).

    Costs are multiples of 1/64 in [0, 1) and P1 / P2 multiples of 1/64 in
    the workload's ranges (P1 in [0.0625, 0.1875], P2 in [0.5, 2.0]): every
    aggregated value is then exact in FP32 as in FP64, so both take the same
    minimum branch at every step, ties included, and the comparison measures
    the gradients themselves.  With arbitrary float costs a few hundred of
    an image's ~2e8 branch decisions fall within FP32 rounding of a tie and
    FP32 routes those elements' gradients differently from FP64 (seen as
    isolated rows and columns off by ~1e-2 in some images)."""
    B, H, W, D, T = 8, 375, 1242, 128, 0.1
    rng = np.random.default_rng(4)
    cost = (rng.integers(0, 64, (B, H, W, D)) / 64.0).astype(np.float32)
    p1 = (rng.integers(4, 13, (B, H, W)) / 64.0).astype(np.float32)
    p2 = (rng.integers(32, 129, (B, H, W)) / 64.0).astype(np.float32)
    gt = rng.uniform(0, D - 1, (B, H, W)).astype(np.float32)
    print(f"[dsgm full size] B{B} {W}x{H} D{D}", flush=True)
    tc = torch.from_numpy(cost).to(cuda).requires_grad_(True)
    t1 = torch.from_numpy(p1).to(cuda).requires_grad_(True)
    t2 = torch.from_numpy(p2).to(cuda).requires_grad_(True)
    disp = diff_sgm(tc, t1, t2, T)
    loss = (disp - torch.from_numpy(gt).to(cuda)).abs().mean(dim=(1, 2)).sum()
    loss.backward()
    torch.cuda.synchronize()
    for b in (0, 3, 5, B - 1):
        print(f"[dsgm full size] FP64 reference of image {b}", flush=True)
        _, d_ref, _ = oracle.dsgm_forward(cost[b], p1[b], p2[b], gt[b], T)
        gC, g1, g2, _ = oracle.dsgm_backward(cost[b], p1[b], p2[b], gt[b], T)
        d = disp[b].detach().cpu().numpy()
        assert np.abs(d - d_ref).max() <= 1e-5 * max(1.0, np.abs(d_ref).max()), "disparity"
        for name, got, ref in (("dC", tc.grad[b], gC), ("dP1", t1.grad[b], g1), ("dP2", t2.grad[b], g2)):
            g = got.cpu().numpy().astype(np.float64)
            rel, rows, cols = _rel(g, ref), _worst_line(g, ref, 0), _worst_line(g, ref, 1)
            print(f"[dsgm full size] image {b} {name}: rel {rel:.2e}, worst row {rows:.2e}, "
                  f"worst column {cols:.2e}, max abs {np.abs(g - ref).max():.2e} of "
                  f"{np.abs(ref).max():.2e}", flush=True)
            assert rel <= 1e-5, f"image {b} {name} rel err {rel:.2e}"
            assert rows <= 2e-5 and cols <= 2e-5, f"image {b} {name} rows {rows:.2e} cols {cols:.2e}"


def test_diffsgm_census_workload_runs_and_repeats(cuda):
    """The workload as the benchmark runs it: census cost / 63 of 8 driving
    pairs at D = 128; forward and backward are finite and bit-repeatable."""
    B, D = 8, 128
    pairs = scenes.driving_batch(B, D=D, workload=4)
    L = torch.from_numpy(np.stack([p.left for p in pairs])).to(cuda)
    R = torch.from_numpy(np.stack([p.right for p in pairs])).to(cuda)
    gt = torch.from_numpy(np.stack([p.disparity for p in pairs])).to(cuda)
    cost = census_cost_volume(L, R, D)
    rng = np.random.default_rng(7)
    p1 = torch.from_numpy(rng.uniform(0.05, 0.2, gt.shape).astype(np.float32)).to(cuda)
    p2 = torch.from_numpy(rng.uniform(0.5, 2.0, gt.shape).astype(np.float32)).to(cuda)
    print("[dsgm census workload] B8 1242x375 D128", flush=True)
    outs = []
    for _ in range(2):
        c = cost.clone().requires_grad_(True)
        a, b = p1.clone().requires_grad_(True), p2.clone().requires_grad_(True)
        disp = diff_sgm(c, a, b, 0.1)
        (disp - gt).abs().mean().backward()
        torch.cuda.synchronize()
        outs.append([x.detach().cpu().numpy() for x in (disp, c.grad, a.grad, b.grad)])
    for x in outs[0]:
        assert np.isfinite(x).all()
    for x, y in zip(*outs):
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))
    # the soft-argmin disparity tracks the scene (ground truth error well under the range)
    err = np.abs(outs[0][0] - gt.cpu().numpy())
    assert np.median(err) < 8.0


@pytest.fixture
def cost_mode():
    from fasterstereocuda_amd import stereo
    before = stereo.get_cost_mode()
    yield stereo.set_cost_mode
    stereo.set_cost_mode(before)


@pytest.mark.parametrize("mode", [1, 2], ids=["stored", "on-the-fly"])
def test_dense_d256_beyond_2gib(cuda, cost_mode, mode):
    """4800 x 1800 at D = 256: each direction volume holds 2.21e9 bytes
    (> 2^31), so every offset past 2 GiB must be 64-bit."""
    W, H, D = 4800, 1800, 256
    assert W * H * D > 2 ** 31
    cost_mode(mode)
    pair = scenes.building(0, H=H, W=W, D=D)
    print(f"[dense D256] {W}x{H} mode={mode}", flush=True)
    st = StereoCuda()
    assert st.Init(W, H, 0, D, CuSGMOption.benchmark()), st.last_error
    dl = np.empty((H, W), np.float32)
    dr = np.empty((H, W), np.float32)
    assert st.Match(pair.left, pair.right, dl, dr), st.last_error
    st.Release()
    print("[dense D256] oracle", flush=True)
    el, er = oracle.match(pair.left, pair.right, 0, D)
    _eq(dl, el, "left map")
    _eq(dr, er, "right map")
    # the rows past 2^31 / (W * D) = 1747 are the ones a 32-bit offset would corrupt
    assert (el[1750:1790] != oracle.INVALID).mean() > 0.3
