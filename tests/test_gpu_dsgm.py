"""Differentiable FP32 SGM on the GPU against the FP64 CPU reference
(oracle/dsgm_oracle.cpp, itself pinned by finite differences in
tests/test_dsgm_oracle.py).  Tolerances: disparity and S within 1e-5
relative, gradients within 1e-4 relative (norm-wise and element-wise against
the largest entry)."""
import numpy as np
import pytest
import torch

import oracle
from fasterstereocuda_amd.diff_sgm import census_cost_volume, diff_sgm

pytestmark = pytest.mark.gpu

T = 0.1


def _inputs(B, H, W, D, seed):
    rng = np.random.default_rng(seed)
    cost = rng.random((B, H, W, D)).astype(np.float32)
    p1 = rng.uniform(0.05, 0.2, (B, H, W)).astype(np.float32)
    p2 = rng.uniform(0.5, 2.0, (B, H, W)).astype(np.float32)
    gt = rng.uniform(0, D - 1, (B, H, W)).astype(np.float32)
    return cost, p1, p2, gt


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


# even and odd widths and heights, lines of fewer steps than the load
# pipeline's depth (rows of 1, 2 and 3 pixels), and D = 64, 128, 256
@pytest.mark.parametrize("B,H,W,D", [(2, 33, 40, 64), (1, 32, 97, 128), (1, 20, 64, 256),
                                     (2, 45, 70, 128), (1, 7, 1, 64), (2, 5, 2, 64),
                                     (1, 9, 3, 128)])
def test_forward_backward_match_fp64(cuda, B, H, W, D):
    print(f"[dsgm] B{B} H{H} W{W} D{D}", flush=True)
    cost, p1, p2, gt = _inputs(B, H, W, D, seed=H + W + D)
    tc = torch.from_numpy(cost).to(cuda).requires_grad_(True)
    t1 = torch.from_numpy(p1).to(cuda).requires_grad_(True)
    t2 = torch.from_numpy(p2).to(cuda).requires_grad_(True)
    tg = torch.from_numpy(gt).to(cuda)
    disp = diff_sgm(tc, t1, t2, T)
    loss = (disp - tg).abs().mean(dim=(1, 2)).sum()     # per-image mean L1
    loss.backward()
    for b in range(B):
        S, d_ref, l_ref = oracle.dsgm_forward(cost[b], p1[b], p2[b], gt[b], T)
        gC, g1, g2, _ = oracle.dsgm_backward(cost[b], p1[b], p2[b], gt[b], T)
        d = disp[b].detach().cpu().numpy()
        assert np.abs(d - d_ref).max() <= 1e-5 * max(1.0, np.abs(d_ref).max()), "disparity"
        for name, got, ref in (("dC", tc.grad[b], gC), ("dP1", t1.grad[b], g1), ("dP2", t2.grad[b], g2)):
            g = got.cpu().numpy().astype(np.float64)
            assert _rel(g, ref) <= 1e-4, f"{name} rel err {_rel(g, ref):.2e}"
            assert np.abs(g - ref).max() <= 1e-4 * np.abs(ref).max(), f"{name} max err"


def test_deterministic_and_repeatable(cuda):
    cost, p1, p2, gt = _inputs(2, 40, 50, 64, seed=1)
    outs = []
    for _ in range(2):
        tc = torch.from_numpy(cost).to(cuda).requires_grad_(True)
        t1 = torch.from_numpy(p1).to(cuda).requires_grad_(True)
        t2 = torch.from_numpy(p2).to(cuda).requires_grad_(True)
        disp = diff_sgm(tc, t1, t2, T)
        (disp - torch.from_numpy(gt).to(cuda)).abs().mean().backward()
        outs.append([x.detach().cpu().numpy() for x in (disp, tc.grad, t1.grad, t2.grad)])
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_census_cost_volume(cuda):
    rng = np.random.default_rng(0)
    left = rng.integers(0, 256, (2, 30, 80), dtype=np.uint8)
    right = rng.integers(0, 256, (2, 30, 80), dtype=np.uint8)
    c = census_cost_volume(torch.from_numpy(left).to(cuda), torch.from_numpy(right).to(cuda), 64)
    for b in range(2):
        cl = oracle.census(left[b], 1, False)
        cr = oracle.census(right[b], 1, True)
        ref = oracle.cost(cl, cr, 0, 64, True, 1).astype(np.float32) / np.float32(63.0)
        np.testing.assert_array_equal(c[b].cpu().numpy(), ref)


def test_bad_shapes_raise(cuda):
    x = torch.zeros((1, 4, 4, 96), device=cuda)
    p = torch.zeros((1, 4, 4), device=cuda)
    with pytest.raises(ValueError):
        diff_sgm(x, p, p)
