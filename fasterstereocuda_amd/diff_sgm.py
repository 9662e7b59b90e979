"""PyTorch-ROCm binding of the differentiable FP32 SGM.

    disp = diff_sgm(cost, p1, p2, temperature=0.1)      # cost [B,H,W,D], p1/p2 [B,H,W]
    loss = (disp - gt).abs().mean(); loss.backward()    # grads for cost, p1, p2

The forward and the hand-written backward are HIP kernels in libfsgm.so
(sgm_dsgm_forward / sgm_dsgm_backward); this module only allocates their
buffers with torch's caching allocator and passes the current stream.  The
forward saves S (the 4-direction sum) and, per direction, the winning term of
every cell (2 bits) and each pixel's argmin byte, about a quarter of one
volume; no per-direction volume is kept.  Each call's transient workspace is
one direction volume plus two per-pixel planes.  The torch-free C++ API never
depends on this module.
"""
from __future__ import annotations

import torch

from ._lib import check, lib
from . import stages


def _ptr(t):
    return t.data_ptr()


def _check_inputs(cost, p1, p2):
    if cost.dim() != 4:
        raise ValueError("cost must be [B, H, W, D]")
    B, H, W, D = cost.shape
    for name, t, shape in (("cost", cost, (B, H, W, D)), ("p1", p1, (B, H, W)), ("p2", p2, (B, H, W))):
        if not t.is_cuda or t.dtype != torch.float32 or tuple(t.shape) != shape:
            raise ValueError(f"{name} must be a float32 CUDA tensor of shape {shape}")
    if D not in (64, 128, 256):
        raise ValueError("disparity range D must be 64, 128 or 256")
    return B, H, W, D


class DiffSGMFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cost, p1, p2, temperature: float):
        cost, p1, p2 = cost.contiguous(), p1.contiguous(), p2.contiguous()
        B, H, W, D = _check_inputs(cost, p1, p2)
        dev = cost.device
        L = lib()
        S = torch.empty_like(cost)
        disp = torch.empty((B, H, W), dtype=torch.float32, device=dev)
        ckpt = torch.empty(L.sgm_dsgm_checkpoint_floats(B, H, W, D), dtype=torch.float32, device=dev)
        work = torch.empty(L.sgm_dsgm_workspace_floats(B, H, W, D), dtype=torch.float32, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        check(L.sgm_dsgm_forward(_ptr(cost), _ptr(p1), _ptr(p2), B, H, W, D, float(temperature),
                                 _ptr(S), _ptr(disp), _ptr(ckpt), _ptr(work), stream))
        del work
        ctx.save_for_backward(cost, p1, p2, S, disp, ckpt)
        ctx.temperature = float(temperature)
        return disp

    @staticmethod
    def backward(ctx, grad_disp):
        cost, p1, p2, S, disp, ckpt = ctx.saved_tensors
        B, H, W, D = cost.shape
        grad_disp = grad_disp.contiguous().to(torch.float32)
        L = lib()
        g_cost = torch.empty_like(cost)
        g_p1 = torch.empty_like(p1)
        g_p2 = torch.empty_like(p2)
        work = torch.empty(L.sgm_dsgm_workspace_floats(B, H, W, D), dtype=torch.float32,
                           device=cost.device)
        stream = torch.cuda.current_stream(cost.device).cuda_stream
        check(L.sgm_dsgm_backward(_ptr(cost), _ptr(p1), _ptr(p2), _ptr(S), _ptr(disp), _ptr(ckpt),
                                  _ptr(grad_disp), B, H, W, D, ctx.temperature, _ptr(g_cost),
                                  _ptr(g_p1), _ptr(g_p2), _ptr(work), stream))
        return g_cost, g_p1, g_p2, None


def diff_sgm(cost: torch.Tensor, p1: torch.Tensor, p2: torch.Tensor,
             temperature: float = 0.1) -> torch.Tensor:
    """Soft-argmin disparity [B, H, W] of the differentiable 4-direction SGM."""
    return DiffSGMFunction.apply(cost, p1, p2, temperature)


def census_cost_volume(left: torch.Tensor, right: torch.Tensor, D: int) -> torch.Tensor:
    """FP32 normalised Hamming cost (9x7 census, cost / 63) of uint8 pairs
    [B, H, W] -> [B, H, W, D], computed by the census / cost HIP kernels."""
    # costs are integers 0..63: map them through a table of correctly rounded
    # quotients so the volume is exactly cost / 63 in float32
    import numpy as np
    table = torch.from_numpy((np.arange(256, dtype=np.float32) / np.float32(63.0))).to(left.device)
    out = []
    for b in range(left.shape[0]):
        cl = stages.census(left[b].contiguous(), 1, False)
        cr = stages.census(right[b].contiguous(), 1, True)
        out.append(table[stages.cost(cl, cr, 0, D, True, 1).long()])
    return torch.stack(out)
