"""The allocation-free peak removal (sgm_stage_remove_peaks_noalloc) against
the workspace form (sgm_stage_remove_peaks), both called through the C ABI:
bit for bit on two calls in a row, the second on a smaller map, so that the
second call runs over what the first left in the fixed scratch; and the
size limit refused before any launch."""
import numpy as np
import pytest
import torch

from fasterstereocuda_amd._lib import SGM_E_INVALID_ARG, SGM_OK, last_error, lib

pytestmark = pytest.mark.gpu

MAX_PIXELS = 1 << 21
INVALID = np.float32(-999999.0)


def _map(seed, H, W):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, 12, size=(H, W)).astype(np.float32) * 0.75
    d = np.repeat(np.repeat(d[::3, ::3], 3, 0), 3, 1)[:H, :W].copy()
    d += rng.choice([0.0, 0.5, 1.25], size=(H, W), p=[0.8, 0.15, 0.05]).astype(np.float32)
    d[rng.random((H, W)) < 0.1] = INVALID
    return d


def _with_workspace(d, strips, ratio, stream):
    H, W = d.shape
    out = d.clone()
    work = torch.empty(2 * H * W, dtype=torch.int32, device=d.device)
    rc = lib().sgm_stage_remove_peaks(out.data_ptr(), W, H, strips, ratio, work.data_ptr(), stream)
    assert rc == SGM_OK, last_error()
    return out


def _no_alloc(d, strips, ratio, stream):
    H, W = d.shape
    out = d.clone()
    rc = lib().sgm_stage_remove_peaks_noalloc(out.data_ptr(), W, H, strips, ratio, stream)
    assert rc == SGM_OK, last_error()
    return out


def _eq(a, b, what):
    a = a.cpu().numpy().view(np.uint32)
    b = b.cpu().numpy().view(np.uint32)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} of {a.size} differ, first at {tuple(bad[0])}"


def test_no_alloc_matches_workspace_form_on_two_calls(cuda):
    stream = torch.cuda.current_stream(cuda).cuda_stream
    cases = [(1080, 1920, 1, 0.0005, 11), (375, 1242, 6, 0.005, 12)]
    assert cases[0][0] * cases[0][1] <= MAX_PIXELS
    assert cases[1][0] * cases[1][1] < cases[0][0] * cases[0][1]
    for H, W, strips, ratio, seed in cases:
        print(f"[no-alloc peaks] {W}x{H} strips={strips} ratio={ratio}", flush=True)
        d = torch.from_numpy(_map(seed, H, W)).to(cuda)
        exp = _with_workspace(d, strips, ratio, stream)
        got = _no_alloc(d, strips, ratio, stream)
        torch.cuda.synchronize()
        _eq(got, exp, f"{W}x{H}")
        removed = int(((d != exp) & (exp == float(INVALID))).sum())
        assert removed > 0, "the case removes no segment"


def test_no_alloc_refuses_maps_above_its_scratch(cuda):
    d = torch.zeros((MAX_PIXELS // 1024 + 1, 1024), dtype=torch.float32, device=cuda)
    rc = lib().sgm_stage_remove_peaks_noalloc(d.data_ptr(), 1024, d.shape[0], 1, 0.001, None)
    assert rc == SGM_E_INVALID_ARG
    assert "2097152" in last_error()
    assert bool((d == 0).all())
