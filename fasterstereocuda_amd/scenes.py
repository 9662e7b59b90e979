"""Deterministic synthetic rectified stereo pairs for the benchmark workloads.

Every generator is seeded (numpy PCG64, seed = 1000 * workload + pair index)
so the host produces the same bytes for the CPU oracle and the GPU.  All
scenes are built the same way: a textured left image, a ground-truth
left-view disparity field, and a right image made by forward-warping the
left image (the larger disparity wins where two pixels land on one column),
filling the columns nobody lands on with fresh noise and adding Gaussian
noise.

Workloads:
  cone_like      450x375, 6 planes in a Voronoi partition, d in [2, 60]
  driving        1242x375, ground ramp 5 -> 60 from row 150 to the bottom,
                 3..8 fronto-parallel boxes with d in U[10, 63]
  building       4800x3409, 40 planes with d in [0, 250]
  uneven batch   64 pairs with sizes from five formats and D in {64, 128}
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class StereoPair:
    left: np.ndarray        # uint8 [H, W]
    right: np.ndarray       # uint8 [H, W]
    disparity: np.ndarray   # float32 [H, W], left-view ground truth
    d_range: int            # disparity range the workload is matched with

    @property
    def shape(self):
        return self.left.shape


def _rng(workload: int, index: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(1000 * workload + index))


def value_noise(rng: np.random.Generator, H: int, W: int, cells=(16, 8, 4, 2)) -> np.ndarray:
    """Sum of octaves of bilinearly upsampled U[0,1) noise, amplitude halving
    per octave, rescaled to [0, 255] (float32).  The interpolation runs along
    x on the coarse rows first, then along y: the same products and sums per
    pixel as the direct bilinear formula, in the same order."""
    acc = np.zeros((H, W), np.float64)
    amp = 1.0
    for s in cells:
        gh, gw = H // s + 2, W // s + 2
        g = rng.random((gh, gw))
        fy = np.arange(H) / s
        fx = np.arange(W) / s
        y0 = np.floor(fy).astype(np.int64)
        x0 = np.floor(fx).astype(np.int64)
        wy = (fy - y0)[:, None]
        wx = (fx - x0)[None, :]
        rows = g[:, x0] * (1 - wx) + g[:, x0 + 1] * wx          # (gh, W)
        acc += amp * (rows[y0] * (1 - wy) + rows[y0 + 1] * wy)
        amp *= 0.5
    lo, hi = acc.min(), acc.max()
    return ((acc - lo) * (255.0 / max(hi - lo, 1e-12))).astype(np.float32)


def _warp_right(rng: np.random.Generator, left: np.ndarray, disp: np.ndarray,
                sigma: float) -> np.ndarray:
    """Forward-warp the left image by its disparity (x_r = round(x - d));
    where two pixels land on one column the larger disparity wins."""
    H, W = left.shape
    ys, xs = np.mgrid[0:H, 0:W]
    xr = np.rint(xs - disp).astype(np.int64)
    ok = (xr >= 0) & (xr < W)
    tgt = (ys * W + xr)[ok]
    dv = disp[ok]
    src = left[ok]
    order = np.lexsort((dv, tgt))          # by target, then by disparity
    tgt, src = tgt[order], src[order]
    last = np.ones(tgt.shape, bool)
    last[:-1] = tgt[1:] != tgt[:-1]        # the largest disparity per target
    right = rng.random((H, W)).astype(np.float32) * 255.0   # occlusion fill
    right.reshape(-1)[tgt[last]] = src[last]
    right = right + rng.normal(0.0, sigma, size=(H, W)).astype(np.float32)
    return np.clip(np.rint(right), 0, 255).astype(np.uint8)


def _planes(rng: np.random.Generator, H: int, W: int, n: int, dmin: float, dmax: float,
            slope: float) -> np.ndarray:
    """Piecewise-planar disparity: n planes on a Voronoi partition (each
    pixel takes the plane of its nearest centre, the lower index on a tie).
    Distances and plane values are float64, evaluated in row blocks."""
    cy = rng.random(n) * H
    cx = rng.random(n) * W
    base = dmin + rng.random(n) * (dmax - dmin)
    gx = (rng.random(n) * 2 - 1) * slope
    gy = (rng.random(n) * 2 - 1) * slope
    xs = np.arange(W, dtype=np.float64)
    dx2 = (xs[None, :] - cx[:, None]) ** 2                       # (n, W)
    disp = np.empty((H, W), np.float64)
    rows = max(1, (1 << 23) // max(1, n * W))
    for y0 in range(0, H, rows):
        ys = np.arange(y0, min(H, y0 + rows), dtype=np.float64)
        dy2 = (ys[None, :] - cy[:, None]) ** 2                   # (n, r)
        k = np.argmin(dy2[:, :, None] + dx2[:, None, :], axis=0)  # (r, W), first minimum
        disp[y0:y0 + len(ys)] = (base[k] + gx[k] * (xs[None, :] - cx[k])
                                 + gy[k] * (ys[:, None] - cy[k]))
    return np.clip(disp, dmin, dmax).astype(np.float32)


def cone_like(index: int = 0, H: int = 375, W: int = 450, D: int = 64,
              workload: int = 0) -> StereoPair:
    """450x375-style scene: 4-octave texture, 6 planes (Voronoi) with
    disparities in [2, 60] (scaled to the range for other D), sigma 2."""
    rng = _rng(workload, index)
    left_f = value_noise(rng, H, W)
    dmax = 60.0 * D / 64.0
    disp = _planes(rng, H, W, 6, 2.0, min(dmax, D - 4.0), 0.04)
    left = np.clip(np.rint(left_f), 0, 255).astype(np.uint8)
    right = _warp_right(rng, left.astype(np.float32), disp, 2.0)
    return StereoPair(left, right, disp, D)


def driving(index: int = 0, H: int = 375, W: int = 1242, D: int = 64,
            workload: int = 1) -> StereoPair:
    """Driving-like scene: ground ramp from 5 at row 150 to 60 at the bottom
    row (5 above the horizon), 3..8 fronto-parallel boxes with disparity
    U[10, 63] and width U[40, 200], 4-octave value-noise texture, sigma 2 on
    the right image."""
    rng = _rng(workload, index)
    left_f = value_noise(rng, H, W)
    horizon = min(150, H - 2)
    rows = np.arange(H, dtype=np.float32)
    ramp = 5.0 + (rows - horizon) * (55.0 / max(H - 1 - horizon, 1))
    ground = np.where(rows < horizon, 5.0, ramp).astype(np.float32)
    disp = np.repeat(ground[:, None], W, axis=1)
    nbox = int(rng.integers(3, 9))
    boxes = []
    for _ in range(nbox):
        d = float(rng.uniform(10.0, 63.0))
        w = int(rng.integers(40, 201))
        h = int(rng.integers(30, 151))
        x0 = int(rng.integers(0, max(W - w, 1)))
        yb = int(rng.integers(horizon, H))
        boxes.append((d, x0, w, max(yb - h, 0), yb))
    for d, x0, w, y0, y1 in sorted(boxes):           # nearer boxes painted last
        disp[y0:y1 + 1, x0:x0 + w] = d
    disp = np.minimum(disp, D - 1.0).astype(np.float32)
    left = np.clip(np.rint(left_f), 0, 255).astype(np.uint8)
    right = _warp_right(rng, left.astype(np.float32), disp, 2.0)
    return StereoPair(left, right, disp, D)


def building(index: int = 0, H: int = 3409, W: int = 4800, D: int = 256,
             workload: int = 2) -> StereoPair:
    """Large scene: 40 planes with disparities in [0, 250], high-frequency
    value-noise texture, sigma 3."""
    rng = _rng(workload, index)
    left_f = value_noise(rng, H, W, cells=(8, 4, 2, 1))
    disp = _planes(rng, H, W, 40, 0.0, 250.0, 0.02)
    left = np.clip(np.rint(left_f), 0, 255).astype(np.uint8)
    right = _warp_right(rng, left.astype(np.float32), disp, 3.0)
    return StereoPair(left, right, disp, D)


UNEVEN_FORMATS = ((450, 375), (640, 480), (1242, 375), (1280, 720), (1920, 1080))


def uneven_specs(n: int = 64, workload: int = 3):
    """(W, H, D) of the uneven batch: sizes uniform over five formats, D
    uniform over {64, 128}."""
    rng = _rng(workload, 999)
    fmt = rng.integers(0, len(UNEVEN_FORMATS), size=n)
    ds = rng.integers(0, 2, size=n)
    return [(UNEVEN_FORMATS[f][0], UNEVEN_FORMATS[f][1], 64 << int(k)) for f, k in zip(fmt, ds)]


def uneven_batch(n: int = 64, workload: int = 3):
    return [cone_like(i, H=h, W=w, D=d, workload=workload)
            for i, (w, h, d) in enumerate(uneven_specs(n, workload))]


def driving_batch(n: int = 32, D: int = 64, workload: int = 1, H: int = 375, W: int = 1242):
    return [driving(i, H=H, W=W, D=D, workload=workload) for i in range(n)]
