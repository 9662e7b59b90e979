"""Batched matching: many stereo pairs (sizes and disparity ranges may
differ) per launch, through sgm_batch_* of the C ABI.

    b = SgmBatch(pairs, CuSGMOption.benchmark())   # pairs: list of dicts of tensors
    b.run()                                        # enqueue on torch's current stream

Each pair dict holds device tensors ``left``, ``right`` (uint8 [H, W]),
``disp_left`` and optionally ``disp_right`` (float32 [H, W]), plus
``min_disparity`` and ``disp_range``.  Per pair the output is byte-identical
to StereoCuda.Match with the same size, range and options.
"""
from __future__ import annotations

import ctypes

from ._lib import SgmPairDesc, check, last_error, lib
from .stereo import CuSGMOption

STAGES = ("census", "cost", "aggregate3", "aggregate_wta", "post")

# scanline schedules of the two aggregation launches
SCHED_GRID, SCHED_JOB_MAJOR, SCHED_LONGEST_FIRST = 0, 1, 2


class SgmBatch:
    def __init__(self, pairs, option: CuSGMOption):
        import torch
        arr = (SgmPairDesc * len(pairs))()
        self._keep = []
        for i, p in enumerate(pairs):
            H, W = p["left"].shape
            for k, dt in (("left", torch.uint8), ("right", torch.uint8), ("disp_left", torch.float32)):
                t = p[k]
                if not t.is_cuda or t.dtype != dt or tuple(t.shape) != (H, W) or not t.is_contiguous():
                    raise ValueError(f"pair {i}: {k} must be a contiguous {dt} CUDA tensor [{H}, {W}]")
            dr = p.get("disp_right")
            if dr is not None and (not dr.is_cuda or dr.dtype != torch.float32
                                   or tuple(dr.shape) != (H, W) or not dr.is_contiguous()):
                raise ValueError(f"pair {i}: disp_right must be a contiguous float32 CUDA tensor "
                                 f"[{H}, {W}]")
            arr[i] = SgmPairDesc(W, H, int(p.get("min_disparity", 0)), int(p["disp_range"]),
                                 p["left"].data_ptr(), p["right"].data_ptr(),
                                 p["disp_left"].data_ptr(), None if dr is None else dr.data_ptr())
            self._keep.append(p)
        opt = option.to_c()
        self._lib = lib()    # kept for __del__, which may run after module teardown
        self._b = self._lib.sgm_batch_create(arr, len(pairs), ctypes.byref(opt))
        if not self._b:
            raise RuntimeError(f"sgm_batch_create failed: {last_error()}")
        self.n = len(pairs)

    def __del__(self):
        b, L = getattr(self, "_b", None), getattr(self, "_lib", None)
        if b and L is not None:
            L.sgm_batch_destroy(b)
            self._b = None

    @property
    def device_bytes(self) -> int:
        return lib().sgm_batch_device_bytes(self._b)

    def run(self, stream=None) -> None:
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        check(lib().sgm_batch_run(self._b, stream))

    def set_schedule(self, policy: int, waves: int = 0) -> None:
        check(lib().sgm_batch_set_schedule(self._b, int(policy), int(waves)))

    @property
    def schedule(self):
        """(policy, resident waves) of the aggregation launches."""
        p, w = ctypes.c_int32(0), ctypes.c_int32(0)
        check(lib().sgm_batch_get_schedule(self._b, ctypes.byref(p), ctypes.byref(w)))
        return p.value, w.value

    def autotune(self, reps: int = 3, force: bool = False, stream=None):
        """Times the candidate schedules, keeps the fastest (remembered per
        batch signature); returns (policy, waves, ms per run; -1 if the
        remembered choice was applied without timing)."""
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        p, w, ms = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_float(0)
        check(lib().sgm_batch_autotune(self._b, stream, int(reps), int(force), ctypes.byref(p),
                                       ctypes.byref(w), ctypes.byref(ms)))
        return p.value, w.value, ms.value

    def use_graph(self, on: bool = True) -> None:
        """Run the whole pipeline as one captured HIP graph."""
        check(lib().sgm_batch_use_graph(self._b, int(on)))

    def set_timing(self, on: bool) -> None:
        check(lib().sgm_batch_set_timing(self._b, int(on)))

    def stage_times(self):
        """(dict stage -> summed ms, number of recorded runs) since the last call."""
        ms = (ctypes.c_float * 5)()
        runs = ctypes.c_int32(0)
        check(lib().sgm_batch_stage_times(self._b, ms, ctypes.byref(runs)))
        return dict(zip(STAGES, [float(v) for v in ms])), runs.value
