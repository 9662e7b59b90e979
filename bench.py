"""Benchmark of the SGM hot path on one MI355X.

Workload (default): 32 independent driving-like pairs, 1242x375, uint8,
D = 64, 9x7 census, P1 = 10, P2 = 150, uniqueness 0.95, LR check (tol 1) and
3x3 median.  One step = the whole pipeline over the 32 pairs, inputs and
outputs resident in HBM.  The headline value is the batched path
(sgm_batch_run: one launch per stage for all pairs); the plain path
(StereoCuda, one Match per pair) is timed beside it.

Prints ONE JSON line:
  value          W*H*D/s summed over the batch (batched path)
  roofline       aggregation + WTA (the launches between the census and the
                 post stage: with the cost computed on the fly, left->right +
                 right->left, top->bottom, then bottom->top fused with the
                 WTA): algorithmic 8 B/cell/view (4 B cost reads + 2 B summed
                 output + 2 B read by the WTA) / their average duration from
                 HIP events recorded around them during the timed steps;
                 `traffic` is the measured HBM bytes of the same launches from
                 the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
                 this workload
  cpu_baseline   the OpenMP oracle on a bounded sample of the same pairs
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads stay on neighbouring cores; set before any
# OpenMP runtime starts (the thread count is the environment's own)
os.environ.setdefault("OMP_PROC_BIND", "close")
# CPUs this process may run on, read before an OpenMP runtime binds the main
# thread to a single place (afterwards the main thread's mask is one CPU)
try:
    CPUS_USABLE = len(os.sched_getaffinity(0))
except (AttributeError, OSError):
    CPUS_USABLE = None

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell and view of aggregation + WTA: 4 direction reads
# of the u8 cost + one u16 4-direction sum write (6 B), and the WTA's read of
# that sum (2 B).  The engine runs them as the horizontal and top->bottom
# launches, then bottom->top fused with the WTA.
AGG_BYTES_PER_CELL = 8.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=32)
    ap.add_argument("--width", type=int, default=1242)
    ap.add_argument("--height", type=int, default=375)
    ap.add_argument("--range", dest="disp_range", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain path")
    ap.add_argument("--cpu-runs", type=int, default=5,
                    help="timed CPU baseline runs (median reported), after one warm-up")
    ap.add_argument("--cpu-pairs", type=int, default=16,
                    help="pairs of the workload per CPU baseline run")
    ap.add_argument("--schedule", default="grid", choices=["grid", "longest-first", "autotune"],
                    help="scanline schedule of the batched aggregation launches")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the pyramid, uneven-batch and differentiable workloads in the "
                         "default line")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a float32 negate of 256 MiB (counter calibration)")
    ap.add_argument("--workload", default="driving",
                    choices=["driving", "building", "uneven", "diffsgm"],
                    help="driving: the default 32-pair batch; the others measure the "
                         "large-image pyramid, the uneven batch and the differentiable form")
    return ap.parse_args()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}, "
              f"WORLD_SIZE={world}); pairs are independent, so several GPUs would each run "
              "this same single-GPU batch on their own share of the pairs.", file=sys.stderr)
        sys.exit(2)

    if args.workload != "driving":
        run_other(args)
        return

    import numpy as np
    import torch

    from fasterstereocuda_amd import scenes
    from fasterstereocuda_amd.batch import SgmBatch
    from fasterstereocuda_amd.stereo import CuSGMOption, StereoCuda

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    W, H, D, NP = args.width, args.height, args.disp_range, args.pairs
    opt = CuSGMOption.benchmark()
    pairs_np = scenes.driving_batch(NP, D=D, H=H, W=W)
    pairs = []
    for p in pairs_np:
        pairs.append({
            "left": torch.from_numpy(p.left).to(dev),
            "right": torch.from_numpy(p.right).to(dev),
            "disp_left": torch.empty((H, W), dtype=torch.float32, device=dev),
            "disp_right": None,
            "min_disparity": 0,
            "disp_range": D,
        })
    batch = SgmBatch(pairs, opt)
    from fasterstereocuda_amd.batch import SCHED_GRID, SCHED_LONGEST_FIRST
    if args.schedule == "autotune":
        batch.autotune(reps=3, force=True)
    else:
        batch.set_schedule({"grid": SCHED_GRID, "longest-first": SCHED_LONGEST_FIRST}[args.schedule])
    schedule = batch.schedule
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        batch.run()
    torch.cuda.synchronize()
    batch.stage_times()                     # drop anything recorded so far
    batch.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    batch.set_timing(False)
    stage_ms, runs = batch.stage_times()
    elapsed = t1 - t0
    cells_per_pair = W * H * D
    value = NP * cells_per_pair * args.steps / elapsed
    ms_per_step = elapsed * 1e3 / args.steps
    stage_avg = {k: v / max(runs, 1) for k, v in stage_ms.items()}
    agg_ms = stage_avg["aggregate3"] + stage_avg["aggregate_wta"]
    agg_bytes = AGG_BYTES_PER_CELL * cells_per_pair * NP * 2      # both views
    achieved = agg_bytes / (agg_ms * 1e-3) / 1e9

    # which aggregation kernels ran: the cost from the census on the fly
    # (D <= 256, 9x7 census) or the stored cost volume
    from fasterstereocuda_amd.stereo import get_cost_mode
    path = "on-the-fly" if (get_cost_mode() == 0 and D <= 256) else "stored"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "aggregate_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if (t.get("pairs"), t.get("width"), t.get("height"), t.get("disp_range"),
                t.get("path")) == (NP, W, H, D, path):
            traffic = t.get("bytes_per_launch")

    # plain path: one StereoCuda Match per pair on the same device buffers
    plain = None
    out_plain0 = None
    if not args.no_plain:
        st = StereoCuda()
        assert st.Init(W, H, 0, D, opt), st.last_error
        outs = [torch.empty((H, W), dtype=torch.float32, device=dev) for _ in range(NP)]
        for p, o in zip(pairs, outs):
            assert st.match_device(p["left"], p["right"], o)
        torch.cuda.synchronize()
        psteps = max(1, args.steps // 4)
        t2 = time.perf_counter()
        for _ in range(psteps):
            for p, o in zip(pairs, outs):
                st.match_device(p["left"], p["right"], o)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        plain = NP * cells_per_pair * psteps / (t3 - t2)
        out_plain0 = outs[0]
        st.Release()

    # streaming rate HBM actually delivers for buffers far larger than the
    # 256 MB last-level cache: torch's copy of 2 GiB (read + write bytes)
    copy_gbs = hbm_copy_rate(dev)

    if args.copy_probe:
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()

    batched0 = pairs[0]["disp_left"].cpu().numpy()
    same_as_plain = None
    if out_plain0 is not None:
        same_as_plain = bool((out_plain0.cpu().numpy().view(np.uint32) == batched0.view(np.uint32)).all())

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pairs_np, D, batched0, args.cpu_runs, args.cpu_pairs)

    # host-pointer path (the reference's calling pattern: copies over PCIe
    # inside every Match), one pair at a time, both reference image sizes
    host = None if args.no_plain else host_match_rates(opt)

    line = {
        "metric": "W*H*D/s bit-exact SGM (plain vs batched)",
        "value": value,
        "unit": "W*H*D/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"driving-like, {NP} pairs {W}x{H}, D={D}, 9x7 census, P1=10 P2=150, "
                        "uniqueness 0.95, LR check tol 1, 3x3 median",
            "pairs": NP, "width": W, "height": H, "disp_range": D, "path": "batched",
            "schedule": {"policy": ["grid", "job-major", "longest-first"][schedule[0]],
                         "waves": schedule[1]},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("wide_h_kernel + wide_v_kernel<false> + wide_v_kernel<true> (bottom->top "
                       "fused with WTA), cost from the census on the fly" if path == "on-the-fly" else
                       "aggregate_kernel<1> + aggregate_kernel<2> (bottom->top fused with WTA), "
                       "stored cost volume"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": agg_bytes,
            "avg_launch_ms": agg_ms,
            "copy_2GiB_GBps": copy_gbs,
            "frac_of_copy": achieved / copy_gbs,
        },
        "cpu_baseline": cpu,
        "plain_value": plain,
        "host_path": host,
        "plain_bit_exact_vs_batched": same_as_plain,
        "stage_ms": stage_avg,
    }
    if not args.no_extra:
        # the other measured workloads, each on its own data, after the headline
        del pairs, batch
        torch.cuda.empty_cache()
        for key, fn in (("diffsgm_batch8_1242x375_d128", diffsgm),
                        ("pyramid_building_4800x3409_d256", building),
                        ("uneven_batch64", uneven)):
            line[key] = fn(args, quick=True)
            torch.cuda.empty_cache()
    print(json.dumps(line), flush=True)


def cpu_baseline(pairs_np, D, gpu_first, runs, per_run):
    """The OpenMP oracle on the host cores: one warm-up, then the median of
    `runs` runs, each matching `per_run` pairs of the workload in sequence."""
    import numpy as np
    import statistics
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle                                    # CPU baseline leg only
    oracle.lib()
    H, W = pairs_np[0].left.shape
    first, _ = oracle.match(pairs_np[0].left, pairs_np[0].right, 0, D)      # warm-up
    exact = bool((first.view(np.uint32) == gpu_first.view(np.uint32)).all())
    rates = []
    for r in range(runs):
        t0 = time.perf_counter()
        for k in range(per_run):
            p = pairs_np[(r * per_run + k) % len(pairs_np)]
            oracle.match(p.left, p.right, 0, D)
        rates.append(per_run * W * H * D / (time.perf_counter() - t0))
    phys, logical = host_cores()
    return {
        "value": statistics.median(rates),
        "unit": "W*H*D/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"median of {runs} runs after 1 warm-up; a run = {per_run} pairs of the "
                  f"workload ({W}x{H}, D={D}) matched in sequence, each pair parallel over "
                  "rows / scanlines",
        "runs": rates,
        "gpu_output_bit_exact_vs_oracle": exact,
        "cpu_model": cpu_model(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
        "cpus_usable": CPUS_USABLE,
        "machine_physical_cores": phys,
        "machine_logical_cpus": logical,
    }


def host_cores():
    """(physical cores, logical CPUs) of the machine from /proc/cpuinfo."""
    phys, logical = set(), 0
    cur = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    if "core id" in cur:
                        phys.add((cur.get("physical id"), cur["core id"]))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "processor":
                    logical += 1
                cur[k] = v.strip()
        if "core id" in cur:
            phys.add((cur.get("physical id"), cur["core id"]))
    except OSError:
        return None, None
    return (len(phys) or None), (logical or None)


def cpu_model() -> str:
    """The CPU's model name with its family / model / stepping numbers (some
    virtual machines report only a short marketing name)."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k in ("model name", "cpu family", "model", "stepping", "vendor_id") and k not in info:
                    info[k] = v.strip()
                if not line.strip() and info:
                    break
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    return (f"{name} ({info.get('vendor_id', '?')} family {info.get('cpu family', '?')} "
            f"model {info.get('model', '?')} stepping {info.get('stepping', '?')})")


def host_match_rates(opt, reps: int = 20):
    """ms per StereoCuda.Match on host buffers, PCIe copies included: page-locked
    buffers (the reference console allocates them with MallocPageLockedPtr) and,
    beside them, ordinary pageable numpy arrays."""
    import numpy as np
    import torch
    from fasterstereocuda_amd import scenes
    from fasterstereocuda_amd.stereo import StereoCuda
    out = {}
    for name, pair in (("cone_450x375_d64", scenes.cone_like(0)),
                       ("driving_1242x375_d64", scenes.driving(0))):
        H, W = pair.left.shape
        st = StereoCuda()
        assert st.Init(W, H, 0, 64, opt), st.last_error

        def rate(left, right, dl, dr):
            t0 = time.perf_counter()
            for _ in range(reps):
                st.Match(left, right, dl, dr)
            return (time.perf_counter() - t0) * 1e3 / reps

        pinned = [torch.from_numpy(pair.left).pin_memory().numpy(),
                  torch.from_numpy(pair.right).pin_memory().numpy(),
                  torch.empty((H, W), dtype=torch.float32).pin_memory().numpy(),
                  torch.empty((H, W), dtype=torch.float32).pin_memory().numpy()]
        pageable = [pair.left, pair.right, np.empty((H, W), np.float32),
                    np.empty((H, W), np.float32)]
        for bufs in (pinned, pageable):           # warm both paths
            for _ in range(3):
                st.Match(*bufs)
        # best of three interleaved rounds: host timings on a shared box are noisy
        ms, ms_pageable = float("inf"), float("inf")
        for _ in range(3):
            ms = min(ms, rate(*pinned))
            ms_pageable = min(ms_pageable, rate(*pageable))
        st.Release()
        out[name] = {"ms_per_pair": ms, "W*H*D/s": W * H * 64 / (ms * 1e-3),
                     "ms_per_pair_pageable": ms_pageable}
    return out


def hbm_copy_rate(dev, gib: int = 2, reps: int = 5) -> float:
    """GB/s (bytes read + written) of torch's device copy over `gib` GiB."""
    import torch
    a = torch.empty(gib << 28, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    rate = 2.0 * a.numel() * 4 * reps / (time.perf_counter() - t0) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return rate


# ---------------------------------------------------------------- other workloads
# building: 4800x3409, D = 256 -- 3-level pyramid (window 16) and the dense
#           single-level D = 256 pipeline (a ~4.2 GB cost volume per view)
# uneven:   64 pairs, five formats, D in {64, 128} -- batched vs plain
# diffsgm:  differentiable FP32 SGM forward + backward, batch 8, 1242x375, D = 128

def _emit(line):
    print(json.dumps(line), flush=True)


def _common(args, metric, value, unit, ms, hib, config, **extra):
    line = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": hib, "scaling": "weak",
            "vs_baseline": None, "data": "synthetic", "config": config}
    line.update(extra)
    return line


def building(args, quick=False):
    import numpy as np
    import torch
    from fasterstereocuda_amd import scenes
    from fasterstereocuda_amd.pyramid import SgmPyramid
    from fasterstereocuda_amd.stereo import CuSGMOption, StereoCuda

    W, H, D = 4800, 3409, 256
    p = scenes.building(0)
    dev = torch.device("cuda:0")
    L = torch.from_numpy(p.left).to(dev)
    R = torch.from_numpy(p.right).to(dev)
    out = torch.empty((H, W), dtype=torch.float32, device=dev)
    opt = CuSGMOption.benchmark()
    pyr = SgmPyramid(W, H, 0, D, 3, 16, opt)
    steps = min(args.steps, 10) if quick else args.steps
    for _ in range(max(1, args.warmup)):
        pyr.match_device(L, R, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pyr.match_device(L, R, out)
    torch.cuda.synchronize()
    pyr_ms = (time.perf_counter() - t0) * 1e3 / steps
    pyr_out = out.cpu().numpy()
    # dense single level, D = 256 (both views, LR check): streams ~4.2 GB per volume
    st = StereoCuda()
    assert st.Init(W, H, 0, D, opt), st.last_error
    dense = torch.empty_like(out)
    st.match_device(L, R, dense)
    torch.cuda.synchronize()
    dsteps = 2 if quick else max(1, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(dsteps):
        st.match_device(L, R, dense)
    torch.cuda.synchronize()
    dense_ms = (time.perf_counter() - t0) * 1e3 / dsteps
    st.Release()
    cells = W * H * D
    valid = pyr_out != -999999.0
    if quick:
        return {"workload": "building 4800x3409, D=256, 3-level pyramid, window 16",
                "ms_per_pair": pyr_ms, "value": cells / (pyr_ms * 1e-3), "unit": "W*H*D/s (nominal D)",
                "steps": steps, "valid_fraction": float(valid.mean()),
                "dense_single_level": {"ms_per_pair": dense_ms, "value": cells / (dense_ms * 1e-3),
                                       "unit": "W*H*D/s", "steps": dsteps}}
    _emit(_common(args, "W*H*D/s (nominal D) large-image pyramid", cells / (pyr_ms * 1e-3), "W*H*D/s",
                  pyr_ms, True,
                  {"workload": "building 4800x3409, D=256, 3-level pyramid, window 16, 9x7 census, "
                               "P1=10 P2=150, LR at the coarsest level, median, dilation",
                   "width": W, "height": H, "disp_range": D},
                  dtype="u8",
                  dense_single_level={"ms": dense_ms, "value": cells / (dense_ms * 1e-3),
                                      "unit": "W*H*D/s"},
                  pyramid_valid_fraction=float(valid.mean())))


def uneven(args, quick=False):
    import numpy as np
    import torch
    from fasterstereocuda_amd import scenes
    from fasterstereocuda_amd.batch import SgmBatch
    from fasterstereocuda_amd.stereo import CuSGMOption, StereoCuda

    dev = torch.device("cuda:0")
    opt = CuSGMOption.benchmark()
    specs = scenes.uneven_specs(64)
    pairs, cells = [], 0
    for i, (w, h, d) in enumerate(specs):
        sc = scenes.cone_like(i, H=h, W=w, D=d, workload=3)
        pairs.append({"left": torch.from_numpy(sc.left).to(dev), "right": torch.from_numpy(sc.right).to(dev),
                      "disp_left": torch.empty((h, w), dtype=torch.float32, device=dev),
                      "min_disparity": 0, "disp_range": d})
        cells += w * h * d
    batch = SgmBatch(pairs, opt)
    steps = min(args.steps, 10) if quick else args.steps
    # grid-schedule eager time, then the autotuned schedule run as one graph
    for _ in range(max(1, args.warmup)):
        batch.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
    torch.cuda.synchronize()
    grid_ms = (time.perf_counter() - t0) * 1e3 / steps
    policy, waves, tuned_ms = batch.autotune(reps=3, force=True)
    batch.use_graph(True)
    for _ in range(max(1, args.warmup)):
        batch.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
    torch.cuda.synchronize()
    b_ms = (time.perf_counter() - t0) * 1e3 / steps
    # plain path: one handle per (W, H, D), one Match per pair
    handles = {}
    outs = []
    for p in pairs:
        h, w = p["left"].shape
        key = (w, h, p["disp_range"])
        if key not in handles:
            st = StereoCuda()
            assert st.Init(w, h, 0, p["disp_range"], opt), st.last_error
            handles[key] = st
        outs.append(torch.empty((h, w), dtype=torch.float32, device=dev))
    def plain_pass():
        for p, o in zip(pairs, outs):
            h, w = p["left"].shape
            handles[(w, h, p["disp_range"])].match_device(p["left"], p["right"], o)
    plain_pass()
    torch.cuda.synchronize()
    psteps = max(1, steps // 2)
    t0 = time.perf_counter()
    for _ in range(psteps):
        plain_pass()
    torch.cuda.synchronize()
    p_ms = (time.perf_counter() - t0) * 1e3 / psteps
    same = all(bool((o.cpu().numpy().view(np.uint32) == p["disp_left"].cpu().numpy().view(np.uint32)).all())
               for o, p in zip(outs, pairs))
    sched = {"policy": ["grid", "job-major", "longest-first"][policy], "waves": waves,
             "autotune_ms": tuned_ms, "grid_eager_ms": grid_ms, "graph": True}
    for st in handles.values():
        st.Release()
    if quick:
        return {"workload": "64 pairs, sizes from {450x375, 640x480, 1242x375, 1280x720, "
                            "1920x1080}, D in {64, 128}",
                "value": cells / (b_ms * 1e-3), "unit": "W*H*D/s", "ms_per_batch": b_ms,
                "plain_value": cells / (p_ms * 1e-3), "plain_ms_per_batch": p_ms,
                "batched_over_plain": p_ms / b_ms, "batched_bit_exact_vs_plain": same,
                "steps": steps, "schedule": sched}
    _emit(_common(args, "W*H*D/s summed over an uneven batch (batched vs plain)", cells / (b_ms * 1e-3),
                  "W*H*D/s", b_ms, True,
                  {"workload": "64 pairs, sizes from {450x375, 640x480, 1242x375, 1280x720, 1920x1080}, "
                               "D in {64, 128}, 9x7 census, P1=10 P2=150, uniqueness 0.95, LR, median",
                   "pairs": 64, "path": "batched"},
                  dtype="u8", plain_value=cells / (p_ms * 1e-3), plain_ms=p_ms,
                  batched_over_plain=p_ms / b_ms, batched_bit_exact_vs_plain=same,
                  schedule=sched, device_bytes=batch.device_bytes))


def diffsgm(args, quick=False):
    import numpy as np
    import torch
    from fasterstereocuda_amd import scenes
    from fasterstereocuda_amd.diff_sgm import census_cost_volume, diff_sgm

    dev = torch.device("cuda:0")
    B, H, W, D = 8, 375, 1242, 128
    sc = scenes.driving_batch(B, D=D, H=H, W=W, workload=4)
    left = torch.from_numpy(np.stack([s.left for s in sc])).to(dev)
    right = torch.from_numpy(np.stack([s.right for s in sc])).to(dev)
    gt = torch.from_numpy(np.stack([s.disparity for s in sc])).to(dev)
    cost = census_cost_volume(left, right, D)
    rng = np.random.default_rng(4000)
    p1 = torch.from_numpy(rng.uniform(0.05, 0.2, (B, H, W)).astype(np.float32)).to(dev)
    p2 = torch.from_numpy(rng.uniform(0.5, 2.0, (B, H, W)).astype(np.float32)).to(dev)

    def step():
        c = cost.detach().requires_grad_(True)
        a = p1.detach().requires_grad_(True)
        b = p2.detach().requires_grad_(True)
        disp = diff_sgm(c, a, b, 0.1)
        loss = (disp - gt).abs().mean()
        loss.backward()
        return loss

    steps = min(args.steps, 10) if quick else args.steps
    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    cells = B * H * W * D
    bytes_alg = 50.0 * cells                      # ~24 B fwd + ~26 B bwd per cell
    roof = {"bound": "hbm", "achieved": bytes_alg / (ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": bytes_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": _profile_traffic("diffsgm_traffic.json", (B, W, H, D)),
            "algorithmic_bytes_per_step": bytes_alg}
    if quick:
        return {"workload": "differentiable FP32 SGM, batch 8, 1242x375, D=128, per-pixel P1/P2, "
                            "soft-argmin T=0.1, L1 loss",
                "ms_per_pair": ms / B, "ms_per_step": ms, "unit": "ms/pair", "steps": steps,
                "loss": float(loss.detach()), "roofline": roof}
    _emit(_common(args, "diff-SGM fwd+bwd ms/pair", ms / B, "ms/pair", ms, False,
                  {"workload": "differentiable FP32 SGM, batch 8, 1242x375, D=128, per-pixel P1/P2, "
                               "soft-argmin T=0.1, L1 loss", "batch": B, "width": W, "height": H,
                   "disp_range": D},
                  dtype="fp32", loss=float(loss), roofline=roof))


def _profile_traffic(name, key):
    """Measured HBM bytes of a workload's step from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary, if one matches the workload."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if tuple(t.get("key", ())) != tuple(key):
        return None
    return t.get("bytes_per_step")


def run_other(args):
    import torch
    torch.cuda.set_device(0)
    {"building": building, "uneven": uneven, "diffsgm": diffsgm}[args.workload](args)
    if args.copy_probe:
        # counter calibration: a float32 negate of 256 MiB (reads and writes 256 MiB)
        a = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
