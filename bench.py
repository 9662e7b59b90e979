"""Benchmark of the SGM hot path on one MI355X.

Workload (default): 32 independent driving-like pairs, 1242x375, uint8,
D = 64, 9x7 census, P1 = 10, P2 = 150, uniqueness 0.95, LR check (tol 1) and
3x3 median.  One step = the whole pipeline over the 32 pairs, inputs and
outputs resident in HBM.  The headline value is the batched path
(sgm_batch_run: one launch per stage for all pairs); the plain path
(StereoCuda, one Match per pair) is timed beside it.

Prints ONE JSON line:
  value          W*H*D/s summed over the batch (batched path)
  roofline       aggregation kernel: algorithmic 6 B/cell/view (4 B cost
                 reads + 2 B summed output) / its average duration from HIP
                 events recorded around the launch during the timed steps
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

HBM_PEAK_GBS = 8000.0
AGG_BYTES_PER_CELL = 6.0     # per view: 4 path reads of the u8 cost + one u16 sum write


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
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a float32 negate of 256 MiB (counter calibration)")
    return ap.parse_args()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}, "
              f"WORLD_SIZE={world}); pairs are independent, so several GPUs would each run "
              "this same single-GPU batch on their own share of the pairs.", file=sys.stderr)
        sys.exit(2)

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
    agg_ms = stage_avg["aggregate"]
    agg_bytes = AGG_BYTES_PER_CELL * cells_per_pair * NP * 2      # both views
    achieved = agg_bytes / (agg_ms * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "aggregate_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if (t.get("pairs"), t.get("width"), t.get("height"), t.get("disp_range")) == (NP, W, H, D):
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
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle                                    # CPU baseline leg only
        oracle.lib()
        oracle.match(pairs_np[0].left, pairs_np[0].right, 0, D)     # warm-up
        done = 0
        tc0 = time.perf_counter()
        exact = None
        while True:
            p = pairs_np[done % NP]
            dl, _ = oracle.match(p.left, p.right, 0, D)
            if done == 0:
                exact = bool((dl.view(np.uint32) == batched0.view(np.uint32)).all())
            done += 1
            if time.perf_counter() - tc0 >= args.cpu_seconds or done >= 4 * NP:
                break
        tc = time.perf_counter() - tc0
        cpu = {
            "value": done * cells_per_pair / tc,
            "unit": "W*H*D/s",
            "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"{done} pairs of the same workload, sequential, {tc:.1f} s",
            "gpu_output_bit_exact_vs_oracle": exact,
        }

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
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "aggregate_kernel",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": agg_bytes,
            "avg_launch_ms": agg_ms,
        },
        "cpu_baseline": cpu,
        "plain_value": plain,
        "plain_bit_exact_vs_batched": same_as_plain,
        "stage_ms": stage_avg,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
