#!/usr/bin/env python3
"""Throughput of the GPU Census-SGM matcher on one MI355X.

A step is one complete match_stereo of the 24 MP pair (6048 x 4032, disparity
span 512, automatic 5-level hierarchy, 8 paths, uint16 costs, the reference's
default options) with the images already resident in HBM.  Prints one JSON
line: pairs/s (value), aggregated Gcell-path/s, per-pair latency percentiles,
the roofline of the aggregation (path) kernels and the CPU oracle's rate on a
bounded band of the same pair.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--config 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pairs/s + Gcell-path/s at 24MP/512d; p50/p99 pair latency; bit-exact"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
WORKLOADS = {
    0: "640x480 pair, disparities 0..63, levels 0 (single full-range level)",
    1: "1920x1080 pair, disparities 0..255, automatic 4-level hierarchy",
    2: "6048x4032 (24 MP) pair, disparity span 512, automatic 5-level hierarchy, 30% low-texture",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else 0.0


def measured_traffic(config):
    """HBM bytes per path launch from the committed FETCH_SIZE / WRITE_SIZE profile
    of this workload (None if there is none)."""
    if config != 2:
        return None
    path = os.path.join(ROOT, "profiles", "k_path_24mp_512d_pmc.md")
    try:
        for line in open(path):
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(pair, options, rows, threads):
    """The oracle on rows [0, rows) of the pair; rate scaled to whole pairs."""
    from oracle import oracle as O
    h = pair.left.shape[0]
    rows = min(rows, h)
    t0 = time.perf_counter()
    r = O.match_stereo(pair.left[:rows], pair.right[:rows], None, None, pair.dmin, pair.nd, options, threads=threads)
    dt = time.perf_counter() - t0
    frac = rows / h
    return {"value": frac / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"rows 0..{rows - 1} of the {pair.left.shape[1]}x{h} pair ({frac:.3f} of it, "
                      f"{r.levels + 1} levels, {r.cells} cells) in {dt:.2f} s, scaled to whole pairs"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-rows", type=int, default=2016, help="rows of the CPU-baseline band (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent, see DESIGN.md for how they would shard)", file=sys.stderr)
        return 2

    import torch

    from regard3d_mvs_amd import StereoOptions, synth
    from regard3d_mvs_amd.native import Matcher

    t = time.perf_counter()
    pair = synth.config_pair(args.config)
    h, w = pair.left.shape
    log(f"bench: generated {WORKLOADS[args.config]} in {time.perf_counter() - t:.1f} s")
    options = StereoOptions(levels=0) if args.config == 0 else StereoOptions()

    dev = torch.device("cuda:0")
    left = torch.from_numpy(pair.left.view(np.int16)).to(dev).view(torch.uint16)
    right = torch.from_numpy(pair.right.view(np.int16)).to(dev).view(torch.uint16)
    disp = torch.empty((h, w), dtype=torch.float32, device=dev)
    conf = torch.empty((h, w), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    m = Matcher(0)
    m.set_profiling(True)
    for i in range(args.warmup):
        r = m.match(left, right, None, None, pair.dmin, pair.nd, options, out=(disp, conf))
        log(f"bench: warmup {i + 1}/{args.warmup}: {r.seconds * 1e3:.1f} ms")
    torch.cuda.synchronize()
    lat, agg_s, agg_b, launches = [], 0.0, 0, 0
    stages = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        r = m.match(left, right, None, None, pair.dmin, pair.nd, options, out=(disp, conf))
        lat.append(time.perf_counter() - s0)
        agg_s += r.stage_seconds["aggregate"]
        agg_b += r.aggregate_bytes
        launches += r.aggregate_launches
        for k, v in r.stage_seconds.items():
            stages[k] = stages.get(k, 0.0) + v
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    steps = args.steps
    pairs_s = steps / elapsed
    cells = r.cells
    valid = float(torch.isfinite(disp).float().mean().item())
    log("bench: stage ms/pair: " + ", ".join(f"{k} {1e3 * v / steps:.2f}" for k, v in stages.items()))
    log(f"bench: {steps} pairs in {elapsed:.3f} s; {cells} cells ({cells / (w * h):.1f}/px), "
        f"{r.tile_pixels} tile px, valid {valid:.3f}")

    achieved = agg_b / agg_s if agg_s > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": measured_traffic(args.config),
                "kernel": "k_path (8 SGM path scans per level)",
                "algorithmic_bytes_per_launch": agg_b / max(launches, 1),
                "avg_launch_ms": 1e3 * agg_s / max(launches, 1)}

    cpu = None
    if args.cpu_rows > 0:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline on {args.cpu_rows} rows with {threads} threads ...")
        cpu = cpu_baseline(pair, options, args.cpu_rows, threads)
        log(f"bench: CPU baseline {cpu['value']:.4f} pairs/s ({cpu['sample']})")

    out = {
        "metric": METRIC,
        "value": pairs_s,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config], "width": w, "height": h, "disparity_min": pair.dmin,
                   "num_disparities": pair.nd, "levels": r.levels, "paths": 8, "form": "plain, one pair per call",
                   "inputs": "uint16 (u8 x 257) resident in HBM", "seed": pair.seed},
        "gcell_path_per_s": cells * 8 * steps / elapsed / 1e9,
        "cells_per_pair": cells,
        "latency_ms": {"p50": 1e3 * percentile(lat, 50), "p95": 1e3 * percentile(lat, 95),
                       "p99": 1e3 * percentile(lat, 99)},
        "stage_ms_per_pair": {k: 1e3 * v / steps for k, v in stages.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
