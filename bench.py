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


def _same(a, b):
    return a.shape == b.shape and bool(np.array_equal(a.view(np.uint32), b.view(np.uint32)))


def bench_rectify(args) -> int:
    """rectify_image on the 24 MP left image (uint16): a radial_k3 camera
    turned by a few tenths of a degree, resampled into a 6200x4150 pinhole
    target, image and mask left in HBM.  HBM-bound gather: the source read
    once (2 B/px), the image and mask written (3 B per target pixel)."""
    import torch

    from regard3d_mvs_amd import Camera, CameraModel, synth
    from regard3d_mvs_amd.native import Matcher

    pair = synth.config_pair(2)
    h, w = pair.left.shape
    ax, ay, az = 0.0017, -0.0044, 0.0044  # radians
    rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    rz = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    cam = Camera(w, h, 5200.0, 5200.0, w / 2 - 0.5, h / 2 - 0.5, CameraModel.radial_k3, (-0.06, 0.012, -0.001),
                 tuple((rz @ ry @ rx).ravel()))
    tgt = Camera(6200, 4150, 5150.0, 5150.0, 3100.0, 2075.0)
    src = torch.from_numpy(pair.left.astype(np.int32)).to(torch.uint16).cuda()
    torch.cuda.synchronize()
    with Matcher(0) as m:
        for _ in range(args.warmup):
            m.rectify(src, cam, tgt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out, valid = m.rectify(src, cam, tgt)
        torch.cuda.synchronize()
        secs = (time.perf_counter() - t0) / args.steps
    tpx = tgt.width * tgt.height
    alg = 2 * w * h + 3 * tpx
    cpu = None
    if args.cpu_rows > 0:
        from oracle import oracle as O
        t1 = time.perf_counter()
        c_out, c_valid = O.rectify(pair.left, cam, tgt)
        csecs = time.perf_counter() - t1
        exact = bool(np.array_equal(out.cpu().numpy(), c_out) and np.array_equal(valid.cpu().numpy(), c_valid))
        cpu = {"value": tpx / csecs / 1e6, "unit": "Mpx/s", "cores": 1, "kind": "port",
               "sample": f"the whole {tgt.width}x{tgt.height} target in {csecs:.2f} s", "bit_exact": exact}
    out_j = {"metric": "rectified target Mpx/s (rectify_image, uint16, radial_k3)", "value": tpx / secs / 1e6,
             "unit": "Mpx/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * secs,
             "higher_is_better": True, "dtype": "u16", "data": "synthetic",
             "config": {"workload": f"{w}x{h} uint16 source -> {tgt.width}x{tgt.height} pinhole target, radial_k3",
                        "inputs": "source resident in HBM, image and mask left in HBM"},
             "roofline": {"bound": "hbm", "achieved": alg / secs / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "frac": alg / secs / HBM_PEAK, "traffic": None, "kernel": "k_rectify<uint16>",
                          "algorithmic_bytes_per_launch": alg,
                          "launch_timing": "host wall time per call (one launch, synchronised)"},
             "cpu_baseline": cpu}
    print(json.dumps(out_j), flush=True)
    return 0


def bench_combine(args) -> int:
    """combine_pairs on a 6048x4536 grid (the reference rig of test_mvs.cpp
    scaled by 18.9) from four rectified pairs of the same size, maps in HBM,
    one pair 6 pixels off.  Per grid pixel: 4 bilinear samples of 4 bytes and
    a 4-byte confidence per pair read, 10 bytes written."""
    import torch

    from regard3d_mvs_amd import PairGeometry
    from regard3d_mvs_amd.multiview import Rig
    from regard3d_mvs_amd.native import Matcher

    offsets = ((0.8, 0.05, 0.1), (-0.7, 0.1, 0.0), (0.1, 0.75, -0.1), (-0.2, -0.7, 0.05))
    rig = Rig(offsets, scale=18.9)
    host = [rig.pair(c, shift=6.0 if i == 2 else 0.0) for i, c in enumerate(rig.partners)]
    dev = [PairGeometry(p.grid_to_rect, p.fb, p.offset, p.partner_center, torch.from_numpy(p.disparity).cuda(),
                        torch.from_numpy(p.confidence).cuda()) for p in host]
    grid = rig.grid()
    torch.cuda.synchronize()
    with Matcher(0) as m:
        for _ in range(args.warmup):
            m.combine_pairs(grid, dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            depth, conf, mask = m.combine_pairs(grid, dev)
        torch.cuda.synchronize()
        secs = (time.perf_counter() - t0) / args.steps
    px = grid.width * grid.height
    alg = px * (len(host) * 8 + 10)  # disparity (4) + confidence (4) per pair read, 10 written
    cpu = None
    if args.cpu_rows > 0:
        from oracle import oracle as O
        t1 = time.perf_counter()
        c = O.combine_pairs(grid, host)
        csecs = time.perf_counter() - t1
        exact = all(np.array_equal(a.cpu().numpy().view(np.uint8), b.view(np.uint8)) for a, b in zip((depth, conf, mask), c))
        cpu = {"value": px / csecs / 1e6, "unit": "Mpx/s", "cores": 1, "kind": "port",
               "sample": f"the whole {grid.width}x{grid.height} grid in {csecs:.2f} s", "bit_exact": bool(exact)}
    valid = float((depth > 0).float().mean().item())
    out_j = {"metric": "combined grid Mpx/s (combine_pairs, 4 pairs)", "value": px / secs / 1e6, "unit": "Mpx/s",
             "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * secs,
             "higher_is_better": True, "dtype": "f64", "data": "synthetic",
             "config": {"workload": f"{grid.width}x{grid.height} grid, 4 rectified pairs of a slanted plane, "
                                    "one 6 px off", "valid": valid,
                        "inputs": "maps resident in HBM, depth / confidence / mask left in HBM"},
             "roofline": {"bound": "hbm", "achieved": alg / secs / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "frac": alg / secs / HBM_PEAK, "traffic": None, "kernel": "k_combine",
                          "algorithmic_bytes_per_launch": alg,
                          "launch_timing": "host wall time per call (one launch, synchronised)"},
             "cpu_baseline": cpu}
    print(json.dumps(out_j), flush=True)
    return 0


def bench_batch(args) -> int:
    """The scene of views x 4 partners (1.5-24 MP, spans 64..512) matched as
    plain r3dg_match calls and as one r3dg_match_batch.  value: images already
    in HBM, maps left in HBM; the host-in / host-out rates are reported beside."""
    import torch

    from regard3d_mvs_amd import StereoOptions, workloads
    from regard3d_mvs_amd.native import Matcher

    t = time.perf_counter()
    src = workloads.source_pair()
    reqs = workloads.scene(src, views=args.views)
    log(f"bench: source {workloads.SOURCE_W}x{workloads.SOURCE_H} and {len(reqs)} pairs in "
        f"{time.perf_counter() - t:.1f} s; {sum(r.left.size for r in reqs) / len(reqs) / 1e6:.1f} MP mean")
    o = StereoOptions()
    dev = torch.device("cuda:0")

    def on_dev(a):
        return torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).to(dev).view(torch.uint16)

    views = {}
    dreqs = []
    for r in reqs:
        if r.view not in views:
            views[r.view] = on_dev(r.left)
        dreqs.append((views[r.view], on_dev(r.right), r.dmin, r.nd))
    torch.cuda.synchronize()
    m = Matcher(0)
    m.set_profiling(True)
    m.match(dreqs[0][0], dreqs[0][1], None, None, dreqs[0][2], dreqs[0][3], o)  # warm-up

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    plain, t_plain = timed(lambda: [m.match(l, r, None, None, a, n, o) for l, r, a, n in dreqs])
    st = {k: sum(p.stage_seconds[k] for p in plain) for k in plain[0].stage_seconds}
    log(f"bench: plain {len(reqs)} pairs in {t_plain:.2f} s; device s: " +
        ", ".join(f"{k} {v:.2f}" for k, v in st.items()))
    got, t_batch = timed(lambda: m.match_batch([dict(left=l, right=r, dmin=a, nd=n) for l, r, a, n in dreqs], o))
    log(f"bench: batched {len(reqs)} pairs in {t_batch:.2f} s; device s: " +
        ", ".join(f"{k} {v:.2f}" for k, v in got[0].stage_seconds.items()))
    exact = all(torch.equal(g.disparity.view(torch.int32), p.disparity.view(torch.int32)) and
                torch.equal(g.confidence.view(torch.int32), p.confidence.view(torch.int32))
                for g, p in zip(got, plain))
    cells = sum(p.cells for p in plain)
    del got, plain
    host_plain, th_plain = timed(lambda: [m.match(r.left, r.right, None, None, r.dmin, r.nd, o) for r in reqs])
    del host_plain
    host_batch, th_batch = timed(lambda: m.match_batch([dict(left=r.left, right=r.right, dmin=r.dmin, nd=r.nd)
                                                        for r in reqs], o))
    del host_batch
    log(f"bench: host in/out: plain {th_plain:.2f} s, batched {th_batch:.2f} s")
    out = {
        "metric": "pairs/s, batched scene vs plain per-pair calls", "unit": "pairs/s", "n_gpus": 1,
        "value": len(reqs) / t_batch, "plain_pairs_per_s": len(reqs) / t_plain,
        "speedup_vs_plain": t_plain / t_batch, "higher_is_better": True, "data": "synthetic", "dtype": "u16",
        "bit_exact_vs_plain": exact, "pairs": len(reqs), "cells": cells,
        "gcell_path_per_s": cells * 8 / t_batch / 1e9, "plain_gcell_path_per_s": cells * 8 / t_plain / 1e9,
        "host_in_out": {"batched_pairs_per_s": len(reqs) / th_batch, "plain_pairs_per_s": len(reqs) / th_plain},
        "config": {"workload": f"{args.views} views x 4 partners, 1.5-24 MP (3:2 or 4:3), right widths +/-15%, "
                               "spans log-uniform 64..512, crops of a 6400x4400 span-512 synthetic pair",
                   "inputs": "uint16 resident in HBM (value); host uint16 in / host float out (host_in_out)"},
    }
    print(json.dumps(out), flush=True)
    m.close()
    return 0


def bench_serve(args) -> int:
    """Poisson request streams through the serving loop (left views cached,
    right images uploaded per request), against one-at-a-time r3dg_match calls
    of the same requests in arrival order."""
    import threading

    from regard3d_mvs_amd import StereoOptions, workloads
    from regard3d_mvs_amd.native import Matcher

    src = workloads.source_pair()
    reqs = workloads.stream(src, args.requests)
    log(f"bench: {len(reqs)} requests, {sum(r.left.size for r in reqs) / len(reqs) / 1e6:.2f} MP mean, "
        f"{len(set(r.view for r in reqs))} views")
    o = StereoOptions()
    m = Matcher(0)
    m.match(reqs[0].left, reqs[0].right, None, None, reqs[0].dmin, reqs[0].nd, o)  # warm-up
    service, plain = [], []
    for r in reqs:
        s0 = time.perf_counter()
        plain.append(m.match(r.left, r.right, None, None, r.dmin, r.nd, o).disparity)
        service.append(time.perf_counter() - s0)
    log(f"bench: plain service {sum(service):.2f} s ({len(reqs) / sum(service):.2f} pairs/s)")
    results = []
    for rate in [float(x) for x in args.rates.split(",")]:
        arrivals = workloads.poisson_arrivals(len(reqs), rate)
        m.serve_start(o, hbm_budget=int(args.hbm_budget * 1e9))
        tickets = [None] * len(reqs)
        late = [0.0] * len(reqs)  # submission after the scheduled arrival (client side)
        put = set()
        t0 = time.perf_counter()

        def client():
            for i, (a, r) in enumerate(zip(arrivals, reqs)):
                d = t0 + a - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                if r.view not in put:
                    m.view_put(r.view, r.left)
                    put.add(r.view)
                late[i] = time.perf_counter() - (t0 + a)
                tickets[i] = m.submit(r.left.shape, r.dmin, r.nd, right=r.right, left_view=r.view)

        th = threading.Thread(target=client)
        th.start()
        lat, exact, done = [], True, 0
        for i in range(len(reqs)):
            while tickets[i] is None:
                time.sleep(0.0005)
            g = m.wait(tickets[i])
            lat.append(late[i] + g.seconds)  # scheduled arrival to completion
            exact = exact and _same(g.disparity, plain[i])
            done = time.perf_counter()
        th.join()
        m.serve_stop()
        for v in put:
            m.view_evict(v)
        span = done - t0
        fl = workloads.fifo_latencies(arrivals, service)
        res = {"rate": rate, "pairs_per_s": len(reqs) / span,
               "latency_ms": {q: 1e3 * percentile(lat, int(q[1:])) for q in ("p50", "p95", "p99")},
               "plain_fifo_latency_ms": {q: 1e3 * percentile(fl, int(q[1:])) for q in ("p50", "p95", "p99")},
               "plain_pairs_per_s": len(reqs) / float(np.max(arrivals + fl)),
               "bit_exact_vs_plain": exact}
        log(f"bench: rate {rate}: {json.dumps(res)}")
        results.append(res)
    out = {"metric": "served pairs/s and p50/p95/p99 latency vs plain", "unit": "pairs/s", "n_gpus": 1,
           "value": results[-1]["pairs_per_s"], "plain_capacity_pairs_per_s": len(reqs) / sum(service),
           "higher_is_better": True, "data": "synthetic", "dtype": "u16", "rates": results,
           "config": {"workload": f"{len(reqs)} requests per rate, Poisson arrivals, 0.3-24 MP log-uniform, "
                                  "spans log-uniform 32..512, views reused by 1-4 partners, crops of a "
                                  "6400x4400 span-512 synthetic pair", "hbm_budget_gb": args.hbm_budget}}
    print(json.dumps(out), flush=True)
    m.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-rows", type=int, default=4032, help="rows of the CPU-baseline band (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", default="pair", choices=["pair", "batch", "serve", "copy", "rectify", "combine"],
                    help="pair: one pair per call (default); batch: the 160-pair scene, batched vs plain; "
                         "serve: a Poisson request stream through the serving loop vs plain; copy: one 1 GiB "
                         "device-to-device copy (calibrates the FETCH_SIZE / WRITE_SIZE counters); rectify: "
                         "rectify_image of the 24 MP image; combine: combine_pairs of 4 pairs on a 24 MP grid")
    ap.add_argument("--views", type=int, default=40, help="batch: views (4 partners each)")
    ap.add_argument("--requests", type=int, default=300, help="serve: requests per rate")
    ap.add_argument("--rates", default="5,20,50", help="serve: arrival rates (pairs/s)")
    ap.add_argument("--hbm-budget", type=float, default=200.0, help="serve: admission budget (GB)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent, see DESIGN.md for how they would shard)", file=sys.stderr)
        return 2
    if args.workload == "copy":
        import torch
        n = 1 << 30
        a = torch.empty(n, dtype=torch.uint8, device="cuda")
        a.fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": n}))
        return 0
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "rectify":
        return bench_rectify(args)
    if args.workload == "combine":
        return bench_combine(args)

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
                # The 8 scans of a level run on two streams (more on small
                # levels), so launches overlap: this is the aggregation's wall
                # time per launch, the union of the launches' intervals.
                "avg_launch_ms": 1e3 * agg_s / max(launches, 1),
                "launch_timing": "aggregation wall time / launches (two streams)"}

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
