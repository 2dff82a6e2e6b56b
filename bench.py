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


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads(args) -> int:
    """All host cores of this process's CPU share (at most 16: the share of
    one GPU on the GPU hosts)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return args.cpu_threads or min(16, n)


def best_of(runs, fn):
    best, res = float("inf"), None
    for _ in range(runs):
        t0 = time.perf_counter()
        res = fn()
        best = min(best, time.perf_counter() - t0)
    return best, res


def cpu_baseline(pair, options, cells_per_pair, args):
    """The oracle (the CPU restatement of the reference matcher) timed on this
    host, best of `cpu_runs`: on all cores over a band of the pair's rows, and
    on one core over a square crop; both keep the pair's level structure, and
    rates are scaled to whole pairs by aggregated cells."""
    from oracle import oracle as O
    h, w = pair.left.shape
    threads = cpu_threads(args)
    rows = min(h, args.cpu_rows)
    t_all, r_all = best_of(args.cpu_runs, lambda: O.match_stereo(
        pair.left[:rows], pair.right[:rows], None, None, pair.dmin, pair.nd, options, threads=threads))
    side = min(h, w, args.cpu_crop)
    cl = np.ascontiguousarray(pair.left[:side, :side])
    cr = np.ascontiguousarray(pair.right[:side, :side])
    t_one, r_one = best_of(args.cpu_runs, lambda: O.match_stereo(cl, cr, None, None, pair.dmin, pair.nd, options,
                                                                  threads=1))
    v_all = r_all.cells / t_all / cells_per_pair
    v_one = r_one.cells / t_one / cells_per_pair
    log(f"bench: CPU baseline {v_all:.4f} pairs/s on {threads} threads, {v_one:.5f} on 1")
    return {"value": v_all, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"rows 0..{rows - 1} of the {w}x{h} pair ({r_all.levels + 1} levels, {r_all.cells} cells) on "
                      f"{threads} threads, best of {args.cpu_runs}: {t_all:.2f} s; scaled to whole pairs by cells "
                      f"({cells_per_pair} per pair)",
            "one_core": {"value": v_one, "unit": "pairs/s", "cores": 1,
                         "sample": f"the {side}x{side} top-left crop ({r_one.levels + 1} levels, {r_one.cells} cells), "
                                   f"best of {args.cpu_runs}: {t_one:.2f} s; scaled by cells"},
            "cpu_model": cpu_model(),
            "host_cpus_visible": os.cpu_count()}


# The reference's own published CPU figures for its bench_stereo default
# (1024x768, 128 disparities; bench/README.md:23-30, clang-cl build, 16-core
# desktop): Mcell/s of full-range cells.
PUBLISHED_MCELLS = {"full_1_thread": 150.0, "full_32_threads": 363.0, "hier_1_thread": 277.0,
                    "hier_32_threads": 515.0}


def calibration(m, args):
    """bench_stereo's default case (1024x768x128, full range and automatic
    levels) through the oracle on 1 and all threads and through the GPU
    matcher, in Mcell/s of full-range cells (w * h * D / s), beside the
    reference's published figures."""
    from oracle import oracle as O
    from regard3d_mvs_amd import StereoOptions, synth
    w, h, nd = 1024, 768, 128
    left, right = synth.bench_stereo_pair(w, h, nd)
    full = w * h * nd
    threads = cpu_threads(args)
    out = {"input": "bench_stereo.cpp default: 1024x768, 128 disparities, multi-scale noise, smooth shift",
           "unit": "Mcell/s (full-range cells w*h*D per second)", "threads": threads,
           "published_reference_clang_cl": PUBLISHED_MCELLS}
    for name, levels in (("full", 0), ("hier", -1)):
        o = StereoOptions(levels=levels)
        t1, r1 = best_of(2, lambda: O.match_stereo(left, right, None, None, 0, nd, o, threads=1))
        tn, _ = best_of(3, lambda: O.match_stereo(left, right, None, None, 0, nd, o, threads=threads))
        m.match(left, right, None, None, 0, nd, o)
        tg, g = best_of(5, lambda: m.match(left, right, None, None, 0, nd, o))
        exact = bool(np.array_equal(g.disparity.view(np.uint32), r1.disparity.view(np.uint32)))
        out[name] = {"levels": r1.levels, "cells_per_pixel": r1.cells / (w * h),
                     "oracle_1_thread": full / t1 / 1e6, f"oracle_{threads}_threads": full / tn / 1e6,
                     "gpu_host_io": full / tg / 1e6, "gpu_bit_exact": exact}
        log(f"bench: calibration {name}: {json.dumps(out[name])}")
    return out


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
    reqs = workloads.scene(views=args.views, workers=args.workers)
    log(f"bench: {len(reqs)} pairs generated in {time.perf_counter() - t:.1f} s; "
        f"{sum(r.left8.size for r in reqs) / len(reqs) / 1e6:.1f} MP mean")
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
                               "spans log-uniform 64..512, every pair from the single-pair generator at its own span",
                   "inputs": "uint16 resident in HBM (value); host uint16 in / host float out (host_in_out)"},
    }
    print(json.dumps(out), flush=True)
    m.close()
    return 0


def bench_serve(args) -> int:
    """Poisson request streams through the serving loop (left views cached,
    right images uploaded per request), against one-at-a-time r3dg_match calls
    of the same requests in arrival order."""
    import concurrent.futures
    import threading

    from regard3d_mvs_amd import StereoOptions, workloads
    from regard3d_mvs_amd.native import Matcher

    import hashlib

    def digest(a):
        return hashlib.blake2b(np.ascontiguousarray(a).view(np.uint8), digest_size=16).digest()

    t = time.perf_counter()
    reqs = workloads.stream(args.requests, workers=args.workers)
    log(f"bench: {len(reqs)} requests generated in {time.perf_counter() - t:.1f} s, "
        f"{sum(r.left8.size for r in reqs) / len(reqs) / 1e6:.2f} MP mean, {len(set(r.view for r in reqs))} views")
    o = StereoOptions()
    m = Matcher(0)
    m.match(reqs[0].left8, reqs[0].right8, None, None, reqs[0].dmin, reqs[0].nd, o)  # warm-up
    # Service times of the match calls alone; the digests that later check
    # the served maps are taken outside the timed calls.
    service, plain = [], []
    for i, r in enumerate(reqs):
        s0 = time.perf_counter()
        g = m.match(r.left8, r.right8, None, None, r.dmin, r.nd, o)
        service.append(time.perf_counter() - s0)
        plain.append(digest(g.disparity))
        if (i + 1) % 250 == 0:
            log(f"bench: plain {i + 1}/{len(reqs)}")
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
                    m.view_put(r.view, r.left8)
                    put.add(r.view)
                late[i] = time.perf_counter() - (t0 + a)
                tickets[i] = m.submit(r.left8.shape, r.dmin, r.nd, right=r.right8, left_view=r.view)

        th = threading.Thread(target=client)
        th.start()
        lat, done = [], 0
        seen = None
        # The served maps are hashed by worker threads (hashlib releases the
        # GIL), so that checking them does not hold up receiving the next.
        checks = []
        pool = concurrent.futures.ThreadPoolExecutor(max_workers=8)
        for i in range(len(reqs)):
            while tickets[i] is None:
                time.sleep(0.0005)
            while (g := m.wait(tickets[i], 5.0)) is None:
                # A line only when the loop advanced (steps or completions).
                st = m.server_stats()
                if (st["steps"], st["completed"]) != seen:
                    seen = (st["steps"], st["completed"])
                    log(f"bench: rate {rate}: waiting for request {i}; loop {json.dumps(st)}")
            lat.append(late[i] + g.seconds)  # scheduled arrival to completion
            checks.append(pool.submit(digest, g.disparity))
            done = time.perf_counter()
            if (i + 1) % 250 == 0:
                log(f"bench: rate {rate}: {i + 1}/{len(reqs)} served")
        th.join()
        exact = all(c.result() == plain[i] for i, c in enumerate(checks))
        pool.shutdown()
        loop_stats = m.server_stats()
        m.serve_stop()
        for v in put:
            m.view_evict(v)
        span = done - t0
        fl = workloads.fifo_latencies(arrivals, service)
        res = {"rate": rate, "pairs_per_s": len(reqs) / span,
               "latency_ms": {q: 1e3 * percentile(lat, int(q[1:])) for q in ("p50", "p95", "p99")},
               "plain_fifo_latency_ms": {q: 1e3 * percentile(fl, int(q[1:])) for q in ("p50", "p95", "p99")},
               "plain_pairs_per_s": len(reqs) / float(np.max(arrivals + fl)),
               "bit_exact_vs_plain": exact, "loop": loop_stats}
        log(f"bench: rate {rate}: {json.dumps(res)}")
        results.append(res)
    out = {"metric": "served pairs/s and p50/p95/p99 latency vs plain", "unit": "pairs/s", "n_gpus": 1,
           "value": results[-1]["pairs_per_s"], "plain_capacity_pairs_per_s": len(reqs) / sum(service),
           "higher_is_better": True, "data": "synthetic", "dtype": "u16", "rates": results,
           "config": {"workload": f"{len(reqs)} requests per rate, Poisson arrivals, 0.3-24 MP log-uniform, "
                                  "spans log-uniform 32..512, each view reused by 1-4 partners, every pair from "
                                  "the single-pair generator at its own span", "hbm_budget_gb": args.hbm_budget,
                      "inputs": "host uint8 images in (views put once, right images per request), host float "
                                "maps out; latency from scheduled arrival to result in host memory"}}
    print(json.dumps(out), flush=True)
    m.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-rows", type=int, default=1536, help="rows of the all-core CPU-baseline band")
    ap.add_argument("--cpu-crop", type=int, default=1536, help="side of the one-core CPU-baseline crop")
    ap.add_argument("--cpu-runs", type=int, default=3, help="best-of runs of the CPU baseline (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-calibrate", dest="calibrate", action="store_false",
                    help="skip the 1024x768x128 calibration against the reference's published CPU figures")
    ap.add_argument("--workload", default="pair", choices=["pair", "batch", "serve", "copy", "rectify", "combine"],
                    help="pair: one pair per call (default); batch: the 160-pair scene, batched vs plain; "
                         "serve: a Poisson request stream through the serving loop vs plain; copy: one 1 GiB "
                         "device-to-device copy (calibrates the FETCH_SIZE / WRITE_SIZE counters); rectify: "
                         "rectify_image of the 24 MP image; combine: combine_pairs of 4 pairs on a 24 MP grid")
    ap.add_argument("--views", type=int, default=40, help="batch: views (4 partners each)")
    ap.add_argument("--requests", type=int, default=2000, help="serve: requests per rate")
    ap.add_argument("--rates", default="5,20,50,200", help="serve: arrival rates (pairs/s; 200 saturates)")
    ap.add_argument("--workers", type=int, default=16, help="batch / serve: processes generating the images")
    ap.add_argument("--hbm-budget", type=float, default=200.0, help="serve: admission budget (GB)")
    args = ap.parse_args()
    if args.cpu_rows <= 0:
        args.cpu_runs = 0
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (pairs are independent: "
              "they would shard across GPUs, one process each, with pairs sharing a view on one GPU)",
              file=sys.stderr)
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

    return bench_pair(args)


def bench_pair(args) -> int:
    """One pair per call (the plain form).  value: images already resident in
    HBM when the timed region starts, maps left in HBM, pairs per wall second.
    Beside it: the same pair with host uint8 images in and host float maps out
    (the PCIe-inclusive rate, and the per-pair latency of that form), and a
    profiled pass for the per-stage device times and the roofline."""
    import torch

    from regard3d_mvs_amd import StereoOptions, synth
    from regard3d_mvs_amd.native import Matcher, host_empty

    t = time.perf_counter()
    pair = synth.config_pair(args.config)
    h, w = pair.left.shape
    log(f"bench: generated {WORKLOADS[args.config]} in {time.perf_counter() - t:.1f} s")
    options = StereoOptions(levels=0) if args.config == 0 else StereoOptions()
    steps = args.steps
    m = Matcher(0)

    # Images resident in HBM, maps left there: the timed region of `value`.
    dev = torch.device("cuda:0")
    dleft = torch.from_numpy(pair.left.view(np.int16)).to(dev).view(torch.uint16)
    dright = torch.from_numpy(pair.right.view(np.int16)).to(dev).view(torch.uint16)
    ddisp = torch.empty((h, w), dtype=torch.float32, device=dev)
    dconf = torch.empty((h, w), dtype=torch.float32, device=dev)
    for i in range(max(args.warmup, 1)):
        r = m.match(dleft, dright, None, None, pair.dmin, pair.nd, options, out=(ddisp, dconf))
        log(f"bench: warmup {i + 1}/{max(args.warmup, 1)}: {r.seconds * 1e3:.1f} ms")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.match(dleft, dright, None, None, pair.dmin, pair.nd, options, out=(ddisp, dconf))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pairs_s = steps / elapsed
    log(f"bench: {steps} pairs resident in HBM in {elapsed:.3f} s ({pairs_s:.2f} pairs/s)")

    # Profiled pass (HIP events around the stages): per-stage device time and
    # the aggregation time of the roofline.
    m.set_profiling(True)
    agg_s, agg_b, launches = 0.0, 0, 0
    stages = {}
    for _ in range(steps):
        r = m.match(dleft, dright, None, None, pair.dmin, pair.nd, options, out=(ddisp, dconf))
        agg_s += r.stage_seconds["aggregate"]
        agg_b += r.aggregate_bytes
        launches += r.aggregate_launches
        for k, v in r.stage_seconds.items():
            stages[k] = stages.get(k, 0.0) + v
    m.set_profiling(False)
    cells, tile_px = r.cells, r.tile_pixels
    log("bench: stage ms/pair: " + ", ".join(f"{k} {1e3 * v / steps:.2f}" for k, v in stages.items()))

    # Host buffers: the uint8 images (u16 = u8 * 257 is formed on the device)
    # and the float maps, in page-locked memory of the library's runtime.
    left8, k1 = host_empty((h, w), np.uint8)
    right8, k2 = host_empty((h, pair.right.shape[1]), np.uint8)
    left8[...] = (pair.left // 257).astype(np.uint8)
    right8[...] = (pair.right // 257).astype(np.uint8)
    hdisp, k3 = host_empty((h, w), np.float32)
    hconf, k4 = host_empty((h, w), np.float32)
    m.match(left8, right8, None, None, pair.dmin, pair.nd, options, out=(hdisp, hconf))
    lat = []
    h0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        m.match(left8, right8, None, None, pair.dmin, pair.nd, options, out=(hdisp, hconf))
        lat.append(time.perf_counter() - s0)
    host_elapsed = time.perf_counter() - h0
    valid = float(np.isfinite(hdisp).mean())
    same = bool(np.array_equal(ddisp.cpu().numpy().view(np.uint32), hdisp.view(np.uint32)) and
                np.array_equal(dconf.cpu().numpy().view(np.uint32), hconf.view(np.uint32)))
    log(f"bench: {steps} pairs host u8 in -> host f32 out in {host_elapsed:.3f} s; {cells} cells "
        f"({cells / (w * h):.1f}/px), {tile_px} tile px, valid {valid:.3f}; maps equal to the HBM run: {same}")

    # Roofline of the aggregation: the path scans' algorithmic bytes per
    # matching direction and level, 30 C + 176 P (C cells, P tile pixels),
    # over the event-timed wall time of the scans.
    alg = 30 * cells + 176 * tile_px
    per_launch = alg / max(r.aggregate_launches, 1)
    achieved = alg * steps / agg_s if agg_s > 0 else 0.0
    own = agg_b / agg_s if agg_s > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": measured_traffic(args.config),
                "kernel": "k_path (8 SGM path scans per level and chunk)",
                "algorithmic_bytes": "30 C + 176 P per pair (C cells, P tile pixels, all levels and both directions)",
                "algorithmic_bytes_per_launch": per_launch,
                # The scans of a level overlap on two or more streams, so a
                # launch's share is the aggregation's wall time per launch,
                # the union of the launches' intervals.
                "avg_launch_ms": 1e3 * agg_s / max(launches, 1),
                "launch_timing": "HIP events around the path launches of each chunk on the library's stream "
                                 "(which waits for the side streams); wall time / launches",
                "own_model": {"bytes_per_launch": agg_b / max(launches, 1), "achieved": own / 1e9,
                              "frac": own / HBM_PEAK,
                              "formula": "(40 - 2 w) C + 80 P: C read by 8 scans, S written by 8 and read by "
                                         "8 - w (w streams), record and intensity per pixel and scan"}}

    cpu = None
    if args.cpu_runs > 0:
        cpu = cpu_baseline(pair, options, cells, args)
    calib = calibration(m, args) if args.calibrate else None

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
                   "inputs": "uint16 images resident in HBM when the timed region starts, maps left in HBM",
                   "seed": pair.seed},
        "gcell_path_per_s": cells * 8 * steps / elapsed / 1e9,
        "cells_per_pair": cells,
        "tile_pixels_per_pair": tile_px,
        "latency_ms": {"p50": 1e3 * percentile(lat, 50), "p95": 1e3 * percentile(lat, 95),
                       "p99": 1e3 * percentile(lat, 99),
                       "form": "host uint8 in -> host float out, per call"},
        "host_in_out": {"pairs_per_s": steps / host_elapsed, "ms_per_pair": 1e3 * host_elapsed / steps,
                        "gcell_path_per_s": cells * 8 * steps / host_elapsed / 1e9,
                        "inputs": "host uint8 images in page-locked memory, widened to u8 * 257 on the device; "
                                  "host float disparity and confidence out (PCIe-inclusive)",
                        "maps_equal_hbm_run": same},
        "stage_ms_per_pair": {k: 1e3 * v / steps for k, v in stages.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "calibration": calib,
    }
    print(json.dumps(out), flush=True)
    m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
