"""Mixed-size pair streams for the batched and served forms of the matcher.

Pairs are crops of one large synthetic rectified pair (synth.make_pair, span
512, slopes <= 0.3, 30% low-texture area): a crop keeps the same rows and the
same first column in both images, so its disparities are those of the source.
Sizes, aspect ratios, right-width jitter and disparity spans follow the
scene (batch) and stream (serve) workloads; each left view is shared by up to
four partners (its right images differ in width).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import synth

SOURCE_W, SOURCE_H = 6400, 4400  # holds 24 MP at 3:2 and at 4:3


@dataclasses.dataclass
class Request:
    left: np.ndarray    # uint16 view into the source
    right: np.ndarray
    dmin: int
    nd: int
    view: int           # id of the left view (shared by its partners)


def source_pair(seed: int = 31337) -> synth.Pair:
    return synth.make_pair(SOURCE_W, SOURCE_H, 512, seed, slope=0.3, low_texture=0.3)


def _crop_size(rng, mp: float) -> tuple[int, int]:
    aspect = 1.5 if rng.random() < 0.5 else 4.0 / 3.0
    h = int(math.sqrt(mp * 1e6 / aspect))
    w = int(h * aspect)
    return min(max(w, 64), SOURCE_W), min(max(h, 48), SOURCE_H)


def make_requests(src: synth.Pair, views: int, partners, mp_range, span_range, seed: int, log_sizes: bool):
    """`partners(rng)` gives the number of pairs of a view; sizes are uniform
    (log-uniform with log_sizes) in mp_range megapixels, spans log-uniform in
    span_range."""
    rng = np.random.default_rng(seed)
    lo, hi = mp_range
    out = []
    for v in range(views):
        mp = math.exp(rng.uniform(math.log(lo), math.log(hi))) if log_sizes else rng.uniform(lo, hi)
        w, h = _crop_size(rng, mp)
        y0 = int(rng.integers(0, SOURCE_H - h + 1))
        x0 = int(rng.integers(0, SOURCE_W - w + 1))
        left = src.left[y0:y0 + h, x0:x0 + w]
        for _ in range(partners(rng)):
            wr = int(w * rng.uniform(0.85, 1.15))
            wr = max(64, min(wr, SOURCE_W - x0))
            right = src.right[y0:y0 + h, x0:x0 + wr]
            nd = int(round(math.exp(rng.uniform(math.log(span_range[0]), math.log(span_range[1])))))
            out.append(Request(left, right, 0, nd, v + 1))
    return out


def scene(src: synth.Pair, views: int = 40, seed: int = 3000):
    """The batched scene: `views` views x 4 partners, 1.5-24 MP (3:2 or 4:3),
    right widths +/-15%, spans log-uniform in 64..512."""
    return make_requests(src, views, lambda rng: 4, (1.5, 24.0), (64, 512), seed, log_sizes=False)


def stream(src: synth.Pair, requests: int, seed: int = 4000):
    """The served stream: sizes log-uniform in 0.3-24 MP, spans log-uniform in
    32..512, each view reused by 1 to 4 partners, cut to `requests` pairs."""
    rng_views = max(1, requests // 2)
    reqs = make_requests(src, rng_views, lambda rng: int(rng.integers(1, 5)), (0.3, 24.0), (32, 512), seed,
                         log_sizes=True)
    while len(reqs) < requests:
        reqs += make_requests(src, rng_views, lambda rng: int(rng.integers(1, 5)), (0.3, 24.0), (32, 512),
                              seed + len(reqs), log_sizes=True)
    return reqs[:requests]


def poisson_arrivals(n: int, rate: float, seed: int = 5000) -> np.ndarray:
    """Arrival times (seconds from 0) of n requests at `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def fifo_latencies(arrivals, service) -> np.ndarray:
    """Latencies of one-at-a-time processing in arrival order."""
    t = 0.0
    out = np.empty(len(arrivals))
    for i, (a, s) in enumerate(zip(arrivals, service)):
        t = max(t, a) + s
        out[i] = t - a
    return out
