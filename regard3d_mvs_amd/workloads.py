"""Mixed-size pair streams for the batched and served forms of the matcher.

Every view is its own synthetic left image (synth.make_view: value noise,
mean 128, sigma 40) and each of its partners is a forward warp of it by its
own piecewise-planar disparity field at the pair's own span
(synth.make_partner: Voronoi planes with slopes <= 0.2, holes filled with
fresh noise, N(0, 2) noise) -- the single-pair generator per pair.  Partners
of a view share the view's Voronoi partition (its surfaces) and draw their
own planes.  Sizes, aspect ratios (3:2 or 4:3), right-width jitter (+/-15%)
and disparity spans follow the batched-scene and served-stream workloads.
Images are kept as uint8; `left` / `right` give the matcher's uint16
(u8 * 257, src/r3d/io/image_io.hpp:25-26).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import synth


@dataclasses.dataclass
class Request:
    left8: np.ndarray   # uint8 left view (shared by the view's partners)
    right8: np.ndarray  # uint8 partner image
    dmin: int
    nd: int
    view: int           # id of the left view
    truth: np.ndarray | None = None  # int16 disparity of the left view for this partner

    @property
    def left(self) -> np.ndarray:
        return self.left8.astype(np.uint16) * np.uint16(257)

    @property
    def right(self) -> np.ndarray:
        return self.right8.astype(np.uint16) * np.uint16(257)


def _size(rng, mp: float) -> tuple[int, int]:
    aspect = 1.5 if rng.random() < 0.5 else 4.0 / 3.0
    h = int(math.sqrt(mp * 1e6 / aspect))
    w = int(h * aspect)
    return max(w, 64), max(h, 48)


def _make_view(spec):
    """Images of one view and its partners from their drawn parameters."""
    w, h, vseed, parts, keep_truth = spec
    left8 = synth.make_view(w, h, vseed)
    fields = synth.view_fields(left8, vseed + 1, int(math.ceil(w * 1.15)) + 1)
    out = []
    for wr, nd, pseed in parts:
        right8, truth = synth.make_partner(left8, nd, pseed, width=wr, fields=fields)
        out.append((right8, nd, truth if keep_truth else None))
    return left8, out


def make_requests(views: int, partners, mp_range, span_range, seed: int, log_sizes: bool, keep_truth=False,
                  view_base: int = 0, workers: int = 1):
    """`partners(rng)` gives the number of pairs of a view; sizes are uniform
    (log-uniform with log_sizes) in mp_range megapixels, spans log-uniform in
    span_range, right widths the view's +/-15%.  The parameters are drawn in
    order from one generator; the images of the views are made by `workers`
    processes (the same images for any count)."""
    rng = np.random.default_rng(seed)
    lo, hi = mp_range
    specs = []
    for _ in range(views):
        mp = math.exp(rng.uniform(math.log(lo), math.log(hi))) if log_sizes else rng.uniform(lo, hi)
        w, h = _size(rng, mp)
        vseed = int(rng.integers(0, 2 ** 62))
        parts = []
        for _ in range(partners(rng)):
            wr = max(64, int(w * rng.uniform(0.85, 1.15)))
            nd = int(round(math.exp(rng.uniform(math.log(span_range[0]), math.log(span_range[1])))))
            parts.append((wr, nd, int(rng.integers(0, 2 ** 62))))
        specs.append((w, h, vseed, parts, keep_truth))
    if workers > 1 and len(specs) > 1:
        import multiprocessing as mp_
        with mp_.get_context("spawn").Pool(min(workers, len(specs))) as pool:
            made = pool.map(_make_view, specs, chunksize=1)
    else:
        made = [_make_view(sp) for sp in specs]
    out = []
    for v, (left8, parts) in enumerate(made):
        for right8, nd, truth in parts:
            out.append(Request(left8, right8, 0, nd, view_base + v + 1, truth))
    return out


def scene(views: int = 40, mp_range=(1.5, 24.0), span_range=(64, 512), seed: int = 3000, workers: int = 1):
    """The batched scene: `views` views x 4 partners, sizes uniform in
    mp_range (1.5-24 MP), spans log-uniform in span_range (64..512)."""
    return make_requests(views, lambda rng: 4, mp_range, span_range, seed, log_sizes=False, workers=workers)


def stream(requests: int, seed: int = 4000, mp_range=(0.3, 24.0), span_range=(32, 512), workers: int = 1):
    """The served stream: sizes log-uniform in 0.3-24 MP, spans log-uniform in
    32..512, each view reused by 1 to 4 partners, cut to `requests` pairs."""
    reqs = []
    k = 0
    while len(reqs) < requests:
        n = max(1, (requests - len(reqs) + 1) // 2)
        reqs += make_requests(n, lambda rng: int(rng.integers(1, 5)), mp_range, span_range, seed + k,
                              log_sizes=True, view_base=1 + max((r.view for r in reqs), default=0), workers=workers)
        k += 1
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
