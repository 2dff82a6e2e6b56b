"""Synthetic rectified pairs for the matcher benchmarks and tests.

Left image: band-limited value noise (4 octaves of bilinearly interpolated
random lattices), mean 128, sigma 40, clipped to uint8.  Disparity: a Voronoi
partition (jittered-grid seeds) with a random plane per cell, rounded to
integers in [0, span).  Right image: forward warp x_r = x_l - d (the larger
disparity wins where two left pixels land on one right pixel), holes filled
with fresh noise, then Gaussian noise.  Optionally a fraction of the area is
low-texture (constant grey plus sigma-1 noise).  uint8 data enters the matcher
as uint16 = u8 * 257, the reference's 8-to-16-bit convention
(src/r3d/io/image_io.hpp:25-26).  Deterministic for a given seed.
"""
from __future__ import annotations

import concurrent.futures
import dataclasses
import os

import numpy as np


@dataclasses.dataclass
class Pair:
    left: np.ndarray    # uint16 (u8 * 257)
    right: np.ndarray   # uint16
    disparity: np.ndarray  # int16 ground truth of the left image (-1: none)
    dmin: int
    nd: int
    seed: int


_POOL = concurrent.futures.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1))


def _bands(h, fn):
    """fn(y0, y1) over row bands in parallel (numpy releases the GIL)."""
    n = min(16, os.cpu_count() or 1)
    edges = [h * i // n for i in range(n + 1)]
    list(_POOL.map(lambda i: fn(edges[i], edges[i + 1]), range(n)))


def _lattice_into(acc, cell, rng):
    h, w = acc.shape
    gh, gw = h // cell + 2, w // cell + 2
    g = rng.standard_normal((gh, gw)).astype(np.float32)
    xs = np.arange(w, dtype=np.float32) / cell
    ix = xs.astype(np.int64)
    fx = xs - ix
    rows = g[:, ix] * (1.0 - fx) + g[:, ix + 1] * fx  # gh x w
    ys = np.arange(h, dtype=np.float32) / cell
    iy = ys.astype(np.int64)
    fy = (ys - iy)[:, None]

    def band(y0, y1):
        acc[y0:y1] += rows[iy[y0:y1]] * (1.0 - fy[y0:y1]) + rows[iy[y0:y1] + 1] * fy[y0:y1]
    _bands(h, band)


def value_noise(h, w, rng, cells=(2, 5, 13, 32), mean=128.0, sigma=40.0):
    acc = np.zeros((h, w), np.float32)
    for c in cells:
        _lattice_into(acc, c, rng)
    acc -= acc.mean()
    acc *= sigma / max(float(acc.std()), 1e-6)
    acc += mean
    return acc


def voronoi_labels(h, w, cell, rng):
    """Nearest jittered-grid seed of every pixel; returns (labels, seeds_x, seeds_y)."""
    gh, gw = h // cell + 1, w // cell + 1
    sx = ((np.arange(gw)[None, :] + rng.random((gh, gw))) * cell).ravel()
    sy = ((np.arange(gh)[:, None] + rng.random((gh, gw))) * cell).ravel()
    lab = np.zeros((h, w), np.int64)
    xs = np.arange(w)[None, :]
    cx = np.minimum(xs // cell, gw - 1)

    def band(y0, y1):
        ys = np.arange(y0, y1)[:, None]
        cy = np.minimum(ys // cell, gh - 1)
        best = np.full((y1 - y0, w), np.inf, np.float32)
        out = lab[y0:y1]
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                idx = np.clip(cy + dy, 0, gh - 1) * gw + np.clip(cx + dx, 0, gw - 1)
                d2 = ((sx[idx] - xs) ** 2 + (sy[idx] - ys) ** 2).astype(np.float32)
                upd = d2 < best
                best[upd] = d2[upd]
                out[upd] = idx[upd]
    _bands(h, band)
    return lab, sx, sy


def make_pair(w, h, span, seed, slope=0.2, low_texture=0.0, cell=None, noise=2.0) -> Pair:
    rng = np.random.default_rng(seed)
    left = value_noise(h, w, rng)
    cell = cell or max(32, int(round(np.sqrt(w * h / 40.0))))
    lab, sx, sy = voronoi_labels(h, w, cell, rng)
    n = sx.size
    d0 = rng.uniform(0, span - 1, n)
    a = rng.uniform(-slope, slope, n)
    b = rng.uniform(-slope, slope, n)
    ys, xs = np.mgrid[0:h, 0:w]
    d = d0[lab] + a[lab] * (xs - sx[lab]) + b[lab] * (ys - sy[lab])
    disp = np.clip(np.rint(d), 0, span - 1).astype(np.int64)
    if low_texture > 0:
        lab2, _, _ = voronoi_labels(h, w, max(64, cell // 2), rng)
        order = rng.permutation(int(lab2.max()) + 1)
        area = np.bincount(lab2.ravel(), minlength=order.size)
        chosen = np.zeros(order.size, bool)
        acc = 0
        for c in order:
            if acc >= low_texture * w * h:
                break
            chosen[c] = True
            acc += area[c]
        flat = chosen[lab2]
        left = np.where(flat, 128.0 + rng.standard_normal((h, w)).astype(np.float32), left)
    left8 = np.clip(np.rint(left), 0, 255).astype(np.uint8)
    # Forward warp: for each right pixel keep the left pixel with the largest d.
    xr = xs - disp
    ok = xr >= 0
    src = np.flatnonzero(ok.ravel())
    dst = (ys * w + xr).ravel()[src]
    # Write in ascending (d, source index) order: the last write (largest d)
    # wins; numpy's stable sort of 16-bit keys is a radix sort.
    order = np.argsort(disp.ravel()[src].astype(np.int16), kind="stable")
    key = np.full(h * w, -1, np.int64)
    key[dst[order]] = src[order]
    fresh = np.clip(np.rint(value_noise(h, w, rng)), 0, 255).astype(np.uint8).ravel()
    right = np.where(key >= 0, left8.ravel()[np.maximum(key, 0)], fresh).reshape(h, w).astype(np.float32)
    right += rng.standard_normal((h, w)).astype(np.float32) * noise
    right8 = np.clip(np.rint(right), 0, 255).astype(np.uint8)
    return Pair(left8.astype(np.uint16) * 257, right8.astype(np.uint16) * 257, disp.astype(np.int16), 0, span, seed)


def config_pair(cfg: int, seed: int | None = None) -> Pair:
    """The single-pair workloads: 0 = 640x480 / 64 d, 1 = 1920x1080 / 256 d,
    2 = 6048x4032 / 512 d with 30% low-texture area."""
    s = (1000 * cfg) if seed is None else seed
    if cfg == 0:
        return make_pair(640, 480, 64, s)
    if cfg == 1:
        return make_pair(1920, 1080, 256, s)
    if cfg == 2:
        return make_pair(6048, 4032, 512, s, slope=0.3, low_texture=0.3)
    raise ValueError(f"no single-pair config {cfg}")
