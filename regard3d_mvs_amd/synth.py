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


# ---------------------------------------------------------------- scenes
# The same generator for scenes of many pairs: one left view (value noise)
# shared by several partners, each partner a forward warp of the view by its
# own piecewise-planar disparity field, at its own span and right width.  The
# rows of a warp are independent, so the field, the warp and the noise run in
# row bands on the thread pool (each band draws from its own seeded stream).
_BLOCK = 64  # rows per band of a partner warp (the noise streams do not depend on the thread count)


def make_view(w: int, h: int, seed: int) -> np.ndarray:
    """Left view: value noise, mean 128, sigma 40, clipped to uint8."""
    rng = np.random.default_rng(seed)
    return np.clip(np.rint(value_noise(h, w, rng)), 0, 255).astype(np.uint8)


def view_fields(left8: np.ndarray, seed: int, max_width: int | None = None, cell: int | None = None):
    """The parts of a view that its partners share: the Voronoi partition of
    the view (the scene's surfaces; each partner draws its own planes) and a
    fresh-noise field of max_width columns for the holes of the warps."""
    h, w = left8.shape
    rng = np.random.default_rng(seed)
    cell = cell or max(32, int(round(np.sqrt(w * h / 40.0))))
    part = voronoi_labels(h, w, cell, rng)
    fresh = np.clip(np.rint(value_noise(h, max_width or w, rng)), 0, 255).astype(np.uint8)
    return part, fresh


def make_partner(left8: np.ndarray, span: int, seed: int, width: int | None = None, slope: float = 0.2,
                 noise: float = 2.0, fields=None):
    """Right image (uint8, `width` columns, default the view's) of the view
    left8 for a Voronoi piecewise-planar integer disparity field in
    [0, span) with plane slopes <= `slope`: x_r = x_l - d, the larger d wins,
    holes (and columns past the view) filled with fresh noise, plus Gaussian
    noise.  `fields` (view_fields) shares the partition and the fill noise
    between the partners of a view.  Returns (right8, disparity int16)."""
    h, w = left8.shape
    wr = w if width is None else int(width)
    rng = np.random.default_rng(seed)
    if fields is None or fields[1].shape[1] < wr:
        fields = view_fields(left8, int(rng.integers(0, 2 ** 63)), wr)
    (lab, sx, sy), fresh = fields
    n = sx.size
    d0 = rng.uniform(0, span - 1, n).astype(np.float32)
    a = rng.uniform(-slope, slope, n).astype(np.float32)
    b = rng.uniform(-slope, slope, n).astype(np.float32)
    sxf, syf = sx.astype(np.float32), sy.astype(np.float32)
    disp = np.empty((h, w), np.int16)
    right = np.empty((h, wr), np.uint8)
    block_seed = int(rng.integers(0, 2 ** 63))

    def band(k):
        y0, y1 = k * _BLOCK, min(h, (k + 1) * _BLOCK)
        L = lab[y0:y1]
        xs = np.arange(w, dtype=np.float32)[None, :]
        ys = np.arange(y0, y1, dtype=np.float32)[:, None]
        d = d0[L] + a[L] * (xs - sxf[L]) + b[L] * (ys - syf[L])
        di = np.clip(np.rint(d), 0, span - 1).astype(np.int16)
        disp[y0:y1] = di
        rows = y1 - y0
        xr = np.arange(w)[None, :] - di
        ok = (xr >= 0) & (xr < wr)
        src = np.flatnonzero(ok.ravel())
        dst = ((np.arange(rows)[:, None] * wr) + xr).ravel()[src]
        order = np.argsort(di.ravel()[src], kind="stable")  # ascending d: the last write wins
        key = np.full(rows * wr, -1, np.int64)
        key[dst[order]] = src[order]
        lv = left8[y0:y1].ravel()
        out = np.where(key >= 0, lv[np.maximum(key, 0)], fresh[y0:y1, :wr].ravel()).astype(np.float32)
        g = np.random.default_rng([block_seed, k])
        out += g.standard_normal(out.size, dtype=np.float32) * noise
        right[y0:y1] = np.clip(np.rint(out), 0, 255).astype(np.uint8).reshape(rows, wr)

    list(_POOL.map(band, range((h + _BLOCK - 1) // _BLOCK)))
    return right, disp


# ---------------------------------------------------------------- bench_stereo input
def _mt_uniform_pm1(bg, n):
    """n draws of uniform_real_distribution<float>(-1, 1) over std::mt19937
    as libstdc++ computes them (generate_canonical<float, 24> from one 32-bit
    draw)."""
    u = bg.random_raw(n).astype(np.float64)
    r = (u.astype(np.float32) / np.float32(4294967296.0)).astype(np.float32)
    r = np.where(r >= 1.0, np.nextafter(np.float32(1.0), np.float32(0.0)), r).astype(np.float32)
    return (np.float32(-1.0) + np.float32(2.0) * r).astype(np.float32)


def bench_stereo_pair(w: int = 1024, h: int = 768, nd: int = 128, seed: int = 1):
    """The reference benchmark's input (bench/bench_stereo.cpp:31-65,
    115-122): multi-scale noise (bilinear random grids of spacing 1..64,
    32768 + 8000 * sum, clamped) and a right image that shows the left one
    shifted by d = nd/8 + (nd/2)(x + y)/(w + h).  Returns (left, right) uint16."""
    bg = np.random.MT19937()
    bg._legacy_seeding(int(seed))
    acc = np.zeros((h, w), np.float32)
    ys, xs = np.arange(h), np.arange(w)
    s = 1
    while s <= 64:
        gw, gh = w // s + 2, h // s + 2
        grid = _mt_uniform_pm1(bg, gw * gh).reshape(gh, gw)
        gy, gx = ys // s, xs // s
        fy = ((ys % s) / np.float32(s)).astype(np.float32)[:, None]
        fx = ((xs % s) / np.float32(s)).astype(np.float32)[None, :]
        a = grid[gy][:, gx]
        b = grid[gy][:, gx + 1]
        c = grid[gy + 1][:, gx]
        e = grid[gy + 1][:, gx + 1]
        top = a + fx * (b - a)
        bottom = c + fx * (e - c)
        acc += top + fy * (bottom - top)
        s *= 2
    left = np.clip(np.float32(32768.0) + acc * np.float32(8000.0), 0, 65535).astype(np.uint16)
    yy, xx = np.mgrid[0:h, 0:w]
    d = nd // 8 + (nd // 2) * (xx + yy) // (w + h)
    right = left[yy, np.minimum(w - 1, xx + d)]
    return left, right
