"""Procedural test scenes of the reference's stereo tests, restated in numpy.

Follows tests/test_util.hpp:22-39 (Rng over std::mt19937 with portable
distributions) and tests/test_stereo.cpp:19-73 (Texture, render_pair,
random_image).  numpy's MT19937 with legacy integer seeding is std::mt19937
seeded with the same value, so the raw streams are identical.
"""
from __future__ import annotations

import numpy as np


class Rng:
    """test_util.hpp:22-39."""

    def __init__(self, seed: int):
        self._bg = np.random.MT19937()
        self._bg._legacy_seeding(int(seed))

    def u32(self, n=None):
        if n is None:
            return int(self._bg.random_raw())
        return self._bg.random_raw(n).astype(np.uint64)

    def uniform(self, lo, hi):
        return lo + (hi - lo) * ((self.u32() + 0.5) / 4294967296.0)

    def uniform_int(self, lo, hi):
        return lo + int(self.u32() % (hi - lo + 1))

    def uniform_int_array(self, lo, hi, n):
        return lo + (self.u32(n) % np.uint64(hi - lo + 1)).astype(np.int64)

    def normal(self, mean, sigma):
        u1 = self.uniform(0.0, 1.0)
        u2 = self.uniform(0.0, 1.0)
        return mean + sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


class Texture:
    """test_stereo.cpp:19-40: sum of 24 random sinusoids, about [0, 65535]."""

    def __init__(self, seed: int):
        rng = Rng(seed)
        self.fx, self.fy, self.phase, self.amp = [], [], [], []
        for i in range(24):
            m = rng.uniform(0.05, 0.9)
            self.fx.append(m * rng.uniform(-1.0, 1.0))
            m = rng.uniform(0.05, 0.9)
            self.fy.append(m * rng.uniform(-1.0, 1.0))
            self.phase.append(rng.uniform(0.0, 6.283))
            self.amp.append(1.0 / (1.0 + i * 0.2))

    def __call__(self, x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        v = np.zeros(np.broadcast(x, y).shape)
        norm = 0.0
        for fx, fy, ph, a in zip(self.fx, self.fy, self.phase, self.amp):
            v = v + a * np.sin(fx * x + fy * y + ph)
            norm += a
        return 32768.0 + 30000.0 * v / norm * 2.0


def to_u16(v):
    return np.floor(np.clip(v, 0.0, 65535.0) + 0.5).astype(np.uint16)


def render_pair(tex, w, h, disparity):
    """test_stereo.cpp:46-65: right pixel xr shows texture(xl) with xl - d(xl) = xr."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    left = to_u16(tex(x, y))
    gt = disparity(x, y).astype(np.float32) * np.ones_like(x, np.float32)
    xl = x.copy()
    for _ in range(20):
        xl = x + disparity(xl, y)
    right = to_u16(tex(xl, y))
    return left, right, gt


def random_image(w, h, seed):
    """test_stereo.cpp:67-73."""
    rng = Rng(seed)
    return rng.uniform_int_array(0, 65535, w * h).astype(np.uint16).reshape(h, w)


def evaluate_disparity(est, gt, mask=None, mask_min=255):
    """disparity_eval.cpp:11-67: density over pixels with ground truth in the
    mask; bad-k and mean |error| over the estimated ones."""
    valid_gt = np.isfinite(gt)
    if mask is not None:
        valid_gt &= mask >= mask_min
    n = int(valid_gt.sum())
    has = valid_gt & np.isfinite(est)
    err = np.abs(est[has].astype(np.float64) - gt[has].astype(np.float64))
    return {
        "num_evaluated": n,
        "density": float(has.sum()) / max(n, 1),
        "bad1": float((err > 1.0).sum()) / max(int(has.sum()), 1),
        "bad2": float((err > 2.0).sum()) / max(int(has.sum()), 1),
        "avg_error": float(err.mean()) if err.size else 0.0,
    }
