"""Pins the CPU oracle's matcher: ranges, filters and end-to-end behaviour.

Restates tests/test_stereo.cpp:191-469 of the reference (its thresholds),
plus brute-force definitions of the filters.
"""
from collections import deque

import numpy as np
import pytest

from oracle import oracle as O
from refscenes import Rng, Texture, evaluate_disparity, render_pair, to_u16
from regard3d_mvs_amd import StereoOptions, SubpixelMethod

INF = np.float32(np.inf)


# ---- ranges_from_coarse KATs (test_stereo.cpp:343-400)
def test_ranges_from_coarse_kats():
    wc, hc = 40, 20
    coarse = np.where(np.arange(wc)[None, :] < wc // 2, 10.0, 30.0).astype(np.float32) * np.ones((hc, 1), np.float32)
    coarse[5, 5] = INF
    coarse[:, 30:] = INF
    w, h, dmin, nd = 2 * wc, 2 * hc, 4, 96
    valid = np.full((h, w), 255, np.uint8)
    valid[1, 1] = 0
    o = StereoOptions(range_radius=1, range_margin=3, range_fill=4)
    lo, num = O.ranges_from_coarse(coarse, w, h, valid, None, 1000, 1, dmin, nd, o)
    lo = lo.astype(int)
    num = num.astype(int)

    def covers(x, y, a, b):
        return lo[y, x] <= a and lo[y, x] + num[y, x] - 1 >= b

    assert (num % 16 == 0).all() and (lo >= 0).all() and (lo + num <= nd).all()
    assert covers(10, 30, 20 - dmin - 3, 20 - dmin + 3) and num[30, 10] == 16
    assert covers(10, 10, 20 - dmin - 3, 20 - dmin + 3) and num[10, 10] == 16
    assert covers(39, 20, 20 - dmin - 3, 60 - dmin + 3)
    assert num[1, 1] == 16
    assert covers(62, 20, 60 - dmin - 3, 60 - dmin + 3) and num[20, 62] == 16
    assert lo[20, 78] == 0 and num[20, 78] == nd
    lo2, num2 = O.ranges_from_coarse(coarse, w, h, valid, None, w, -1, dmin, nd, o)
    lo2 = lo2.astype(int)
    num2 = num2.astype(int)
    assert num2[30, 10] == 16
    assert lo2[30, 30] <= 20 - dmin - 3 and lo2[30, 30] + num2[30, 30] - 1 >= 20 - dmin + 3
    assert lo2[20, 78] + num2[20, 78] - 1 < 78 - dmin + 16


def test_ranges_mirrored_is_mirror():
    rng = np.random.default_rng(3)
    coarse = rng.uniform(0, 30, size=(12, 17)).astype(np.float32)
    coarse[rng.random(coarse.shape) < 0.2] = INF
    a = O.ranges_from_coarse(coarse, 33, 24, None, None, 40, 1, -3, 80)
    b = O.ranges_from_coarse(coarse, 33, 24, None, None, 40, 1, -3, 80, mirrored=True)
    np.testing.assert_array_equal(a[0][:, ::-1], b[0])
    np.testing.assert_array_equal(a[1][:, ::-1], b[1])


# ---- filters (test_stereo.cpp:451-469 + definitions)
def test_speckle_and_median_kats():
    d = np.full((10, 20), 5.0, np.float32)
    d[3, 3] = 20.0
    d[5, 10] = INF
    d[0, 15:20] = 30.0
    s = O.remove_speckles(d, 6, 1.0)
    assert not np.isfinite(s[3, 3]) and not np.isfinite(s[0, 17]) and s[0, 0] == 5.0
    s = O.remove_speckles(d, 5, 1.0)
    assert s[0, 17] == 30.0
    m = O.median_filter_3x3(d)
    assert m[3, 3] == 5.0 and not np.isfinite(m[5, 10]) and m[5, 11] == 5.0


def bfs_speckles(d, min_size, max_diff):
    h, w = d.shape
    out = d.copy()
    seen = np.zeros(d.shape, bool)
    for sy in range(h):
        for sx in range(w):
            if seen[sy, sx] or not np.isfinite(d[sy, sx]):
                continue
            comp = [(sy, sx)]
            seen[sy, sx] = True
            q = deque(comp)
            while q:
                y, x = q.popleft()
                for ny, nx in ((y, x - 1), (y, x + 1), (y - 1, x), (y + 1, x)):
                    if 0 <= ny < h and 0 <= nx < w and not seen[ny, nx] and abs(d[ny, nx] - d[y, x]) <= max_diff:
                        seen[ny, nx] = True
                        comp.append((ny, nx))
                        q.append((ny, nx))
            if len(comp) < min_size:
                for y, x in comp:
                    out[y, x] = INF
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_speckles_are_connected_components(seed):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, 4, size=(23, 31)).astype(np.float32) * 0.75
    d[rng.random(d.shape) < 0.15] = INF
    np.testing.assert_array_equal(O.remove_speckles(d, 7, 1.0), bfs_speckles(d, 7, np.float32(1.0)))


def test_median_is_window_median():
    rng = np.random.default_rng(4)
    d = rng.uniform(0, 10, size=(13, 17)).astype(np.float32)
    d[rng.random(d.shape) < 0.2] = INF
    m = O.median_filter_3x3(d)
    h, w = d.shape
    for y in range(h):
        for x in range(w):
            if not np.isfinite(d[y, x]):
                assert not np.isfinite(m[y, x])
                continue
            v = sorted(float(d[yy, xx]) for yy in range(max(0, y - 1), min(h, y + 2))
                       for xx in range(max(0, x - 1), min(w, x + 2)) if np.isfinite(d[yy, xx]))
            n = len(v)
            e = np.float32(v[n // 2]) if n % 2 else np.float32(0.5) * (np.float32(v[n // 2]) + np.float32(v[n // 2 - 1]))
            assert m[y, x] == e, (y, x)


def test_automatic_levels():
    assert O.automatic_levels(640, 480, 64) == 1
    assert O.automatic_levels(1920, 1080, 256) == 3
    assert O.automatic_levels(6048, 4032, 512) == 4
    assert O.automatic_levels(1024, 768, 128) == 2
    assert O.automatic_levels(150, 60, 32) == 0


# ---- end-to-end behaviour (test_stereo.cpp:191-449)
@pytest.mark.parametrize("method", [SubpixelMethod.parabola, SubpixelMethod.equiangular])
def test_slanted_plane_subpixel(method):
    left, right, gt = render_pair(Texture(11), 200, 120, lambda x, y: 12.3 + 0.04 * x + 0.02 * y)
    r = O.match_stereo(left, right, None, None, 0, 32, StereoOptions(subpixel=method))
    assert r.num_disparities == 32
    mask = np.zeros((120, 200), np.uint8)
    mask[:, 30:] = 255
    s = evaluate_disparity(r.disparity, gt, mask)
    assert s["density"] > 0.95 and s["bad1"] < 0.01 and s["avg_error"] < 0.2


def test_tiles_and_threads_consistent():
    left, right, gt = render_pair(Texture(5), 150, 90, lambda x, y: np.where(x < 70, 6.0, 14.5))
    o = StereoOptions(tile_size=48, tile_margin=8, tile_blend=8)
    ref = O.match_stereo(left, right, None, None, 0, 32, o, threads=8)
    single = O.match_stereo(left, right, None, None, 0, 32, o, threads=1)
    np.testing.assert_array_equal(ref.disparity, single.disparity)
    o.tile_size = 1024
    whole = O.match_stereo(left, right, None, None, 0, 32, o)
    mask = np.zeros((90, 150), np.uint8)
    xs = np.arange(150)
    mask[:, (xs >= 20) & (np.abs(xs - 70) > 12)] = 255
    for d in (ref.disparity, whole.disparity):
        s = evaluate_disparity(d, gt, mask)
        assert s["density"] > 0.95 and s["bad1"] < 0.01


def test_masks():
    left, right, _ = render_pair(Texture(8), 120, 60, lambda x, y: 10.0 + 0 * x)
    lm = np.full((60, 120), 255, np.uint8)
    rm = np.full((60, 120), 255, np.uint8)
    lm[:, 100:] = 0
    rm[:, :40] = 0
    r = O.match_stereo(left, right, lm, rm, 0, 32)
    assert not np.isfinite(r.disparity[:, 100:]).any()
    assert not np.isfinite(r.disparity[:, :40]).any()
    good = (np.abs(r.disparity[:, 55:95] - 10.0) < 0.25).sum()
    assert good > 0.95 * 60 * 40


@pytest.mark.parametrize("full", [False, True])
def test_masked_left_does_not_disturb_lr(full):
    left, right, _ = render_pair(Texture(21), 160, 60, lambda x, y: 10.0 + 0 * x)
    rng = Rng(4)
    left = left.astype(np.float64)
    right = right.astype(np.float64)
    for y in range(60):
        for x in range(160):
            left[y, x] += rng.normal(0.0, 3000.0)
            right[y, x] += rng.normal(0.0, 3000.0)
    left, right = to_u16(left), to_u16(right)
    yy, xx = np.mgrid[0:60, 0:160]
    lm = np.where(xx < 70 + 2 * yy, 255, 0).astype(np.uint8)
    valid = int(((xx < 70 + 2 * yy) & (xx >= 20)).sum())
    r = O.match_stereo(left, right, lm, None, 0, 256, StereoOptions(lr_full_match=full, tile_size=64))
    sub = (lm[:, 20:] != 0) & (np.abs(r.disparity[:, 20:] - 10.0) < 0.5)
    assert sub.sum() > 0.75 * valid


@pytest.mark.parametrize("levels", [0, 1])
@pytest.mark.parametrize("full", [False, True])
def test_lr_unequal_widths_negative_disparity(levels, full):
    tex = Texture(23)
    wl, wr, h, d = 150, 190, 60, -7.0
    yy, xl = np.mgrid[0:h, 0:wl].astype(np.float64)
    _, xr = np.mgrid[0:h, 0:wr].astype(np.float64)
    left = to_u16(tex(xl, yy))
    right = to_u16(tex(xr + d, np.mgrid[0:h, 0:wr][0].astype(np.float64)))
    r = O.match_stereo(left, right, None, None, -20, 32, StereoOptions(lr_full_match=full, levels=levels))
    assert r.levels == levels
    assert (np.abs(r.disparity - d) < 0.25).sum() > 0.9 * wl * h


def hierarchical_scene():
    tex, hidden = Texture(29), Texture(31)
    w, h, step = 480, 300, 300
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    d_fg = lambda x, y: 150.0 + 0.1 * y - 0.05 * x  # noqa: E731
    d_bg = lambda x, y: 40.0 + 0.02 * x  # noqa: E731
    left = to_u16(tex(x, y))
    gt = np.where(x < step, d_fg(x, y), d_bg(x, y)).astype(np.float32)
    xf = x.copy()
    for _ in range(30):
        xf = x + d_fg(xf, y)
    xb = x.copy()
    for _ in range(30):
        xb = x + d_bg(xb, y)
    right = np.where(xf < step, to_u16(tex(xf, y)), np.where(xb >= step, to_u16(tex(xb, y)), to_u16(hidden(xb, y))))
    mask = np.zeros((h, w), np.uint8)
    xs = np.arange(w)
    mask[:, (xs >= 200) & (np.abs(xs - step) > 12)] = 255
    return left, right.astype(np.uint16), gt, mask


def test_hierarchical_equals_full_range():
    left, right, gt, mask = hierarchical_scene()
    full = O.match_stereo(left, right, None, None, 0, 256, StereoOptions(levels=0))
    hier = O.match_stereo(left, right, None, None, 0, 256, StereoOptions(levels=-1))
    assert full.levels == 0 and hier.levels >= 1
    assert hier.cells * 3 < full.cells
    sf = evaluate_disparity(full.disparity, gt, mask)
    sh = evaluate_disparity(hier.disparity, gt, mask)
    assert sf["density"] > 0.95 and sf["bad1"] < 0.01
    assert sh["density"] > 0.95 and sh["bad1"] < 0.01
    assert sh["density"] >= sf["density"] - 0.01
    assert sh["avg_error"] <= sf["avg_error"] + 0.02
