"""Closed-form pins of the oracle's float composition: selection, sub-pixel
fit, confidence, the diagonal and full left-right checks, and tile blending.

Each check recomputes the quantity from its definition (numpy, mostly in
float64) on inputs where the definition decides the answer, so that a
dropped term, a swapped neighbour, a wrong offset or a wrong rounding mode in
the oracle fails:

* selection (stereo_matcher.cpp:446-525) on a single-tile match, whose blend
  is the identity: first argmin and second-best outside best +- 1 by brute
  force over the aggregated volume, the right-column and uniqueness
  rejections, the sub-pixel disparity within 1/2 ulp(d) + 1e-7 of its FP64
  value (parabola and equiangular fits, stereo_matcher.cpp:426-436), and the
  confidence 1 - best/second within 2.5e-7 of FP64;
* the diagonal left-right check (sgm_kernels.cpp:427-435) against a brute
  force right-image winner per diagonal (first x on ties);
* the full left-right check (stereo_matcher.cpp:777-805) against a per-pixel
  loop with lround's half-away-from-zero rounding, on values that sit on .5;
* the blend (stereo_matcher.cpp:530-543, 688-766): one tile is the identity,
  two tiles with the same map return it within 4 ulp, the weights of two
  overlapping tiles follow the FP64 ramp (centred on the core boundary, pixel
  centres at x + 0.5), the valid weight must reach half of the total, and a
  mean more than 1 away from the dominant tile's value is replaced by it.
"""
import numpy as np
import pytest

from oracle import oracle as O
from regard3d_mvs_amd import StereoOptions, SubpixelMethod, synth

INVALID = 0xFFFF


def single_tile(**kw):
    base = dict(levels=0, tile_size=4096, speckle_size=0, median=False)
    base.update(kw)
    return StereoOptions(**base)


def volume(left, right, dmin, nd, o):
    h, w = left.shape
    D = (nd + 15) // 16 * 16
    return O.cost_volume(left, right, dmin, nd, o).reshape(h, w, D).astype(np.int64)


def brute_wta(S):
    best = np.argmin(S, axis=2)  # first argmin
    bcost = np.take_along_axis(S, best[..., None], 2)[..., 0]
    k = np.arange(S.shape[2])[None, None, :]
    outside = (k + 1 < best[..., None]) | (k > best[..., None] + 1)
    second = np.where(outside, S, INVALID).min(axis=2)
    return best, bcost, second


def subpixel64(method, sm, s0, sp):
    sm, s0, sp = (np.asarray(v, np.float64) for v in (sm, s0, sp))
    if method == SubpixelMethod.parabola:
        den = sm - 2 * s0 + sp
        delta = np.where(den > 0, (sm - sp) / np.where(den > 0, 2 * den, 1), 0.0)
    elif method == SubpixelMethod.equiangular:
        slope = np.maximum(sm - s0, sp - s0)
        delta = np.where(slope > 0, (sm - sp) / np.where(slope > 0, 2 * slope, 1), 0.0)
    else:
        delta = np.zeros_like(s0)
    return np.clip(delta, -0.5, 0.5)


def pair(seed, w=150, h=60, span=40, wr=None):
    p = synth.make_pair(w, h, span, seed)
    right = p.right if wr is None else np.ascontiguousarray(p.right[:, :wr])
    return p.left, right


@pytest.mark.parametrize("method", [SubpixelMethod.parabola, SubpixelMethod.equiangular, SubpixelMethod.none])
@pytest.mark.parametrize("dmin,nd,wr,uniq", [(0, 48, None, 0.0), (-6, 40, 130, 0.0), (2, 32, None, 0.15)])
def test_select_subpixel_confidence(method, dmin, nd, wr, uniq):
    left, right = pair(11 + nd, wr=wr)
    o = single_tile(subpixel=method, lr_max_diff=-1, uniqueness=uniq)
    g = O.match_stereo(left, right, None, None, dmin, nd, o, threads=2)
    S = volume(left, right, dmin, nd, o)
    h, w, D = S.shape
    best, bcost, second = brute_wta(S)
    X = np.arange(w)[None, :]
    xr = X - dmin - best
    ok = (xr >= 0) & (xr <= right.shape[1] - 1)
    if uniq > 0:
        ratio = bcost.astype(np.float64) / np.where(second == INVALID, 1, second)
        rej = (second != INVALID) & (ratio > 1.0 - uniq)
        # Ratios of integer costs can sit on the threshold, where float32
        # rounding of (1 - u) * second decides: leave those pixels out.
        tie = (second != INVALID) & (np.abs(ratio - (1.0 - uniq)) < 1e-6)
        ok &= ~rej
        assert rej.sum() > 0 and tie.sum() < 0.01 * tie.size
    else:
        tie = np.zeros(ok.shape, bool)
    assert np.array_equal(np.isfinite(g.disparity)[~tie], ok[~tie])
    ok = np.isfinite(g.disparity) & ~tie
    assert np.array_equal(g.labels[~tie], np.where(ok, dmin + best, np.iinfo(np.int16).min).astype(np.int16)[~tie])
    # Sub-pixel disparity against FP64.
    j = best
    inner = ok & (j > 0) & (j < D - 1)
    jm, jp = np.clip(j - 1, 0, D - 1), np.clip(j + 1, 0, D - 1)
    sm = np.take_along_axis(S, jm[..., None], 2)[..., 0]
    sp = np.take_along_axis(S, jp[..., None], 2)[..., 0]
    d64 = dmin + best + np.where(inner, subpixel64(method, sm, bcost, sp), 0.0)
    got = g.disparity.astype(np.float64)
    bound = 0.5 * np.spacing(np.abs(g.disparity)).astype(np.float64) + 1e-7
    err = np.abs(got - d64)
    assert np.all(err[ok] <= bound[ok]), f"max error {err[ok].max()} px"
    if method != SubpixelMethod.none:
        assert np.any(np.abs(d64[inner] - np.rint(d64[inner])) > 0.05)
    # Confidence against FP64.
    c64 = np.where((second != INVALID) & (second > 0), 1.0 - bcost / np.maximum(second, 1), 0.0)
    cerr = np.abs(g.confidence.astype(np.float64) - c64)
    assert np.all(cerr[ok] <= 2.5e-7), f"max confidence error {cerr[ok].max()}"
    assert np.all(g.confidence[~np.isfinite(g.disparity)] == 0.0)


@pytest.mark.parametrize("max_diff", [0, 1, 3])
def test_diagonal_lr_check(max_diff):
    left, right = pair(5, w=140, h=40, span=32)
    dmin, nd = 0, 32
    o = single_tile(subpixel=SubpixelMethod.none, lr_full_match=False, lr_max_diff=max_diff)
    g = O.match_stereo(left, right, None, None, dmin, nd, o, threads=2)
    S = volume(left, right, dmin, nd, o)
    h, w, D = S.shape
    best, _, _ = brute_wta(S)
    ok = (np.arange(w)[None, :] - dmin - best) >= 0
    xs = np.arange(w)
    for y in range(h):
        # M[i, x] = S(x, x - (i - D + 1)) on the diagonal i; first x wins ties.
        M = np.full((w + D - 1, w), 0x10000, np.int64)
        for k in range(D):
            M[xs - k + D - 1, xs] = S[y, :, k]
        xstar = np.argmin(M, axis=1)
        mval = M[np.arange(w + D - 1), xstar]
        rbest = np.where(mval < INVALID, xstar - (np.arange(w + D - 1) - D + 1), INVALID)
        i = xs - best[y] + D - 1
        ok[y] &= np.abs(rbest[i] - best[y]) <= max_diff
    assert np.array_equal(np.isfinite(g.disparity), ok)
    assert 0 < (~ok).sum() < ok.size


def lround(v):
    v = np.asarray(v, np.float64)
    return (np.sign(v) * np.floor(np.abs(v) + 0.5)).astype(np.int64)


@pytest.mark.parametrize("sign", [-1, 1])
def test_full_lr_check_rounding(sign):
    rng = np.random.default_rng(3)
    h, w, wo = 9, 40, 37
    d = (rng.integers(0, 16, (h, w)) + rng.choice([0.0, 0.5, 0.25, 0.75], (h, w))).astype(np.float32)
    if sign > 0:
        d = -d
    d[rng.random((h, w)) < 0.1] = np.inf
    other = d[:, :wo].copy() + rng.choice([0.0, 0.5, 1.0, 1.5, -1.0], (h, wo)).astype(np.float32)
    other[rng.random((h, wo)) < 0.1] = np.inf
    for max_diff in (0, 1):
        got = O.lr_mismatch(d, other, sign, max_diff)
        exp = np.zeros((h, w), np.uint8)
        for y in range(h):
            for x in range(w):
                if not np.isfinite(d[y, x]):
                    continue
                xo = int(lround(np.float32(x) + np.float32(sign) * d[y, x]))
                do = other[y, xo] if 0 <= xo < wo else np.inf
                exp[y, x] = 1 if (not np.isfinite(do) or abs(float(d[y, x]) - float(do)) > max_diff) else 0
        assert np.array_equal(got, exp)
    # Half-way cases decide between rounding modes: x - d = -0.5 is outside.
    d1 = np.zeros((1, 4), np.float32)
    d1[0, 0] = 0.5
    d1[0, 2] = 1.5  # 2 - 1.5 = 0.5 -> column 1 (away from zero), not 0 (to even)
    oth = np.array([[np.inf, 1.5, 7.0, 7.0]], np.float32)
    m = O.lr_mismatch(d1, oth, -1, 0)
    assert m[0, 0] == 1 and m[0, 2] == 0


def ramp(v, c, b, rising):
    pos = np.asarray(v, np.float64) + 0.5
    a = (pos - c + 0.5 * b) / b if rising else (c + 0.5 * b - pos) / b
    return np.clip(a, 0.0, 1.0)


def two_tiles(width=200, height=30, c=100, ext=24):
    """Tile 0 covers [0, c + ext) with core [0, c); tile 1 covers [c - ext, width)
    with core [c, width)."""
    return [(0, 0, (0, c, 0, height), c + ext), (c - ext, 0, (c, width, 0, height), width - c + ext)]


def test_blend_single_tile_is_identity():
    rng = np.random.default_rng(1)
    d = rng.uniform(-5, 60, (25, 40)).astype(np.float32)
    d[rng.random(d.shape) < 0.2] = np.inf
    c = rng.random((25, 40)).astype(np.float32)
    od, oc = O.blend(40, 25, 32, [(0, 0, (0, 40, 0, 25), d, c)])
    assert np.array_equal(od.view(np.uint32), d.view(np.uint32))
    assert np.array_equal(oc[np.isfinite(d)].view(np.uint32), c[np.isfinite(d)].view(np.uint32))
    assert np.all(oc[~np.isfinite(d)] == 0)


@pytest.mark.parametrize("blend", [32, 16, 8])
def test_blend_two_tiles(blend):
    width, height, c = 200, 30, 100
    tiles = two_tiles(width, height, c)
    x = np.arange(width)
    w0 = ramp(x, c, blend, rising=False)  # tile 0's weight (its right core edge at c)
    w1 = ramp(x, c, blend, rising=True)
    assert np.allclose(w0 + w1, 1.0)
    # Same map in both tiles: returned within 4 ulp.
    yy, xx = np.mgrid[0:height, 0:width]
    full = (3.0 + 0.013 * xx + 0.007 * yy).astype(np.float32)
    spec = [(x0, y0, core, full[:, x0:x0 + tw], np.full((height, tw), 0.5, np.float32))
            for x0, y0, core, tw in tiles]
    od, oc = O.blend(width, height, blend, spec)
    assert np.all(np.abs(od - full) <= 4 * np.spacing(full))
    # Maps 0 and 1: the result is tile 1's weight share, the FP64 ramp.
    spec = [(x0, y0, core, np.full((height, tw), float(i), np.float32), np.full((height, tw), float(i), np.float32))
            for i, (x0, y0, core, tw) in enumerate(tiles)]
    od, oc = O.blend(width, height, blend, spec)
    share = w1 / (w0 + w1)
    assert np.all(np.abs(od - share[None, :]) <= 1e-6)
    assert np.all(np.abs(oc - share[None, :]) <= 1e-6)
    # Tile 1 invalid in the overlap: valid only where tile 0 holds at least half the weight.
    spec = [(x0, y0, core, np.full((height, tw), np.inf if i == 1 else 5.0, np.float32),
             np.zeros((height, tw), np.float32)) for i, (x0, y0, core, tw) in enumerate(tiles)]
    od, _ = O.blend(width, height, blend, spec)
    decided = np.abs(w0 - w1) > 1e-6
    assert np.array_equal(np.isfinite(od[0])[decided], (w0 >= 0.5 * (w0 + w1))[decided])
    # Values 10 and 14: the mean where it is within 1 of the dominant tile, else that tile's value.
    spec = [(x0, y0, core, np.full((height, tw), 10.0 + 4 * i, np.float32), np.zeros((height, tw), np.float32))
            for i, (x0, y0, core, tw) in enumerate(tiles)]
    od, _ = O.blend(width, height, blend, spec)
    mean = (10.0 * w0 + 14.0 * w1) / (w0 + w1)
    dom = np.where(w1 > w0, 14.0, 10.0)  # the first tile wins ties
    exp = np.where(np.abs(mean - dom) > 1.0, dom, mean)
    margin = np.abs(np.abs(mean - dom) - 1.0) > 1e-5
    assert np.all(np.abs(od[0] - exp)[margin] <= 1e-5)
    assert np.any(exp != dom) and np.any(exp == dom)
