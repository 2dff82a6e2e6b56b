"""GPU parity: every stage of libr3dg.so against the CPU oracle, bit-exact.

Integer outputs (Census, pyramids, aggregated costs, ranges, integer labels,
LR masks) must be identical; float maps (disparity, confidence, filters) must
be bit-identical too, since every float step is the same IEEE operation
sequence as the reference (-ffp-contract=off on both sides).
"""
import numpy as np
import pytest

from oracle import oracle as O
from refscenes import Texture, random_image, render_pair, to_u16
from regard3d_mvs_amd import StereoOptions, SubpixelMethod, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from regard3d_mvs_amd.native import Matcher
    with Matcher(0) as m:
        yield m


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    if a.dtype == np.float32:
        bad = a.view(np.uint32) != b.view(np.uint32)
    else:
        bad = a != b
    n = int(bad.sum())
    assert n == 0, f"{n} of {a.size} differ; first at {np.argwhere(bad)[:5].tolist()}"


# ---- Census
@pytest.mark.parametrize("w", [1, 5, 16, 31, 64, 77])
def test_census(gpu, w):
    rng = np.random.default_rng(w)
    img = rng.integers(0, 8 if w % 2 == 0 else 65536, size=(13, w)).astype(np.uint16)
    same_bits(gpu.census(img), O.census(img))


def test_census_kats(gpu):
    img = np.full((7, 9), 100, np.uint16)
    img[0, 0] = 50
    assert int(gpu.census(img)[3, 4]) == 1 << 61
    img = np.tile((10 * np.arange(20)).astype(np.uint16), (10, 1))
    assert gpu.census(img)[5, 0] == 0


def test_census_large(gpu):
    p = synth.config_pair(0)
    same_bits(gpu.census(p.left), O.census(p.left))


# ---- pyramid
@pytest.mark.parametrize("shape", [(41, 77), (1, 1), (2, 3), (480, 640)])
def test_pyramid(gpu, shape):
    rng = np.random.default_rng(1)
    a16 = rng.integers(0, 65536, size=shape).astype(np.uint16)
    a8 = rng.integers(0, 256, size=shape).astype(np.uint8)
    same_bits(gpu.pyramid_down(a16), O.pyramid_down(a16))
    same_bits(gpu.pyramid_down(a8), O.pyramid_down(a8))


# ---- aggregated costs
@pytest.mark.parametrize("p2_adapt", [0, 12])
@pytest.mark.parametrize("dmin,nd", [(-3, 16), (2, 48)])
def test_cost_volume_reference_cases(gpu, p2_adapt, dmin, nd):
    left, right, _ = render_pair(Texture(3), 41, 23, lambda x, y: 5.0 + 0.1 * x)
    right = right ^ (random_image(41, 23, 9) & 0x0F00)
    o = StereoOptions(p2_adapt=p2_adapt, p2=120)
    same_bits(gpu.cost_volume(left, right, dmin, nd, o), O.textbook_volume(left, right, dmin, nd, o))


def _ranges(w, h, nd, seed, maxk=3):
    rng = np.random.default_rng(seed)
    num = (16 * rng.integers(1, maxk + 1, size=(h, w))).astype(np.uint16)
    lo = (rng.random((h, w)) * (nd - num + 1)).astype(np.uint16)
    return lo, num


def test_cost_volume_ranges(gpu):
    w, h, nd = 37, 19, 96
    left, right, _ = render_pair(Texture(13), w, h, lambda x, y: 20.0 + 0.5 * x - 0.3 * y)
    lo, num = _ranges(w, h, nd, 31)
    o = StereoOptions(p2=150, p2_adapt=8)
    same_bits(gpu.cost_volume(left, right, -2, nd, o, (lo, num)), O.textbook_volume(left, right, -2, nd, o, (lo, num)))


@pytest.mark.parametrize("seed", [1, 2])
def test_cost_volume_large_random(gpu, seed):
    rng = np.random.default_rng(seed)
    h, wl, wr, nd = 97, 201, 230, 80
    left = rng.integers(0, 65536, size=(h, wl)).astype(np.uint16)
    right = rng.integers(0, 65536, size=(h, wr)).astype(np.uint16)
    o = StereoOptions(p1=11, p2=300, p2_adapt=6, invalid_cost=40)
    same_bits(gpu.cost_volume(left, right, -7, nd, o), O.cost_volume(left, right, -7, nd, o))
    lo, num = _ranges(wl, h, nd, seed + 10, maxk=5)
    same_bits(gpu.cost_volume(left, right, -7, nd, o, (lo, num)), O.cost_volume(left, right, -7, nd, o, (lo, num)))


def test_cost_volume_saturation(gpu):
    rng = np.random.default_rng(9)
    h, w, nd = 31, 45, 32
    left = rng.integers(0, 65536, size=(h, w)).astype(np.uint16)
    right = rng.integers(0, 65536, size=(h, w)).astype(np.uint16)
    o = StereoOptions(p1=7900, p2=8000, p2_adapt=0, invalid_cost=255)
    lo, num = _ranges(w, h, nd, 3, maxk=2)
    same_bits(gpu.cost_volume(left, right, 3, nd, o, (lo, num)), O.textbook_volume(left, right, 3, nd, o, (lo, num)))


def test_cost_volume_full_range_512(gpu):
    p = synth.make_pair(96, 40, 512, 5)
    same_bits(gpu.cost_volume(p.left, p.right, 0, 512), O.cost_volume(p.left, p.right, 0, 512))


def test_cost_volume_max_labels(gpu):
    # 4000 labels, the most a pixel may hold on the GPU: label counts up to
    # 4000 (count / 16 = 250) and first labels up to 3984 go through the path
    # scans' packed per-step records; every path cost lives in scratch rows.
    w, h, nd = 22, 5, 4000
    rng = np.random.default_rng(77)
    left = rng.integers(0, 65536, size=(h, w)).astype(np.uint16)
    right = rng.integers(0, 65536, size=(h, w + 30)).astype(np.uint16)
    o = StereoOptions(p2=300, p2_adapt=6)
    same_bits(gpu.cost_volume(left, right, -4010, nd, o), O.cost_volume(left, right, -4010, nd, o))
    num = (16 * rng.integers(1, 251, size=(h, w))).astype(np.uint16)
    lo = (rng.random((h, w)) * (nd - num + 1)).astype(np.uint16)
    num[0, :3] = nd
    lo[0, :3] = 0
    lo[1, :3] = nd - 16
    num[1, :3] = 16
    same_bits(gpu.cost_volume(left, right, -4010, nd, o, (lo, num)),
              O.cost_volume(left, right, -4010, nd, o, (lo, num)))


def test_cost_volume_mixed_wide_ranges(gpu):
    # Label counts from 16 to 512 in random order along every path: the path
    # costs of a pixel move between LDS rows (narrow) and scratch rows (wide)
    # in both directions.
    w, h, nd = 53, 21, 512
    left, right, _ = render_pair(Texture(29), w + 40, h, lambda x, y: 60.0 + 0.4 * x - 0.2 * y)
    lo, num = _ranges(w + 40, h, nd, 41, maxk=32)
    o = StereoOptions(p2=150, p2_adapt=8)
    same_bits(gpu.cost_volume(left, right, 3, nd, o, (lo, num)), O.cost_volume(left, right, 3, nd, o, (lo, num)))


def test_cost_volume_wide_row(gpu):
    # A single tile wider than the cost pass stages in LDS, with a partner
    # window wider than its LDS window: the pass reads everything from HBM.
    w, h, nd = 700, 6, 640
    left, right, _ = render_pair(Texture(17), w, h, lambda x, y: 200.0 + 0.1 * x)
    o = StereoOptions(p2=120)
    same_bits(gpu.cost_volume(left, right, -20, nd, o), O.cost_volume(left, right, -20, nd, o))


# ---- ranges and filters
@pytest.mark.parametrize("mirrored", [False, True])
@pytest.mark.parametrize("sign", [-1, 1])
def test_ranges_from_coarse(gpu, mirrored, sign):
    rng = np.random.default_rng(3)
    wc, hc = 40, 23
    coarse = rng.uniform(-5, 40, size=(hc, wc)).astype(np.float32)
    coarse[rng.random(coarse.shape) < 0.3] = np.inf
    coarse[:, 25:] = np.inf
    w, h = 2 * wc - 1, 2 * hc
    fv = np.full((h, w), 255, np.uint8)
    fv[rng.random(fv.shape) < 0.05] = 0
    pv = np.full((h, 70), 255, np.uint8)
    pv[:, :9] = 0
    pv[5] = 0
    o = StereoOptions(range_radius=2, range_margin=3, range_fill=8)
    a = gpu.ranges_from_coarse(coarse, w, h, fv, pv, 70, sign, -4, 96, o, mirrored)
    b = O.ranges_from_coarse(coarse, w, h, fv, pv, 70, sign, -4, 96, o, mirrored)
    same_bits(a[0], b[0])
    same_bits(a[1], b[1])


@pytest.mark.parametrize("fill", [8, 64, 256])
def test_ranges_from_coarse_fill_levels(gpu, fill):
    """A coarse grid of several 64 x 64 regions with a hole that only high
    min / max pyramid levels reach across: levels up to 6 are built in one
    launch, those past 6 (range_fill 256) one launch each."""
    rng = np.random.default_rng(11)
    wc, hc = 600, 400
    yy, xx = np.mgrid[0:hc, 0:wc]
    coarse = (25 + 12 * np.sin(xx / 23.0) + 8 * np.cos(yy / 17.0) + rng.uniform(0, 1, (hc, wc))).astype(np.float32)
    coarse[rng.random(coarse.shape) < 0.2] = np.inf
    coarse[5:395, 10:590] = np.inf
    w, h = 2 * wc, 2 * hc - 1
    fv = np.full((h, w), 255, np.uint8)
    pv = np.full((h, w), 255, np.uint8)
    o = StereoOptions(range_radius=3, range_margin=4, range_fill=fill)
    a = gpu.ranges_from_coarse(coarse, w, h, fv, pv, w, -1, 0, 128, o, False)
    b = O.ranges_from_coarse(coarse, w, h, fv, pv, w, -1, 0, 128, o, False)
    same_bits(a[0], b[0])
    same_bits(a[1], b[1])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_speckles_and_median(gpu, seed):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, 5, size=(61, 83)).astype(np.float32) * 0.75 + rng.random((61, 83)).astype(np.float32) * 0.1
    d[rng.random(d.shape) < 0.15] = np.inf
    for size in (2, 7, 40):
        same_bits(gpu.remove_speckles(d, size, 1.0), O.remove_speckles(d, size, 1.0))
    # An infinite step joins invalid (+inf) pixels to their finite neighbours'
    # segments, as |inf - v| <= inf in the reference's segment growth.
    for size in (3, 200, 3000):
        same_bits(gpu.remove_speckles(d, size, np.inf), O.remove_speckles(d, size, np.inf))
    same_bits(gpu.median_filter_3x3(d), O.median_filter_3x3(d))


def test_speckles_large_segments(gpu):
    """Segments spanning many blocks (wave-aggregated counting near the
    saturation threshold)."""
    rng = np.random.default_rng(7)
    h, w = 300, 520
    d = np.full((h, w), 10.0, np.float32)
    d[:, 200:] += np.cumsum(rng.random((h, w - 200)).astype(np.float32) * 0.2, axis=1)
    d[rng.random(d.shape) < 0.05] = np.inf
    for size in (50, 4000, 60000, 150000):
        same_bits(gpu.remove_speckles(d, size, 0.3), O.remove_speckles(d, size, 0.3))


# ---- complete matcher
def _match_both(gpu, left, right, lm, rm, dmin, nd, o):
    g = gpu.match(left, right, lm, rm, dmin, nd, o, debug=True)
    c = O.match_stereo(left, right, lm, rm, dmin, nd, o)
    assert (g.levels, g.cells, g.num_disparities) == (c.levels, c.cells, c.num_disparities)
    same_bits(g.disparity, c.disparity)
    same_bits(g.confidence, c.confidence)
    same_bits(g.labels, c.labels)
    same_bits(g.lr_mask, c.lr_mask)
    return g, c


@pytest.mark.parametrize("method", [SubpixelMethod.parabola, SubpixelMethod.equiangular, SubpixelMethod.none])
def test_match_slanted_plane(gpu, method):
    left, right, _ = render_pair(Texture(11), 200, 120, lambda x, y: 12.3 + 0.04 * x + 0.02 * y)
    _match_both(gpu, left, right, None, None, 0, 32, StereoOptions(subpixel=method))


def test_match_small_tiles(gpu):
    left, right, _ = render_pair(Texture(5), 150, 90, lambda x, y: np.where(x < 70, 6.0, 14.5))
    _match_both(gpu, left, right, None, None, 0, 32, StereoOptions(tile_size=48, tile_margin=8, tile_blend=8))


def test_match_masks(gpu):
    left, right, _ = render_pair(Texture(8), 120, 60, lambda x, y: 10.0 + 0 * x)
    lm = np.full((60, 120), 255, np.uint8)
    rm = np.full((60, 120), 255, np.uint8)
    lm[:, 100:] = 0
    rm[:, :40] = 0
    _match_both(gpu, left, right, lm, rm, 0, 32, StereoOptions())


@pytest.mark.parametrize("full", [False, True])
def test_match_masked_left_lr(gpu, full):
    left, right, _ = render_pair(Texture(21), 160, 60, lambda x, y: 10.0 + 0 * x)
    rng = np.random.default_rng(4)
    left = to_u16(left + rng.normal(0, 3000, left.shape))
    right = to_u16(right + rng.normal(0, 3000, right.shape))
    yy, xx = np.mgrid[0:60, 0:160]
    lm = np.where(xx < 70 + 2 * yy, 255, 0).astype(np.uint8)
    _match_both(gpu, left, right, lm, None, 0, 256, StereoOptions(lr_full_match=full, tile_size=64))


@pytest.mark.parametrize("levels", [0, 1])
@pytest.mark.parametrize("full", [False, True])
def test_match_unequal_widths(gpu, levels, full):
    tex = Texture(23)
    yl, xl = np.mgrid[0:60, 0:150].astype(np.float64)
    yr, xr = np.mgrid[0:60, 0:190].astype(np.float64)
    left, right = to_u16(tex(xl, yl)), to_u16(tex(xr - 7.0, yr))
    _match_both(gpu, left, right, None, None, -20, 32, StereoOptions(lr_full_match=full, levels=levels))


def test_match_options_variants(gpu):
    left, right, _ = render_pair(Texture(17), 170, 97, lambda x, y: 9.0 + 0.03 * x - 0.02 * y)
    for o in (StereoOptions(uniqueness=0.1, median=False), StereoOptions(speckle_size=0, p2_adapt=0),
              StereoOptions(lr_max_diff=-1), StereoOptions(levels=2, range_fill=4, tile_size=64),
              StereoOptions(tile_blend=0, tile_size=80, memory_budget=1 << 20)):
        _match_both(gpu, left, right, None, None, -3, 40, o)


def test_match_config0_synthetic(gpu):
    p = synth.config_pair(0)
    for levels in (0, -1):
        _match_both(gpu, p.left, p.right, None, None, p.dmin, p.nd, StereoOptions(levels=levels))


def test_match_chunked_volume(gpu):
    """A context whose memory budget holds one tile's volume at a time runs the
    aggregation in chunks of tiles; the maps are the same as in one chunk."""
    from regard3d_mvs_amd.native import Matcher
    p = synth.config_pair(0)
    o = StereoOptions(levels=0)
    ref = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o, debug=True)
    with Matcher(0, hbm_budget=64 << 20) as small:
        got = small.match(p.left, p.right, None, None, p.dmin, p.nd, o, debug=True)
    assert got.cells == ref.cells
    for a, b in ((got.disparity, ref.disparity), (got.confidence, ref.confidence), (got.labels, ref.labels),
                 (got.lr_mask, ref.lr_mask)):
        same_bits(a, b)


def test_match_wide_tiles(gpu):
    # Tiles wider than the selection and cost passes stage in LDS.
    left, right, _ = render_pair(Texture(23), 1100, 24, lambda x, y: 20.0 + 0.01 * x)
    _match_both(gpu, left, right, None, None, 0, 48, StereoOptions(tile_size=2048, levels=0))


def test_match_hierarchical_wide_range(gpu):
    from test_oracle_matcher import hierarchical_scene
    left, right, _, _ = hierarchical_scene()
    _match_both(gpu, left, right, None, None, 0, 256, StereoOptions())


@pytest.mark.slow
def test_match_config1_synthetic(gpu):
    p = synth.config_pair(1)
    _match_both(gpu, p.left, p.right, None, None, p.dmin, p.nd, StereoOptions())
