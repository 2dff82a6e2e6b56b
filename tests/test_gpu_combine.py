"""GPU combine_pairs (r3dg_combine_pairs) against the CPU oracle
(oracle/combine.cpp, pinned by test_oracle_combine.py): depth, confidence and
pair mask bit-exact on the numpy rig, with invalid and noisy disparities,
disagreeing pairs and up to the 16-pair limit."""
import numpy as np
import pytest

from regard3d_mvs_amd.multiview import Rig
from oracle import oracle as O
from regard3d_mvs_amd import MultiStereoOptions, PairGeometry

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
    bad = a.view(np.uint32) != b.view(np.uint32) if a.dtype == np.float32 else a != b
    assert int(bad.sum()) == 0, f"{int(bad.sum())} of {a.size} differ; first at {np.argwhere(bad)[:5].tolist()}"


def noisy(p, rng, sigma, holes, conf=True):
    d = p.disparity + rng.normal(0.0, sigma, p.disparity.shape).astype(np.float32)
    d[rng.random(d.shape) < holes] = np.inf
    c = rng.random(d.shape).astype(np.float32) if conf else None
    return PairGeometry(p.grid_to_rect, p.fb, p.offset, p.partner_center, d, c)


def check(gpu, grid, pairs, options=None):
    g = gpu.combine_pairs(grid, pairs, options)
    c = O.combine_pairs(grid, pairs, options)
    for a, b in zip(g, c):
        same_bits(a, b)
    return c


def test_plane_three_pairs(gpu):
    rig = Rig()
    pairs = [rig.pair(rig.partners[0]), rig.pair(rig.partners[1]), rig.pair(rig.partners[2], shift=6.0)]
    depth, _, _ = check(gpu, rig.grid(), pairs)
    assert (depth > 0).mean() > 0.5


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_noisy_pairs(gpu, seed):
    rng = np.random.default_rng(seed)
    rig = Rig()
    base = [rig.pair(c) for c in rig.partners]
    pairs = [noisy(p, rng, 0.7, 0.1, conf=(i != 1)) for i, p in enumerate(base)]
    for opt in (None, MultiStereoOptions(inlier_px=1.0, max_reprojection_px=0.5, min_angle_deg=2.0, min_pairs=2)):
        check(gpu, rig.grid(), pairs, opt)


def test_sixteen_pairs_and_limit(gpu):
    rng = np.random.default_rng(7)
    offsets = [(0.6 * np.cos(a), 0.6 * np.sin(a), 0.05 * np.cos(3 * a)) for a in np.linspace(0, 2 * np.pi, 16,
                                                                                               endpoint=False)]
    rig = Rig(offsets)
    pairs = [noisy(rig.pair(c, shift=float(rng.choice([0.0, 0.0, 0.0, 4.0]))), rng, 0.3, 0.05) for c in rig.partners]
    depth, _, mask = check(gpu, rig.grid(), pairs)
    assert (depth > 0).mean() > 0.5 and mask.max() > 255  # pairs beyond the eighth are used
    from regard3d_mvs_amd.native import R3dgError
    with pytest.raises(R3dgError):
        gpu.combine_pairs(rig.grid(), pairs + pairs[:1])


def test_device_maps(gpu):
    import torch
    rig = Rig()
    pairs = [rig.pair(c) for c in rig.partners[:2]]
    dev = [PairGeometry(p.grid_to_rect, p.fb, p.offset, p.partner_center, torch.from_numpy(p.disparity).cuda(),
                        torch.from_numpy(p.confidence).cuda()) for p in pairs]
    g = gpu.combine_pairs(rig.grid(), dev)
    c = O.combine_pairs(rig.grid(), pairs)
    for a, b in zip(g, c):
        same_bits(a.cpu().numpy(), b)
