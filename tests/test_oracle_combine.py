"""Pins the CPU oracle's combine_pairs (mvs/multi_stereo.cpp:44-216) with the
reference's own tests restated on a rig built in numpy
(regard3d_mvs_amd/multiview.py): sample_disparity KATs (test_mvs.cpp:217-227), a
slanted plane triangulated by three pairs with the wrong one outvoted
(:130-172), and the rejection of ties, single pairs under min_pairs = 2 and
narrow triangulation angles (:174-215)."""
import numpy as np

from regard3d_mvs_amd.multiview import Rig
from oracle import oracle as O
from regard3d_mvs_amd import MultiStereoOptions


def test_sample_disparity_kats():
    d = np.full((2, 4), 10.0, np.float32)
    d[0, 2] = 10.5
    d[0, 3] = 30.0
    d[1, 3] = np.inf
    assert abs(O.sample_disparity(d, 1.5, 0.0) - 10.25) < 1e-6
    assert O.sample_disparity(d, 2.4, 0.3) == 10.5  # jump to 30: nearest pixel
    assert O.sample_disparity(d, 2.6, 0.6) == np.inf  # nearest pixel invalid
    assert O.sample_disparity(d, -0.6, 0.0) == np.inf


def test_plane_from_three_pairs():
    rig = Rig()
    pairs = [rig.pair(rig.partners[0]), rig.pair(rig.partners[1]), rig.pair(rig.partners[2], shift=6.0)]
    depth, conf, mask = O.combine_pairs(rig.grid(), pairs)
    valid = depth > 0
    npairs = np.vectorize(lambda m: bin(int(m)).count("1"))(mask)
    two = valid & (npairs >= 2)
    assert valid.mean() > 0.5 and two.sum() > 0.8 * valid.sum()
    assert not (mask[two] & 4).any()  # the wrong pair is outvoted wherever two agree
    ys, xs = np.nonzero(two)
    rel = [abs(depth[y, x] / rig.depth(x, y) - 1.0) for y, x in zip(ys[::37], xs[::37])]
    assert max(rel) < 1e-3
    assert np.all(conf[valid] == 0.5)


def test_ties_min_pairs_and_narrow_angles():
    rig = Rig()
    a = rig.pair(rig.partners[0])
    b = rig.pair(rig.partners[1], shift=5.0)
    single = (O.combine_pairs(rig.grid(), [a])[0] > 0).sum()
    tie_d, _, tie_m = O.combine_pairs(rig.grid(), [a, b])
    assert single > 0 and (tie_d > 0).sum() < single / 2
    assert all(bin(int(m)).count("1") == 1 for m in tie_m[tie_d > 0])
    assert (O.combine_pairs(rig.grid(), [a], MultiStereoOptions(min_pairs=2))[0] > 0).sum() == 0
    close = rig.cb + rig.rb.T @ np.array([0.15, 0.0, 0.0])
    c = rig.pair(close)
    assert (O.combine_pairs(rig.grid(), [c])[0] > 0).sum() == 0
    assert (O.combine_pairs(rig.grid(), [c], MultiStereoOptions(min_angle_deg=1.0))[0] > 0).sum() > 0


def test_too_many_pairs():
    import pytest
    rig = Rig()
    a = rig.pair(rig.partners[0])
    with pytest.raises(ValueError):
        O.combine_pairs(rig.grid(), [a] * 17)
