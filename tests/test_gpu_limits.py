"""Size limits of the GPU matcher that surface as r3d errors instead of
wrong results: a tile volume of 2^31 cells or more (the path scans address a
tile with 32-bit byte offsets) is refused before any volume is allocated."""
import numpy as np
import pytest

from regard3d_mvs_amd import StereoOptions

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from regard3d_mvs_amd.native import Matcher
    with Matcher(0) as m:
        yield m


def test_tile_of_2_31_cells_is_refused(gpu):
    # detail::sgm_cost_volume matches the whole image as one tile:
    # 4096 x 2048 pixels x 256 labels = 2^31 cells.
    rng = np.random.default_rng(3)
    left = (rng.integers(0, 256, (2048, 4096)) * 257).astype(np.uint16)
    right = np.ascontiguousarray(left)
    with pytest.raises(Exception, match="2\\^31"):
        gpu.cost_volume(left, right, 0, 256, StereoOptions())
    # The context stays usable.
    small = gpu.cost_volume(left[:32, :64].copy(), right[:32, :64].copy(), 0, 16, StereoOptions())
    assert small.size == 32 * 64 * 16


def test_more_than_4000_labels_per_pixel_is_an_error(gpu):
    # The step records of the path scans hold a pixel's label count / 16 in
    # 8 bits; a single full-range level of 4016 labels is refused with
    # r3d::Error (the reference accepts it: stereo_matcher.hpp:73-76).
    from regard3d_mvs_amd.native import R3dgError
    from regard3d_mvs_amd._abi import Status
    rng = np.random.default_rng(4)
    left = (rng.integers(0, 256, (8, 64)) * 257).astype(np.uint16)
    with pytest.raises(R3dgError) as e:
        gpu.match(left, left.copy(), None, None, 0, 4016, StereoOptions(levels=0))
    assert e.value.status == Status.INVALID_ARGUMENT and "4000 labels" in str(e.value)
    got = gpu.match(left, left.copy(), None, None, 0, 4000, StereoOptions(levels=0))
    assert got.disparity.shape == left.shape
