"""The GPU library reproduces the committed golden fixtures (tests/golden/,
the oracle's outputs on small seeded pairs) bit for bit: Census, aggregated
volumes and complete matches."""
import pytest

from golden_fixtures import cases, load, options, same_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from regard3d_mvs_amd.native import Matcher
    with Matcher(0) as m:
        yield m


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_gpu_reproduces_fixture(gpu, case):
    o = options(case)
    f = case["files"]
    left, right = load(case, "left"), load(case, "right")
    if "census_left" in f:
        same_bits(gpu.census(left), load(case, "census_left"))
        same_bits(gpu.census(right), load(case, "census_right"))
    if "volume" in f:
        ranges = (load(case, "lo"), load(case, "num")) if "lo" in f else None
        same_bits(gpu.cost_volume(left, right, case["dmin"], case["nd"], o, ranges=ranges), load(case, "volume"))
    if "disparity" in f:
        lv = load(case, "left_valid") if "left_valid" in f else None
        rv = load(case, "right_valid") if "right_valid" in f else None
        g = gpu.match(left, right, lv, rv, case["dmin"], case["nd"], o, debug=True)
        for k in ("disparity", "confidence", "labels", "lr_mask"):
            same_bits(getattr(g, k), load(case, k))
        assert g.cells == case["cells"] and g.levels == case["levels"]
