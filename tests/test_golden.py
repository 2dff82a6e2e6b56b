"""The oracle reproduces the committed golden fixtures (tests/golden/), bit
for bit: Census, aggregated volumes (full range and per-pixel ranges) and
complete matches (disparity, confidence, labels, left-right mask, cell
counts).  The PFM maps are read with the library's reader (io/pfm.cpp)."""
import numpy as np
import pytest

from golden_fixtures import cases, load, options, same_bits
from oracle import oracle as O


@pytest.fixture(scope="module", autouse=True)
def oracle_built():
    O.build()


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_oracle_reproduces_fixture(case):
    o = options(case)
    f = case["files"]
    left, right = load(case, "left"), load(case, "right")
    if "census_left" in f:
        same_bits(O.census(left), load(case, "census_left"))
        same_bits(O.census(right), load(case, "census_right"))
    if "volume" in f:
        ranges = (load(case, "lo"), load(case, "num")) if "lo" in f else None
        same_bits(O.cost_volume(left, right, case["dmin"], case["nd"], o, ranges=ranges), load(case, "volume"))
    if "disparity" in f:
        lv = load(case, "left_valid") if "left_valid" in f else None
        rv = load(case, "right_valid") if "right_valid" in f else None
        g = O.match_stereo(left, right, lv, rv, case["dmin"], case["nd"], o, threads=2)
        for k in ("disparity", "confidence", "labels", "lr_mask"):
            same_bits(getattr(g, k), load(case, k))
        assert g.cells == case["cells"] and g.levels == case["levels"]
        assert np.isfinite(g.disparity).mean() > 0.3
