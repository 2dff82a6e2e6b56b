"""Failure paths forced from the host, every kernel exiting normally:
cancellation requested before a match, a device-allocation failure injected
through the arena at each allocation of a match and of a served pair, and a
request refused at admission because it exceeds the context's device
budget.  After each failure the context gives the maps it gave before, bit
for bit."""
import numpy as np
import pytest

from regard3d_mvs_amd import StereoOptions, synth
from regard3d_mvs_amd._abi import Status

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from regard3d_mvs_amd.native import Matcher
    with Matcher(0) as m:
        yield m


@pytest.fixture(scope="module")
def case():
    return synth.make_pair(320, 200, 64, 77), StereoOptions()


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    assert int((a.view(np.uint32) != b.view(np.uint32)).sum()) == 0


def test_cancel_before_match(gpu, case):
    from regard3d_mvs_amd.native import Cancelled
    p, o = case
    ref = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    gpu.cancel(True)
    try:
        with pytest.raises(Cancelled):
            gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    finally:
        gpu.cancel(False)
    got = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    same_bits(got.disparity, ref.disparity)
    same_bits(got.confidence, ref.confidence)


def test_allocation_failure_at_every_site(gpu, case):
    from regard3d_mvs_amd.native import R3dgError
    p, o = case
    ref = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    failures = 0
    for k in range(2000):
        gpu.inject_alloc_failure(k)
        try:
            got = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
        except R3dgError as e:
            assert e.status == Status.OUT_OF_MEMORY, str(e)
            failures += 1
            continue
        finally:
            gpu.inject_alloc_failure(-1)
        # The first k at which the match completes: every allocation before it failed once.
        same_bits(got.disparity, ref.disparity)
        break
    else:
        pytest.fail("the match never completed")
    assert failures == k and failures > 10
    got = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    same_bits(got.disparity, ref.disparity)
    same_bits(got.confidence, ref.confidence)


def test_allocation_failure_in_the_request_loop(gpu, case):
    from regard3d_mvs_amd.native import R3dgError
    p, o = case
    ref = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    gpu.serve_start(o)
    try:
        gpu.inject_alloc_failure(5)
        t_bad = gpu.submit(p.left.shape, p.dmin, p.nd, left=p.left, right=p.right)
        with pytest.raises(R3dgError) as e:
            gpu.wait(t_bad)
        assert e.value.status == Status.OUT_OF_MEMORY
        gpu.inject_alloc_failure(-1)
        t_ok = gpu.submit(p.left.shape, p.dmin, p.nd, left=p.left, right=p.right)
        g = gpu.wait(t_ok)
    finally:
        gpu.inject_alloc_failure(-1)
        gpu.serve_stop()
    same_bits(g.disparity, ref.disparity)


def test_admission_refused_at_the_device_budget(hip_runtime):
    from regard3d_mvs_amd.native import Matcher, R3dgError
    small = synth.make_pair(160, 96, 32, 5)
    o = StereoOptions()
    with Matcher(0, hbm_budget=64 << 20) as m:
        alone = m.match(small.left, small.right, None, None, small.dmin, small.nd, o)
        m.serve_start(o)
        try:
            big = np.zeros((1024, 2048), np.uint16)
            with pytest.raises(R3dgError) as e:
                m.submit(big.shape, 0, 64, left=big, right=big)
            assert e.value.status == Status.BUDGET
            t = m.submit(small.left.shape, small.dmin, small.nd, left=small.left, right=small.right)
            g = m.wait(t)
        finally:
            m.serve_stop()
    same_bits(g.disparity, alone.disparity)
