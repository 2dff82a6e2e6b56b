"""r3dg_remove_speckles_static (the speckle filter allocating nothing, its
scratch a fixed device array per GPU) against r3dg_remove_speckles, bit for
bit, called through the C ABI on torch device maps: two calls in a row, the
second on a smaller map, so the second finds the first call's contents in
the shared array.  Also the maps the entry point refuses."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from regard3d_mvs_amd.native import Matcher
    with Matcher(0) as m:
        yield m


def speckled(w, h, seed):
    """Piecewise-flat disparities with small islands, noise and holes."""
    rng = np.random.default_rng(seed)
    d = np.repeat(np.repeat(rng.integers(0, 40, (h // 8 + 1, w // 8 + 1)), 8, 0), 8, 1)[:h, :w].astype(np.float32)
    d += (rng.random((h, w)) < 0.05) * rng.integers(3, 9, (h, w))
    d += rng.random((h, w)).astype(np.float32) * 0.4
    d[rng.random((h, w)) < 0.03] = np.inf
    return np.ascontiguousarray(d)


def call_static(gpu, t, min_size, max_diff):
    from regard3d_mvs_amd import native
    from regard3d_mvs_amd._abi import R3DG_DEVICE
    h, w = t.shape
    img = native.CImage(ctypes.c_void_p(t.data_ptr()), w, h, t.stride(0), R3DG_DEVICE, 0)
    lib = native.lib()
    lib.r3dg_remove_speckles_static.argtypes = [ctypes.c_void_p, native.CImage, ctypes.c_int32, ctypes.c_float]
    return lib.r3dg_remove_speckles_static(gpu._ctx, img, min_size, ctypes.c_float(max_diff))


@pytest.mark.parametrize("max_diff", [1.0, float("inf")])
def test_static_equals_allocating_entry_point(gpu, max_diff):
    import torch
    for (w, h), seed, min_size in (((2048, 1536), 1, 100), ((640, 480), 2, 40)):
        d = speckled(w, h, seed)
        ref = gpu.remove_speckles(d, min_size, max_diff)
        t = torch.from_numpy(d).cuda()
        torch.cuda.synchronize()
        assert call_static(gpu, t, min_size, max_diff) == 0
        got = t.cpu().numpy()
        bad = got.view(np.uint32) != ref.view(np.uint32)
        assert int(bad.sum()) == 0, f"{w}x{h}: {int(bad.sum())} pixels differ"
        if max_diff < float("inf"):
            assert np.isinf(got).sum() > np.isinf(d).sum()  # the filter removed something


def test_static_refuses_what_it_cannot_hold(gpu):
    import torch
    from regard3d_mvs_amd import native
    from regard3d_mvs_amd._abi import R3DG_HOST
    big = torch.zeros((2049, 2048), dtype=torch.float32, device="cuda")  # 1 row past the fixed scratch
    assert call_static(gpu, big, 10, 1.0) == 1
    strided = torch.zeros((64, 128), dtype=torch.float32, device="cuda")[:, :100]
    assert call_static(gpu, strided, 10, 1.0) == 1
    host = np.zeros((8, 8), np.float32)
    img = native.CImage(ctypes.c_void_p(host.ctypes.data), 8, 8, 8, R3DG_HOST, 0)
    assert native.lib().r3dg_remove_speckles_static(gpu._ctx, img, 10, ctypes.c_float(1.0)) == 1
    # The context still works after the refusals.
    t = torch.from_numpy(speckled(256, 128, 5)).cuda()
    ref = gpu.remove_speckles(t.cpu().numpy(), 20, 1.0)
    assert call_static(gpu, t, 20, 1.0) == 0
    assert np.array_equal(t.cpu().numpy().view(np.uint32), ref.view(np.uint32))
