"""GPU rectify_image (r3dg_rectify) against the CPU oracle
(oracle/rectify.cpp, pinned by test_oracle_rectify.py): bit-exact for the
polynomial lens models in every element type, with pyramid levels, rotations
and out-of-image pixels; the fisheye model goes through atan, whose last bit
the GPU and host libraries may round differently, so it is checked within one
grey level (and 1e-4 relative for float) and a tiny fraction of mask pixels."""
import numpy as np
import pytest

from oracle import oracle as O
from regard3d_mvs_amd import Camera, CameraModel
from test_oracle_rectify import rot

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


def source(dtype, h, w, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    base = 0.5 + 0.35 * np.sin(x / 7.0) * np.cos(y / 5.0) + 0.15 * rng.random((h, w))
    if dtype == np.float32:
        return (1000.0 * base).astype(np.float32)
    return np.round(base * np.iinfo(dtype).max).astype(dtype)


MODELS = [
    (CameraModel.pinhole, ()),
    (CameraModel.radial_k1, (-0.12,)),
    (CameraModel.radial_k3, (-0.1, 0.03, -0.004)),
    (CameraModel.brown_t2, (-0.08, 0.02, 0.001, 0.001, -0.0015)),
]


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("model,params", MODELS)
def test_rectify_bit_exact(gpu, dtype, model, params):
    h, w = 150, 211
    src = source(dtype, h, w, 3)
    cam = Camera(w, h, 180.0, 181.5, 104.3, 76.1, model, params, rot(0.02, -0.03, 0.01).ravel(), (0.5, -1.0, 2.0))
    tgt = Camera(230, 160, 175.0, 175.0, 118.0, 77.0, CameraModel.pinhole, (), rot(-0.01, 0.02, 0.03).ravel(),
                 (0.5, -1.0, 2.0))
    g_out, g_valid = gpu.rectify(src, cam, tgt)
    c_out, c_valid = O.rectify(src, cam, tgt)
    same_bits(g_out, c_out)
    same_bits(g_valid, c_valid)
    assert 0 < (c_valid == 255).mean() < 1  # both inside and outside pixels


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("shrink", [1.6, 3.5, 9.0])
def test_rectify_pyramid_levels(gpu, dtype, shrink):
    # Targets of lower resolution sample pyramid levels 1, 2 and 3.
    h, w = 300, 401
    src = source(dtype, h, w, 7)
    cam = Camera(w, h, 320.0, 320.0, 200.0, 150.0, CameraModel.radial_k3, (-0.05, 0.01, 0.0))
    f = 320.0 / shrink
    tgt = Camera(int(w / shrink) + 3, int(h / shrink) + 2, f, f, w / shrink / 2, h / shrink / 2, CameraModel.pinhole, (),
                 rot(0.0, 0.01, -0.02).ravel())
    g_out, g_valid = gpu.rectify(src, cam, tgt)
    c_out, c_valid = O.rectify(src, cam, tgt)
    same_bits(g_out, c_out)
    same_bits(g_valid, c_valid)


def test_rectify_fisheye_within_rounding(gpu):
    h, w = 140, 190
    src = source(np.uint16, h, w, 11)
    cam = Camera(w, h, 120.0, 120.0, 95.0, 70.0, CameraModel.fisheye, (0.02, -0.01, 0.003, -0.0005),
                 rot(0.03, 0.0, 0.0).ravel())
    tgt = Camera(200, 150, 110.0, 110.0, 100.0, 75.0)
    g_out, g_valid = gpu.rectify(src, cam, tgt)
    c_out, c_valid = O.rectify(src, cam, tgt)
    assert (g_valid != c_valid).mean() < 1e-3
    both = (g_valid == 255) & (c_valid == 255)
    assert np.abs(g_out[both].astype(np.int64) - c_out[both].astype(np.int64)).max() <= 1
    f_out, _ = gpu.rectify(src.astype(np.float32), cam, tgt)
    cf_out, _ = O.rectify(src.astype(np.float32), cam, tgt)
    np.testing.assert_allclose(f_out[both], cf_out[both], rtol=1e-4, atol=1e-2)


def test_rectify_device_tensors_and_no_mask(gpu):
    import torch
    h, w = 64, 96
    src = source(np.uint16, h, w, 13)
    cam = Camera(w, h, 80.0, 80.0, 47.5, 31.5, CameraModel.radial_k1, (-0.1,))
    tgt = Camera(90, 60, 78.0, 78.0, 45.0, 30.0)
    d_out, d_valid = gpu.rectify(torch.from_numpy(src.astype(np.int32)).to(torch.uint16).cuda(), cam, tgt)
    h_out, _ = gpu.rectify(src, cam, tgt, want_valid=False)
    c_out, c_valid = O.rectify(src, cam, tgt)
    same_bits(d_out.cpu().numpy(), c_out)
    same_bits(d_valid.cpu().numpy(), c_valid)
    same_bits(h_out, c_out)


def test_rectify_errors(gpu):
    from regard3d_mvs_amd.native import R3dgError
    src = np.zeros((10, 12), np.uint16)
    with pytest.raises(R3dgError):
        gpu.rectify(src, Camera(13, 10, 1, 1, 0, 0), Camera(12, 10, 1, 1, 0, 0))
    with pytest.raises(R3dgError):
        gpu.rectify(src, Camera(12, 10, 1, 1, 0, 0), Camera(12, 10, 1, 1, 0, 0, center=(0.0, 0.0, 1e-3)))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_rectify_unaligned_rows(gpu, dtype):
    # A device view starting one element into its rows (and rows of an odd
    # length): the tap loads take the element-wise path.
    import torch
    h, w = 70, 101
    full = source(dtype, h, w + 1, 17)
    tdt = torch.uint8 if dtype == np.uint8 else torch.uint16
    view = torch.from_numpy(full.astype(np.int32)).to(tdt).cuda()[:, 1:]
    src = full[:, 1:].copy()
    cam = Camera(w, h, 90.0, 90.0, 50.0, 35.0, CameraModel.brown_t2, (-0.05, 0.01, 0.0, 0.001, 0.001),
                 rot(0.01, 0.0, -0.01).ravel())
    tgt = Camera(105, 72, 88.0, 88.0, 52.0, 36.0)
    d_out, d_valid = gpu.rectify(view, cam, tgt)
    c_out, c_valid = O.rectify(src, cam, tgt)
    same_bits(d_out.cpu().numpy(), c_out)
    same_bits(d_valid.cpu().numpy(), c_valid)
