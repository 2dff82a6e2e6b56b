"""Transfers at the boundary: uint8 images widened on the device
(R3DG_FORMAT_U8X257), page-locked host buffers moved by DMA without staging,
and torch device tensors written by torch kernels right before the call.
Every form gives the maps of the plain host uint16 call, bit for bit."""
import numpy as np
import pytest

from regard3d_mvs_amd import StereoOptions, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hip_runtime):
    from regard3d_mvs_amd.native import Matcher
    with Matcher(0) as m:
        yield m


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape
    bad = a.view(np.uint32) != b.view(np.uint32)
    assert int(bad.sum()) == 0, f"{int(bad.sum())} of {a.size} differ"


@pytest.fixture(scope="module")
def case():
    p = synth.make_pair(700, 300, 96, 42)
    return p, StereoOptions()


def test_u8_input_equals_u16(gpu, case):
    p, o = case
    ref = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    l8, r8 = (p.left // 257).astype(np.uint8), (p.right // 257).astype(np.uint8)
    got = gpu.match(l8, r8, None, None, p.dmin, p.nd, o)
    same_bits(got.disparity, ref.disparity)
    same_bits(got.confidence, ref.confidence)
    # A strided view (every other row of a taller buffer).
    tall = np.zeros((2 * l8.shape[0], l8.shape[1] + 13), np.uint8)
    tall[::2, 5:5 + l8.shape[1]] = l8
    got2 = gpu.match(tall[::2, 5:5 + l8.shape[1]], r8, None, None, p.dmin, p.nd, o)
    same_bits(got2.disparity, ref.disparity)


def test_page_locked_buffers(gpu, case):
    from regard3d_mvs_amd.native import host_empty
    p, o = case
    ref = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    h, w = p.left.shape
    l8, k1 = host_empty((h, w), np.uint8)
    r8, k2 = host_empty(p.right.shape, np.uint8)
    l8[...] = p.left // 257
    r8[...] = p.right // 257
    d, k3 = host_empty((h, w), np.float32)
    c, k4 = host_empty((h, w), np.float32)
    for _ in range(2):
        d[...] = 0
        gpu.match(l8, r8, None, None, p.dmin, p.nd, o, out=(d, c))
        same_bits(d, ref.disparity)
        same_bits(c, ref.confidence)


def test_torch_inputs_written_just_before(gpu, case):
    import torch
    p, o = case
    ref = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    dev = torch.device("cuda:0")
    l8 = torch.from_numpy((p.left // 257).astype(np.uint8)).to(dev)
    r8 = torch.from_numpy((p.right // 257).astype(np.uint8)).to(dev)
    for _ in range(3):
        # Widened by torch kernels queued on torch's stream immediately before the call.
        left = (l8.to(torch.int32) * 257).to(torch.uint16)
        right = (r8.to(torch.int32) * 257).to(torch.uint16)
        got = gpu.match(left, right, None, None, p.dmin, p.nd, o)
        same_bits(got.disparity.cpu().numpy(), ref.disparity)
        del left, right
    got8 = gpu.match(l8, r8, None, None, p.dmin, p.nd, o)
    same_bits(got8.disparity.cpu().numpy(), ref.disparity)
    same_bits(got8.confidence.cpu().numpy(), ref.confidence)


def test_batch_and_views_take_u8(gpu, case):
    p, o = case
    ref = gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o)
    l8, r8 = (p.left // 257).astype(np.uint8), (p.right // 257).astype(np.uint8)
    got = gpu.match_batch([dict(left=l8, right=r8, dmin=p.dmin, nd=p.nd)], o)[0]
    same_bits(got.disparity, ref.disparity)
    gpu.view_put(6, l8)  # before the loop runs: uploaded and widened on the device
    gpu.serve_start(o)
    try:
        gpu.view_put(5, l8)  # while the loop runs: widened on the host, uploaded by the loop
        t5 = gpu.submit(l8.shape, p.dmin, p.nd, right=r8, left_view=5)
        t6 = gpu.submit(l8.shape, p.dmin, p.nd, right=r8, left_view=6)
        g5, g6 = gpu.wait(t5), gpu.wait(t6)
    finally:
        gpu.serve_stop()
        gpu.view_evict(5)
        gpu.view_evict(6)
    same_bits(g5.disparity, ref.disparity)
    same_bits(g6.disparity, ref.disparity)
