"""Batched matching (r3dg_match_batch): every pair of a batch gets the same
maps, bit for bit, as when it is matched alone, whatever the mix of sizes,
disparity ranges, masks and level counts; the single-pair path is itself
pinned to the CPU oracle by test_gpu_parity.py."""
import numpy as np
import pytest

from refscenes import Texture, render_pair
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
    assert a.shape == b.shape and a.dtype == b.dtype
    bad = a.view(np.uint32) != b.view(np.uint32) if a.dtype == np.float32 else a != b
    assert int(bad.sum()) == 0, f"{int(bad.sum())} of {a.size} differ; first at {np.argwhere(bad)[:5].tolist()}"


def mixed_pairs():
    """Pairs of different sizes, widths, spans and masks (levels 0 to 3)."""
    out = []
    for i, (w, h, span) in enumerate([(160, 96, 32), (333, 211, 64), (97, 150, 48), (420, 260, 128)]):
        p = synth.make_pair(w, h, span, 500 + i)
        q = {"left": p.left, "right": p.right, "dmin": p.dmin, "nd": p.nd}
        if i == 2:
            m = np.full(p.left.shape, 255, np.uint8)
            m[:, :7] = 0
            m[40:60, 30:50] = 0
            q["left_valid"] = m
        out.append(q)
    left, right, _ = render_pair(Texture(5), 190, 70, lambda x, y: 6.0 + 0.02 * x)
    out.append({"left": left, "right": right[:, :180].copy(), "dmin": -2, "nd": 24})
    return out


def check_against_single(gpu, pairs, options, got):
    for q, g in zip(pairs, got):
        ref = gpu.match(q["left"], q["right"], q.get("left_valid"), q.get("right_valid"), q["dmin"], q["nd"],
                        options)
        assert g.cells == ref.cells and g.levels == ref.levels and g.tile_pixels == ref.tile_pixels
        same_bits(g.disparity, ref.disparity)
        same_bits(g.confidence, ref.confidence)


@pytest.mark.parametrize("options", [StereoOptions(), StereoOptions(lr_full_match=False, tile_size=96),
                                     StereoOptions(levels=0, median=False)])
def test_batch_equals_single(gpu, options):
    pairs = mixed_pairs()
    got = gpu.match_batch(pairs, options)
    check_against_single(gpu, pairs, options, got)


def test_batch_small_budget_groups_and_chunks(gpu):
    from regard3d_mvs_amd.native import Matcher
    pairs = mixed_pairs()
    with Matcher(0, hbm_budget=48 << 20) as small:
        got = small.match_batch(pairs, StereoOptions())
    check_against_single(gpu, pairs, StereoOptions(), got)


def test_batch_invalid_pair_is_skipped(gpu):
    from regard3d_mvs_amd.native import R3dgError
    pairs = mixed_pairs()[:3]
    bad = dict(pairs[1])
    bad["right"] = bad["right"][:-3].copy()  # unequal heights
    batch = [pairs[0], bad, pairs[2]]
    with pytest.raises(R3dgError):
        gpu.match_batch(batch, StereoOptions())
    got = gpu.match_batch(batch, StereoOptions(), raise_errors=False)
    assert got[1] is None
    check_against_single(gpu, [pairs[0], pairs[2]], StereoOptions(), [got[0], got[2]])


def test_batch_empty(gpu):
    assert gpu.match_batch([], StereoOptions()) == []


def test_batch_outputs_in_every_memory(gpu):
    """A group's maps leave on a copy stream: staged through pinned blocks for
    pageable host buffers of 1 MB or more, straight into page-locked host
    buffers and into device buffers.  Each form equals the plain call."""
    import torch
    from regard3d_mvs_amd.native import host_empty
    o = StereoOptions()
    pairs = [synth.make_pair(640, 480, 48, 700 + i) for i in range(3)]
    ref = [gpu.match(p.left, p.right, None, None, p.dmin, p.nd, o) for p in pairs]
    pinned = [host_empty((480, 640), np.float32) for _ in range(2)]
    dev = torch.device("cuda:0")
    reqs = [dict(left=pairs[0].left, right=pairs[0].right, dmin=pairs[0].dmin, nd=pairs[0].nd),
            dict(left=pairs[1].left, right=pairs[1].right, dmin=pairs[1].dmin, nd=pairs[1].nd,
                 out=(pinned[0][0], pinned[1][0])),
            dict(left=torch.from_numpy(pairs[2].left.astype(np.int32)).to(dev).to(torch.uint16),
                 right=torch.from_numpy(pairs[2].right.astype(np.int32)).to(dev).to(torch.uint16),
                 dmin=pairs[2].dmin, nd=pairs[2].nd)]
    got = gpu.match_batch(reqs, o)
    for g, r in zip(got, ref):
        d, c = g.disparity, g.confidence
        if hasattr(d, "cpu"):
            d, c = d.cpu().numpy(), c.cpu().numpy()
        same_bits(d, r.disparity)
        same_bits(c, r.confidence)
