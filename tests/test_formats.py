"""PFM and .r3dd files (io/pfm.cpp:30-102, mvs/depth_map.cpp:94-221) through
the library's host-only entry points: the reference's round-trip tests
(test_stereo.cpp:496-503, test_mvs.cpp:356-381) restated, byte-level
known-answer files written here by hand, and the reference's error cases."""
import struct

import numpy as np
import pytest

from regard3d_mvs_amd import DepthMapFile
from regard3d_mvs_amd.native import R3dgError, read_depth_map, read_pfm, write_depth_map, write_pfm


def test_pfm_round_trip(tmp_path):
    img = np.array([[x * 10 + y + 0.25 for x in range(5)] for y in range(3)], np.float32)
    img[1, 2] = np.inf
    write_pfm(tmp_path / "a.pfm", img)
    back = read_pfm(tmp_path / "a.pfm")
    assert back.dtype == np.float32 and np.array_equal(back.view(np.uint32), img.view(np.uint32))
    data = (tmp_path / "a.pfm").read_bytes()
    assert data.startswith(b"Pf\n5 3\n-1\n")  # little-endian, rows bottom to top
    assert struct.unpack("<f", data[10:14])[0] == img[2, 0]


def test_pfm_big_endian_and_colour(tmp_path):
    rows = [[1.5, -2.0], [3.25, 1e-3]]  # top row first
    body = b"".join(struct.pack(">ff", *r) for r in reversed(rows))
    (tmp_path / "b.pfm").write_bytes(b"Pf\n2 2\n1.0\n" + body)
    assert np.array_equal(read_pfm(tmp_path / "b.pfm"), np.array(rows, np.float32))
    col = np.arange(2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3)
    write_pfm(tmp_path / "c.pfm", col)
    assert (tmp_path / "c.pfm").read_bytes().startswith(b"PF\n3 2\n-1\n")
    assert np.array_equal(read_pfm(tmp_path / "c.pfm"), col)


def test_pfm_errors(tmp_path):
    (tmp_path / "x.pfm").write_bytes(b"P6\n2 2\n255\n")
    with pytest.raises(R3dgError):
        read_pfm(tmp_path / "x.pfm")
    (tmp_path / "t.pfm").write_bytes(b"Pf\n4 4\n-1\n" + b"\0" * 20)
    with pytest.raises(R3dgError):
        read_pfm(tmp_path / "t.pfm")
    with pytest.raises(R3dgError):
        read_pfm(tmp_path / "missing.pfm")
    with pytest.raises(R3dgError):
        write_pfm(tmp_path / "bad.pfm", np.zeros((2, 2, 2), np.float32))


def sample_map():
    h, w = 9, 11
    rng = np.random.default_rng(3)
    depth = (5.0 + rng.random((h, w))).astype(np.float32)
    depth[2, 1] = 0.0
    pm = np.ones((h, w), np.uint16)
    pm[6, 5] = 3
    sup = np.zeros((h, w), np.uint8)
    sup[8, 7] = 4
    col = np.zeros((h, w, 3), np.uint8)
    col[..., 0], col[..., 1], col[..., 2] = 200, 100, 50
    rot = (0.99, -0.1, 0.05, 0.1, 0.99, 0.0, -0.05, 0.0, 0.99)
    return DepthMapFile(7, w, h, 280.0, 281.0, 5.2, 4.1, rot, (0.1, -0.2, 0.3), (3, 9), depth,
                        np.full((h, w), 0.5, np.float32), pm, sup, col)


def test_depth_map_round_trip(tmp_path):
    m = sample_map()
    write_depth_map(tmp_path / "d.r3dd", m)
    r = read_depth_map(tmp_path / "d.r3dd")
    assert (r.view, r.partners, r.cx, r.rotation, r.center) == (7, (3, 9), 5.2, m.rotation, m.center)
    for a, b in ((r.depth, m.depth), (r.confidence, m.confidence), (r.pair_mask, m.pair_mask), (r.support, m.support),
                 (r.color, m.color)):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    assert bin(int(r.pair_mask[6, 5])).count("1") == 2


def test_depth_map_bytes(tmp_path):
    m = sample_map()
    write_depth_map(tmp_path / "d.r3dd", m)
    data = (tmp_path / "d.r3dd").read_bytes()
    head = struct.pack("<4sIIii4d9d3dI2Iii", b"R3DD", 2, 7, 11, 9, 280.0, 281.0, 5.2, 4.1, *m.rotation, *m.center, 2,
                       3, 9, 11, 9)
    assert data[:len(head)] == head
    assert len(data) == len(head) + 11 * 9 * (4 + 4 + 2 + 1 + 3)
    assert data[len(head):len(head) + 4 * 99] == m.depth.tobytes()


def test_depth_map_errors(tmp_path):
    m = sample_map()
    write_depth_map(tmp_path / "d.r3dd", m)
    data = bytearray((tmp_path / "d.r3dd").read_bytes())
    (tmp_path / "short.r3dd").write_bytes(bytes(data[:-5]))
    with pytest.raises(R3dgError):
        read_depth_map(tmp_path / "short.r3dd")
    bad = data.copy()
    bad[4:8] = struct.pack("<I", 3)
    (tmp_path / "v3.r3dd").write_bytes(bytes(bad))
    with pytest.raises(R3dgError):
        read_depth_map(tmp_path / "v3.r3dd")
    off = 4 + 4 + 4 + 8 + 4 * 8 + 9 * 8 + 3 * 8
    bad = data.copy()
    bad[off:off + 4] = struct.pack("<I", 17)
    (tmp_path / "p17.r3dd").write_bytes(bytes(bad))
    with pytest.raises(R3dgError):
        read_depth_map(tmp_path / "p17.r3dd")
