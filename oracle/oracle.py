"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker of the GPU library.  Arrays are
C-contiguous numpy arrays; images are indexed [y, x].
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from regard3d_mvs_amd._abi import CMatchInfo, DisparityMap, StereoOptions, align16

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_last_error.restype = ctypes.c_char_p
        _lib.oracle_automatic_levels.restype = ctypes.c_int
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def _check(rc):
    if rc != 0:
        raise ValueError(lib().oracle_last_error().decode())


def census(img):
    img = _c(img, np.uint16)
    h, w = img.shape
    out = np.empty((h, w), np.uint64)
    _check(lib().oracle_census(_ptr(img), w, h, _ptr(out)))
    return out


def pyramid_down(img):
    img = np.ascontiguousarray(img)
    h, w = img.shape
    out = np.empty(((h + 1) // 2, (w + 1) // 2), img.dtype)
    fn = {np.dtype(np.uint16): lib().oracle_pyramid_down_u16, np.dtype(np.uint8): lib().oracle_pyramid_down_u8}
    _check(fn[img.dtype](_ptr(img), w, h, _ptr(out)))
    return out


def rectify(src, camera, target):
    """rectify_image (rectification.cpp:288-350), one channel: (image, valid)."""
    src = np.ascontiguousarray(src)
    fn = {np.dtype(np.uint8): lib().oracle_rectify_u8, np.dtype(np.uint16): lib().oracle_rectify_u16,
          np.dtype(np.float32): lib().oracle_rectify_f32}[src.dtype]
    out = np.empty((target.height, target.width), src.dtype)
    valid = np.empty((target.height, target.width), np.uint8)
    c, t = camera.to_c(), target.to_c()
    _check(fn(_ptr(src), src.shape[1], src.shape[0], ctypes.byref(c), ctypes.byref(t), _ptr(out), _ptr(valid)))
    return out, valid


def sample_disparity(d, u, v):
    """sample_disparity (mvs/multi_stereo.cpp:44-72)."""
    d = _c(d, np.float32)
    h, w = d.shape
    fn = lib().oracle_sample_disparity
    fn.restype = ctypes.c_float
    return float(fn(_ptr(d), w, h, ctypes.c_double(u), ctypes.c_double(v)))


def combine_pairs(grid, pairs, options=None):
    """combine_pairs (mvs/multi_stereo.cpp:74-216): (depth, confidence, pair mask)."""
    from regard3d_mvs_amd._abi import CImage, CPairGeometry, MultiStereoOptions
    arr = (CPairGeometry * max(len(pairs), 1))()
    keep = []
    for i, p in enumerate(pairs):
        d = _c(p.disparity, np.float32)
        keep.append(d)
        dv = CImage(d.ctypes.data_as(ctypes.c_void_p), d.shape[1], d.shape[0], d.shape[1], 0, 0)
        cv = CImage(None, 0, 0, 0, 0, 0)
        if p.confidence is not None:
            c = _c(p.confidence, np.float32)
            keep.append(c)
            cv = CImage(c.ctypes.data_as(ctypes.c_void_p), c.shape[1], c.shape[0], c.shape[1], 0, 0)
        arr[i] = CPairGeometry((ctypes.c_double * 9)(*[float(v) for v in p.grid_to_rect]), float(p.fb),
                               float(p.offset), (ctypes.c_double * 3)(*[float(v) for v in p.partner_center]), dv, cv)
    h, w = int(grid.height), int(grid.width)
    depth = np.empty((h, w), np.float32)
    conf = np.empty((h, w), np.float32)
    mask = np.empty((h, w), np.uint16)
    g = grid.to_c()
    o = (options or MultiStereoOptions()).to_c()
    _check(lib().oracle_combine_pairs(ctypes.byref(g), arr, len(pairs), ctypes.byref(o), _ptr(depth), _ptr(conf),
                                      _ptr(mask)))
    return depth, conf, mask


def reduce_mask(mask):
    mask = _c(mask, np.uint8)
    h, w = mask.shape
    out = np.empty(((h + 1) // 2, (w + 1) // 2), np.uint8)
    _check(lib().oracle_reduce_mask(_ptr(mask), w, h, _ptr(out)))
    return out


def automatic_levels(w, h, nd):
    return int(lib().oracle_automatic_levels(int(w), int(h), int(nd)))


def _volume(fn, left, right, dmin, nd, options, ranges):
    left = _c(left, np.uint16)
    right = _c(right, np.uint16)
    h, wl = left.shape
    wr = right.shape[1]
    o = (options or StereoOptions()).to_c()
    lo = num = None
    if ranges is not None:
        lo, num = _c(ranges[0], np.uint16), _c(ranges[1], np.uint16)
        cells = int(num.astype(np.int64).sum())
    else:
        cells = wl * h * align16(nd)
    out = np.empty(cells, np.uint16)
    _check(fn(_ptr(left), wl, h, _ptr(right), wr, int(dmin), int(nd), ctypes.byref(o), _ptr(lo), _ptr(num),
              _ptr(out), ctypes.c_uint64(cells)))
    return out


def cost_volume(left, right, dmin, nd, options=None, ranges=None):
    """sgm_cost_volume (stereo_matcher.cpp:1269-1278) via the row-pass dataflow."""
    return _volume(lib().oracle_cost_volume, left, right, dmin, nd, options, ranges)


def textbook_volume(left, right, dmin, nd, options=None, ranges=None):
    """sgm_cost_volume_reference (stereo_matcher.cpp:1280-1362)."""
    return _volume(lib().oracle_textbook_volume, left, right, dmin, nd, options, ranges)


def wta_row(sums, lo, num, right_shift=None, right_size=0):
    lo = _c(lo, np.int32)
    num = _c(num, np.uint16)
    sums = _c(sums, np.uint16)
    w = lo.shape[0]
    best = np.empty(w, np.uint16)
    cost = np.empty(w, np.uint16)
    second = np.empty(w, np.uint16)
    rc = rb = None
    if right_shift is not None:
        rc = np.full(right_size, 0xFFFF, np.uint16)
        rb = np.full(right_size, 0xFFFF, np.uint16)
    _check(lib().oracle_wta_row(_ptr(sums), w, _ptr(lo), _ptr(num), _ptr(best), _ptr(cost), _ptr(second), _ptr(rc),
                                _ptr(rb), ctypes.c_int64(right_shift or 0)))
    return best, cost, second, rc, rb


def ranges_from_coarse(coarse, width, height, fine_valid, partner_valid, partner_width, partner_sign, dmin, nd,
                       options=None, mirrored=False):
    coarse = _c(coarse, np.float32)
    hc, wc = coarse.shape
    fv = _c(fine_valid, np.uint8)
    pv = _c(partner_valid, np.uint8)
    lo = np.empty((height, width), np.uint16)
    num = np.empty((height, width), np.uint16)
    o = (options or StereoOptions()).to_c()
    _check(lib().oracle_ranges_from_coarse(_ptr(coarse), wc, hc, int(width), int(height), _ptr(fv), _ptr(pv),
                                           int(partner_width), int(partner_sign), int(dmin), int(nd), ctypes.byref(o),
                                           int(bool(mirrored)), _ptr(lo), _ptr(num)))
    return lo, num


def remove_speckles(d, min_size, max_diff):
    d = np.array(d, dtype=np.float32, copy=True, order="C")
    h, w = d.shape
    _check(lib().oracle_remove_speckles(_ptr(d), w, h, int(min_size), ctypes.c_float(max_diff)))
    return d


def median_filter_3x3(d):
    d = _c(d, np.float32)
    h, w = d.shape
    out = np.empty_like(d)
    _check(lib().oracle_median3x3(_ptr(d), w, h, _ptr(out)))
    return out


def blend(width, height, blend_px, tiles, confidence=True):
    """blend_tiles: `tiles` is a list of (x0, y0, core (c0, c1, r0, r1),
    disparity, confidence) with float32 maps of the tile's size."""
    n = len(tiles)
    ext = np.zeros((max(n, 1), 8), np.int32)
    ds, cs = [], []
    for i, (x0, y0, core, d, c) in enumerate(tiles):
        d = _c(d, np.float32)
        ext[i] = (x0, y0, d.shape[1], d.shape[0], *core)
        ds.append(d)
        cs.append(_c(c, np.float32) if confidence else None)
    dp = (ctypes.c_void_p * max(n, 1))(*[a.ctypes.data for a in ds])
    cp = (ctypes.c_void_p * max(n, 1))(*[a.ctypes.data for a in cs]) if confidence else None
    out_d = np.empty((height, width), np.float32)
    out_c = np.empty((height, width), np.float32) if confidence else None
    _check(lib().oracle_blend(width, height, blend_px, n, _ptr(ext), dp, cp, _ptr(out_d), _ptr(out_c)))
    return out_d, out_c


def lr_mismatch(disp, other, sign, max_diff):
    """1 where a finite disparity fails the left-right check against `other`
    (x_o = lround(x + sign * d))."""
    disp = _c(disp, np.float32)
    other = _c(other, np.float32)
    h, w = disp.shape
    out = np.empty((h, w), np.uint8)
    _check(lib().oracle_lr_mismatch(_ptr(disp), w, h, _ptr(other), other.shape[1], sign, max_diff, _ptr(out)))
    return out


def match_stereo(left, right, left_valid=None, right_valid=None, dmin=0, nd=64, options=None, threads=None):
    """match_stereo (stereo_matcher.cpp:1023-1095) with the level-0 integer outputs."""
    left = _c(left, np.uint16)
    right = _c(right, np.uint16)
    h, wl = left.shape
    wr = right.shape[1]
    lv = _c(left_valid, np.uint8)
    rv = _c(right_valid, np.uint8)
    o = (options or StereoOptions()).to_c()
    disp = np.empty((h, wl), np.float32)
    conf = np.empty((h, wl), np.float32)
    labels = np.empty((h, wl), np.int16)
    lr = np.empty((h, wl), np.uint8)
    info = CMatchInfo()
    nthreads = threads if threads is not None else (os.cpu_count() or 1)
    _check(lib().oracle_match(_ptr(left), wl, h, _ptr(right), wr, _ptr(lv), _ptr(rv), int(dmin), int(nd),
                              ctypes.byref(o), int(nthreads), _ptr(disp), _ptr(conf), _ptr(labels), _ptr(lr),
                              ctypes.byref(info)))
    return DisparityMap(disp, conf, info.disparity_min, info.num_disparities, info.levels, int(info.cells),
                        int(info.tile_pixels), info.tiles_level0, labels, lr)
