"""Python binding of libr3dg.so (the C ABI in include/r3dg.h).

Accepts numpy arrays (host memory) or torch tensors on the GPU (device
memory, used in place).  Loading fails loudly when the library is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

from ._abi import (R3DG_DEVICE, R3DG_FORMAT_U8X257, R3DG_HOST, STAGES, CImage, CMatchDebug, CMatchInfo, CPair,
                   CServerStats, DisparityMap, Status, StereoOptions, align16)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libr3dg.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "r3dg.h")

_lib = None


class R3dgError(RuntimeError):
    """r3d::Error analogue: a non-zero r3dg_status."""

    def __init__(self, status, message):
        super().__init__(f"{Status(status).name}: {message}")
        self.status = Status(status)


class Cancelled(R3dgError):
    pass


def declared_symbols() -> list[str]:
    """Every function declared in include/r3dg.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(r3dg_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libr3dg.so is not built ({LIB_PATH}); run `python -m regard3d_mvs_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name in declared_symbols():
            fn = getattr(L, name)
            fn.restype = ctypes.c_int32
        L.r3dg_default_options.restype = None
        _lib = L
    return _lib


def _raise(ctx, status):
    buf = ctypes.create_string_buffer(4096)
    lib().r3dg_last_error(ctx, buf, ctypes.c_size_t(len(buf)))
    msg = buf.value.decode(errors="replace")
    if status == Status.CANCELLED:
        raise Cancelled(status, msg)
    raise R3dgError(status, msg)


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _image(a, dtype, fmt=0) -> tuple[CImage, object]:
    """CImage view of a 2-D array; returns (view, keep-alive).  A torch GPU
    tensor is used in place: torch's current stream on its device is
    synchronised first, so that the library (on its own HIP stream) sees
    finished inputs and never writes an output block torch work still uses."""
    if a is None:
        return CImage(None, 0, 0, 0, R3DG_HOST, 0), None
    if _is_torch(a):
        import torch
        want = {np.uint8: torch.uint8, np.uint16: torch.uint16, np.float32: torch.float32,
                np.int16: torch.int16, np.uint64: torch.uint64}[dtype]
        if a.dtype != want:
            raise TypeError(f"expected {want}, got {a.dtype}")
        if a.dim() != 2 or a.stride(1) != 1:
            raise ValueError("expected a 2-D tensor with unit column stride")
        mem = R3DG_DEVICE if a.is_cuda else R3DG_HOST
        if a.is_cuda:
            torch.cuda.current_stream(a.device).synchronize()
        return CImage(ctypes.c_void_p(a.data_ptr()), a.shape[1], a.shape[0], a.stride(0), mem, fmt), a
    arr = np.asarray(a)
    if arr.dtype != np.dtype(dtype):
        arr = arr.astype(dtype)
    if arr.ndim != 2:
        raise ValueError("expected a 2-D array")
    if arr.strides[1] != arr.itemsize or arr.strides[0] % arr.itemsize != 0:
        arr = np.ascontiguousarray(arr)
    return CImage(arr.ctypes.data_as(ctypes.c_void_p), arr.shape[1], arr.shape[0], arr.strides[0] // arr.itemsize,
                  R3DG_HOST, fmt), arr


def _is_u8(a):
    return a is not None and str(a.dtype).split(".")[-1] == "uint8"


def _image16(a):
    """A uint16 image input; uint8 arrays are passed as bytes and widened to
    u8 * 257 on the device (R3DG_FORMAT_U8X257)."""
    if _is_u8(a):
        return _image(a, np.uint8, R3DG_FORMAT_U8X257)
    return _image(a, np.uint16)


class HostBuffer:
    """Page-locked host memory of the library's runtime (r3dg_host_alloc):
    numpy arrays over it move to and from the device by DMA, unstaged."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        _global_check(lib().r3dg_host_alloc(ctypes.c_size_t(max(int(nbytes), 1)), ctypes.byref(self.ptr)))
        self.nbytes = int(nbytes)

    def array(self, shape, dtype):
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        if n > self.nbytes:
            raise ValueError("host buffer too small")
        buf = (ctypes.c_char * max(n, 1)).from_address(self.ptr.value)
        a = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)
        return a

    def close(self):
        if self.ptr:
            lib().r3dg_host_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def host_empty(shape, dtype):
    """(array, owner): an uninitialised numpy array in page-locked memory; keep
    `owner` alive while the array is used."""
    b = HostBuffer(int(np.prod(shape)) * np.dtype(dtype).itemsize)
    return b.array(shape, dtype), b


def _like(ref, shape, dtype):
    """Output buffer in the same memory as `ref` (torch device tensor or numpy)."""
    if ref is not None and _is_torch(ref) and ref.is_cuda:
        import torch
        tdt = {np.uint8: torch.uint8, np.uint16: torch.uint16, np.float32: torch.float32, np.int16: torch.int16,
               np.uint64: torch.uint64}[dtype]
        return torch.empty(shape, dtype=tdt, device=ref.device)
    return np.empty(shape, dtype)


def _shape(a):
    return (int(a.shape[0]), int(a.shape[1]))


def _global_check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(4096)
        lib().r3dg_global_error(buf, ctypes.c_size_t(len(buf)))
        raise R3dgError(rc, buf.value.decode(errors="replace"))


# ---- file formats (io/pfm.hpp, mvs/depth_map.hpp); host only
def read_pfm(path):
    """read_pfm: (h, w) float32, or (h, w, 3) for a colour PFM."""
    w, h, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    p = os.fsencode(path)
    _global_check(lib().r3dg_pfm_header(p, ctypes.byref(w), ctypes.byref(h), ctypes.byref(c)))
    out = np.empty((h.value, w.value * c.value), np.float32)
    _global_check(lib().r3dg_read_pfm(p, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(out.shape[1])))
    return out if c.value == 1 else out.reshape(h.value, w.value, 3)


def write_pfm(path, img):
    """write_pfm of a (h, w) or (h, w, 3) float image."""
    a = np.ascontiguousarray(img, np.float32)
    ch = 1 if a.ndim == 2 else a.shape[2]
    _global_check(lib().r3dg_write_pfm(os.fsencode(path), a.ctypes.data_as(ctypes.c_void_p), a.shape[1], a.shape[0],
                                       ch, ctypes.c_int64(a.shape[1] * ch)))


def read_depth_map(path):
    """read_depth_map (.r3dd, version 2) into a DepthMapFile."""
    from ._abi import CDepthMap, DepthMapFile
    p = os.fsencode(path)
    m = CDepthMap()
    _global_check(lib().r3dg_read_depth_map_header(p, ctypes.byref(m)))
    h, w = m.height, m.width
    imgs = (np.empty((h, w), np.float32), np.empty((h, w), np.float32), np.empty((h, w), np.uint16),
            np.empty((h, w), np.uint8), np.empty((h, w, 3), np.uint8))
    m.depth, m.confidence, m.pair_mask, m.support, m.color = [a.ctypes.data for a in imgs]
    _global_check(lib().r3dg_read_depth_map(p, ctypes.byref(m)))
    return DepthMapFile(m.view, m.camera_width, m.camera_height, m.fx, m.fy, m.cx, m.cy, tuple(m.rotation),
                        tuple(m.center), tuple(m.partners[:m.num_partners]), *imgs)


def write_depth_map(path, dm):
    """write_depth_map (.r3dd, version 2) of a DepthMapFile."""
    from ._abi import CDepthMap
    h, w = np.asarray(dm.depth).shape
    imgs = [np.ascontiguousarray(dm.depth, np.float32), np.ascontiguousarray(dm.confidence, np.float32),
            np.ascontiguousarray(dm.pair_mask, np.uint16), np.ascontiguousarray(dm.support, np.uint8),
            np.ascontiguousarray(dm.color, np.uint8)]
    if imgs[4].shape != (h, w, 3) or any(a.shape != (h, w) for a in imgs[:4]):
        raise ValueError("write_depth_map: images of different sizes")
    if len(dm.partners) > 16:
        raise ValueError("write_depth_map: at most 16 partners")
    parts = list(dm.partners) + [0] * (16 - len(dm.partners))
    m = CDepthMap(int(dm.view), int(dm.camera_width), int(dm.camera_height), len(dm.partners), float(dm.fx),
                  float(dm.fy), float(dm.cx), float(dm.cy), (ctypes.c_double * 9)(*[float(v) for v in dm.rotation]),
                  (ctypes.c_double * 3)(*[float(v) for v in dm.center]), (ctypes.c_uint32 * 16)(*parts), w, h,
                  *[a.ctypes.data for a in imgs])
    _global_check(lib().r3dg_write_depth_map(os.fsencode(path), ctypes.byref(m)))


class Matcher:
    """One context on one GPU (r3dg_ctx); the methods mirror the reference API."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._ctx = ctypes.c_void_p()
        rc = lib().r3dg_ctx_create(int(device), ctypes.c_uint64(hbm_budget), ctypes.byref(self._ctx))
        if rc != 0:
            buf = ctypes.create_string_buffer(4096)
            lib().r3dg_global_error(buf, ctypes.c_size_t(len(buf)))
            raise R3dgError(rc, buf.value.decode(errors="replace"))

    def close(self):
        if self._ctx:
            lib().r3dg_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            _raise(self._ctx, rc)

    def cancel(self, requested=True):
        self._check(lib().r3dg_cancel(self._ctx, int(bool(requested))))

    def inject_alloc_failure(self, after):
        """Test hook: the device allocation after `after` more fails once
        (R3DG_ERROR_OUT_OF_MEMORY); negative disarms it."""
        self._check(lib().r3dg_inject_alloc_failure(self._ctx, ctypes.c_int64(int(after))))

    # census_transform (census.hpp:20)
    def census(self, img):
        v, keep = _image16(img)
        out = _like(img, _shape(img), np.uint64)
        o, keep2 = _image(out, np.uint64)
        self._check(lib().r3dg_census(self._ctx, v, o))
        return out

    # rectify_image<T> (rectify/rectification.hpp:80-82), one channel of uint8,
    # uint16 or float32: (rectified image, valid mask).
    def rectify(self, src, camera, target, want_valid=True):
        dts = {"uint8": (np.uint8, 0), "uint16": (np.uint16, 1), "float32": (np.float32, 2)}
        key = str(src.dtype).split(".")[-1]
        if key not in dts:
            raise TypeError(f"rectify: unsupported element type {src.dtype}")
        dt, code = dts[key]
        v, keep = _image(src, dt)
        shape = (int(target.height), int(target.width))
        out = _like(src, shape, dt)
        o, keep2 = _image(out, dt)
        valid = _like(src, shape, np.uint8) if want_valid else None
        vv, keep3 = _image(valid, np.uint8)
        cam, tgt = camera.to_c(), target.to_c()
        self._check(lib().r3dg_rectify(self._ctx, code, v, ctypes.byref(cam), ctypes.byref(tgt), o, vv))
        return out, valid

    # combine_pairs (mvs/multi_stereo.hpp:43-44): (depth, confidence, pair mask)
    # on the grid; `pairs` are PairGeometry records with their maps.
    def combine_pairs(self, grid, pairs, options=None):
        from ._abi import CPairGeometry, MultiStereoOptions
        arr = (CPairGeometry * max(len(pairs), 1))()
        keep = []
        for i, p in enumerate(pairs):
            dv, k1 = _image(p.disparity, np.float32)
            cv, k2 = _image(p.confidence, np.float32)
            keep += [k1, k2]
            arr[i] = CPairGeometry((ctypes.c_double * 9)(*[float(v) for v in p.grid_to_rect]), float(p.fb),
                                   float(p.offset), (ctypes.c_double * 3)(*[float(v) for v in p.partner_center]),
                                   dv, cv)
        shape = (int(grid.height), int(grid.width))
        ref = pairs[0].disparity if pairs else None
        depth, conf, mask = _like(ref, shape, np.float32), _like(ref, shape, np.float32), _like(ref, shape, np.uint16)
        d, k3 = _image(depth, np.float32)
        c, k4 = _image(conf, np.float32)
        m, k5 = _image(mask, np.uint16)
        g = grid.to_c()
        o = (options or MultiStereoOptions()).to_c()
        self._check(lib().r3dg_combine_pairs(self._ctx, ctypes.byref(g), arr, len(pairs), ctypes.byref(o), d, c, m))
        return depth, conf, mask

    # pyramid_down (pyramid.hpp:20)
    def pyramid_down(self, img):
        dt = np.uint8 if str(img.dtype).endswith("uint8") else np.uint16
        h, w = _shape(img)
        v, keep = _image(img, dt)
        out = _like(img, ((h + 1) // 2, (w + 1) // 2), dt)
        o, keep2 = _image(out, dt)
        fn = lib().r3dg_pyramid_down_u8 if dt == np.uint8 else lib().r3dg_pyramid_down_u16
        self._check(fn(self._ctx, v, o))
        return out

    # detail::sgm_cost_volume (stereo_matcher.hpp:98-101)
    def cost_volume(self, left, right, dmin, nd, options=None, ranges=None):
        o = (options or StereoOptions()).to_c()
        lv, k1 = _image16(left)
        rv, k2 = _image16(right)
        h, w = _shape(left)
        if ranges is not None:
            lo, num = ranges
            lov, k3 = _image(lo, np.uint16)
            nv, k4 = _image(num, np.uint16)
            cells = int(np.asarray(num).astype(np.int64).sum())
        else:
            lov, k3 = _image(None, np.uint16)
            nv, k4 = _image(None, np.uint16)
            cells = w * h * align16(int(nd))
        out = np.empty(cells, np.uint16)
        self._check(lib().r3dg_cost_volume(self._ctx, lv, rv, int(dmin), int(nd), ctypes.byref(o), lov, nv,
                                           out.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(cells), R3DG_HOST))
        return out

    # match_stereo (stereo_matcher.hpp:82-85)
    def match(self, left, right, left_valid=None, right_valid=None, dmin=0, nd=64, options=None, debug=False,
              out=None):
        o = (options or StereoOptions()).to_c()
        h, w = _shape(left)
        lv, k1 = _image16(left)
        rv, k2 = _image16(right)
        mlv, k3 = _image(left_valid, np.uint8)
        mrv, k4 = _image(right_valid, np.uint8)
        if out is not None:
            disp, conf = out
        else:
            disp = _like(left, (h, w), np.float32)
            conf = _like(left, (h, w), np.float32)
        dv, k5 = _image(disp, np.float32)
        cv, k6 = _image(conf, np.float32)
        dbg = None
        labels = lr = None
        if debug:
            labels = _like(left, (h, w), np.int16)
            lr = _like(left, (h, w), np.uint8)
            lab_v, k7 = _image(labels, np.int16)
            lr_v, k8 = _image(lr, np.uint8)
            dbg = CMatchDebug(lab_v, lr_v)
        info = CMatchInfo()
        self._check(lib().r3dg_match(self._ctx, lv, rv, mlv, mrv, int(dmin), int(nd), ctypes.byref(o), dv, cv,
                                     ctypes.byref(dbg) if dbg is not None else None, ctypes.byref(info)))
        return DisparityMap(disp, conf, info.disparity_min, info.num_disparities, info.levels, int(info.cells),
                            int(info.tile_pixels), info.tiles_level0, labels, lr, info.seconds,
                            int(info.aggregate_launches), int(info.aggregate_bytes),
                            {k: float(info.stage_seconds[i]) for i, k in enumerate(STAGES)})

    def match_batch(self, pairs, options=None, raise_errors=True):
        """r3dg_match_batch: `pairs` is a sequence of dicts with left, right and
        optional left_valid, right_valid, dmin (0), nd (64), out (disparity,
        confidence buffers).  Returns one
        DisparityMap per pair (None for a pair that failed when raise_errors
        is False)."""
        o = (options or StereoOptions()).to_c()
        n = len(pairs)
        arr = (CPair * max(n, 1))()
        keep, outs = [], []
        for i, q in enumerate(pairs):
            left = q["left"]
            h, w = _shape(left)
            if q.get("out") is not None:
                disp, conf = q["out"]
            else:
                disp = _like(left, (h, w), np.float32)
                conf = _like(left, (h, w), np.float32)
            views = [_image16(left), _image16(q["right"]), _image(q.get("left_valid"), np.uint8),
                     _image(q.get("right_valid"), np.uint8), _image(disp, np.float32), _image(conf, np.float32)]
            keep.append(views)
            outs.append((disp, conf))
            p = arr[i]
            p.left, p.right, p.left_valid, p.right_valid = (v[0] for v in views[:4])
            p.disparity, p.confidence = views[4][0], views[5][0]
            p.disparity_min = int(q.get("dmin", 0))
            p.num_disparities = int(q.get("nd", 64))
        rc = lib().r3dg_match_batch(self._ctx, arr, n, ctypes.byref(o))
        if rc != 0 and raise_errors:
            self._check(rc)
        res = []
        for i in range(n):
            info = arr[i].info
            if arr[i].status != 0:
                res.append(None)
                continue
            disp, conf = outs[i]
            res.append(DisparityMap(disp, conf, info.disparity_min, info.num_disparities, info.levels,
                                    int(info.cells), int(info.tile_pixels), info.tiles_level0, None, None,
                                    info.seconds, int(info.aggregate_launches), int(info.aggregate_bytes),
                                    {k: float(info.stage_seconds[j]) for j, k in enumerate(STAGES)}))
        return res

    # ---- serving: request loop, tickets, cached views
    def serve_start(self, options=None, hbm_budget=0):
        self._serve_opts = (options or StereoOptions()).to_c()
        self._tickets = {}
        self._check(lib().r3dg_server_start(self._ctx, ctypes.byref(self._serve_opts), ctypes.c_uint64(hbm_budget)))

    def serve_stop(self):
        self._check(lib().r3dg_server_stop(self._ctx))

    def inject_step_failure(self, after):
        """Test hook: in the next loop step with two or more pairs in flight,
        the device allocation after `after` more fails once."""
        self._check(lib().r3dg_server_inject_step_failure(self._ctx, ctypes.c_int64(int(after))))

    def server_stats(self) -> dict:
        """Counters of the running request loop (r3dg_server_stats_get)."""
        st = CServerStats()
        self._check(lib().r3dg_server_stats_get(self._ctx, ctypes.byref(st)))
        return st.as_dict()

    def view_put(self, view_id, image, valid=None):
        v, k1 = _image16(image)
        m, k2 = _image(valid, np.uint8)
        self._check(lib().r3dg_view_put(self._ctx, ctypes.c_uint64(view_id), v, m))

    def view_evict(self, view_id):
        self._check(lib().r3dg_view_evict(self._ctx, ctypes.c_uint64(view_id)))

    def submit(self, shape, dmin=0, nd=64, left=None, right=None, left_valid=None, right_valid=None, left_view=0,
               right_view=0):
        """Queues a pair; `shape` is (height, width) of the left image (the
        output size).  Images not given are taken from the cached views."""
        h, w = shape
        disp = np.empty((h, w), np.float32)
        conf = np.empty((h, w), np.float32)
        views = [_image16(left), _image16(right), _image(left_valid, np.uint8),
                 _image(right_valid, np.uint8), _image(disp, np.float32), _image(conf, np.float32)]
        pair = CPair()
        pair.left, pair.right, pair.left_valid, pair.right_valid = (v[0] for v in views[:4])
        pair.disparity, pair.confidence = views[4][0], views[5][0]
        pair.disparity_min = int(dmin)
        pair.num_disparities = int(nd)
        t = ctypes.c_uint64()
        self._check(lib().r3dg_submit(self._ctx, ctypes.byref(pair), ctypes.c_uint64(left_view),
                                      ctypes.c_uint64(right_view), ctypes.byref(t)))
        self._tickets[t.value] = (pair, views, disp, conf, left, right, left_valid, right_valid)
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """DisparityMap of a finished pair (info.seconds: submission to completion),
        None on timeout; raises R3dgError if the pair failed."""
        rc = lib().r3dg_wait(self._ctx, ctypes.c_uint64(ticket), ctypes.c_double(timeout))
        if rc == Status.PENDING:
            return None
        pair, views, disp, conf = self._tickets.pop(ticket)[:4]
        self._check(rc)
        info = pair.info
        return DisparityMap(disp, conf, info.disparity_min, info.num_disparities, info.levels, int(info.cells),
                            int(info.tile_pixels), info.tiles_level0, None, None, info.seconds,
                            int(info.aggregate_launches), int(info.aggregate_bytes),
                            {k: float(info.stage_seconds[j]) for j, k in enumerate(STAGES)})

    def set_profiling(self, enabled=True):
        self._check(lib().r3dg_set_profiling(self._ctx, int(bool(enabled))))

    # remove_speckles (stereo_matcher.hpp:89): returns a filtered copy
    def remove_speckles(self, d, min_size, max_diff):
        d = np.array(d, dtype=np.float32, copy=True, order="C")
        v, keep = _image(d, np.float32)
        self._check(lib().r3dg_remove_speckles(self._ctx, v, int(min_size), ctypes.c_float(max_diff)))
        return d

    # median_filter_3x3 (stereo_matcher.hpp:91)
    def median_filter_3x3(self, d):
        v, keep = _image(d, np.float32)
        out = _like(d, _shape(d), np.float32)
        o, keep2 = _image(out, np.float32)
        self._check(lib().r3dg_median_filter_3x3(self._ctx, v, o))
        return out

    # detail::ranges_from_coarse (stereo_matcher.hpp:121-125)
    def ranges_from_coarse(self, coarse, width, height, fine_valid, partner_valid, partner_width, partner_sign, dmin,
                           nd, options=None, mirrored=False):
        o = (options or StereoOptions()).to_c()
        cv, k1 = _image(coarse, np.float32)
        fv, k2 = _image(fine_valid, np.uint8)
        pv, k3 = _image(partner_valid, np.uint8)
        lo = np.empty((height, width), np.uint16)
        num = np.empty((height, width), np.uint16)
        lov, k4 = _image(lo, np.uint16)
        nv, k5 = _image(num, np.uint16)
        self._check(lib().r3dg_ranges_from_coarse(self._ctx, cv, int(width), int(height), fv, pv, int(partner_width),
                                                  int(partner_sign), int(dmin), int(nd), ctypes.byref(o),
                                                  int(bool(mirrored)), lov, nv))
        return lo, num


def automatic_levels(w, h, nd):
    """detail::automatic_levels (stereo_matcher.hpp:128)."""
    return int(lib().r3dg_automatic_levels(int(w), int(h), int(nd)))


def default_options() -> StereoOptions:
    return StereoOptions()
