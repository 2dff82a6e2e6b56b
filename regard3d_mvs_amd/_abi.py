"""ctypes mirror of the POD types in include/r3dg.h (C ABI of the GPU matcher).

Pure data definitions: no library is loaded here.
"""
from __future__ import annotations

import ctypes
import dataclasses
import enum

R3DG_HOST = 0
R3DG_DEVICE = 1
R3DG_FORMAT_NATIVE = 0
R3DG_FORMAT_U8X257 = 1


class Status(enum.IntEnum):
    OK = 0
    INVALID_ARGUMENT = 1
    OUT_OF_MEMORY = 2
    DEVICE = 3
    CANCELLED = 4
    BUDGET = 5
    NOT_FOUND = 6
    INTERNAL = 7
    PENDING = 8


class SubpixelMethod(enum.IntEnum):
    """r3d::SubpixelMethod (stereo_matcher.hpp:23-27)."""
    none = 0
    parabola = 1
    equiangular = 2


class CImage(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("stride", ctypes.c_int64), ("memory", ctypes.c_int32), ("format", ctypes.c_int32)]


class COptions(ctypes.Structure):
    _fields_ = [("p1", ctypes.c_int32), ("p2", ctypes.c_int32), ("p2_adapt", ctypes.c_int32),
                ("invalid_cost", ctypes.c_int32), ("subpixel", ctypes.c_int32), ("uniqueness", ctypes.c_float),
                ("lr_max_diff", ctypes.c_int32), ("lr_full_match", ctypes.c_int32),
                ("speckle_size", ctypes.c_int32), ("speckle_diff", ctypes.c_float), ("median", ctypes.c_int32),
                ("levels", ctypes.c_int32), ("range_radius", ctypes.c_int32), ("range_margin", ctypes.c_int32),
                ("range_fill", ctypes.c_int32), ("tile_size", ctypes.c_int32), ("tile_margin", ctypes.c_int32),
                ("tile_blend", ctypes.c_int32), ("memory_budget", ctypes.c_uint64)]


class CameraModel(enum.IntEnum):
    """r3d::CameraModel (core/camera_model.hpp:13-20)."""
    pinhole = 0
    radial_k1 = 1
    radial_k3 = 2
    brown_t2 = 3
    fisheye = 4


class CCamera(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("fx", ctypes.c_double), ("fy", ctypes.c_double),
                ("cx", ctypes.c_double), ("cy", ctypes.c_double), ("params", ctypes.c_double * 5),
                ("rotation", ctypes.c_double * 9), ("center", ctypes.c_double * 3)]


@dataclasses.dataclass
class Camera:
    """CameraIntrinsics with its Pose (core/camera_model.hpp:32-41, core/camera.hpp:9-12):
    rotation maps world to camera axes (row-major 3x3), centre in world
    coordinates.  A PinholeCamera is a Camera of model pinhole."""
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    model: CameraModel = CameraModel.pinhole
    params: tuple = (0.0, 0.0, 0.0, 0.0, 0.0)
    rotation: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    center: tuple = (0.0, 0.0, 0.0)

    def to_c(self) -> CCamera:
        def flat(v, n):
            out = []
            for e in v:
                if hasattr(e, "__len__"):
                    out.extend(float(x) for x in e)
                else:
                    out.append(float(e))
            if len(out) != n:
                raise ValueError(f"expected {n} values, got {len(out)}")
            return out
        params = [float(v) for v in self.params]
        if len(params) > 5:
            raise ValueError("at most 5 distortion parameters")
        params += [0.0] * (5 - len(params))
        return CCamera(int(self.model), int(self.width), int(self.height), 0, float(self.fx), float(self.fy),
                       float(self.cx), float(self.cy), (ctypes.c_double * 5)(*params),
                       (ctypes.c_double * 9)(*flat(self.rotation, 9)), (ctypes.c_double * 3)(*flat(self.center, 3)))


class CPairGeometry(ctypes.Structure):
    _fields_ = [("grid_to_rect", ctypes.c_double * 9), ("fb", ctypes.c_double), ("offset", ctypes.c_double),
                ("partner_center", ctypes.c_double * 3), ("disparity", CImage), ("confidence", CImage)]


class CGrid(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("inv_k", ctypes.c_double * 9),
                ("rot_t", ctypes.c_double * 9), ("center", ctypes.c_double * 3)]


class CMultiOptions(ctypes.Structure):
    _fields_ = [("inlier_px", ctypes.c_double), ("max_reprojection_px", ctypes.c_double),
                ("cos_min_angle", ctypes.c_double), ("min_pairs", ctypes.c_int32), ("reserved", ctypes.c_int32)]


@dataclasses.dataclass
class MultiStereoOptions:
    """r3d::MultiStereoOptions (mvs/multi_stereo.hpp:25-30)."""
    inlier_px: float = 2.0
    max_reprojection_px: float = 1.0
    min_angle_deg: float = 3.0
    min_pairs: int = 1

    def to_c(self) -> CMultiOptions:
        import math
        return CMultiOptions(float(self.inlier_px), float(self.max_reprojection_px),
                             math.cos(self.min_angle_deg * math.pi / 180.0), int(self.min_pairs), 0)


@dataclasses.dataclass
class GridCamera:
    """The depth map's camera as combine_pairs uses it: K^-1 and R^T
    (row-major) and the centre."""
    width: int
    height: int
    inv_k: tuple
    rot_t: tuple
    center: tuple

    def to_c(self) -> CGrid:
        return CGrid(int(self.width), int(self.height), (ctypes.c_double * 9)(*[float(v) for v in self.inv_k]),
                     (ctypes.c_double * 9)(*[float(v) for v in self.rot_t]),
                     (ctypes.c_double * 3)(*[float(v) for v in self.center]))


@dataclasses.dataclass
class PairGeometry:
    """PairGeometry (mvs/multi_stereo.cpp:26-33) with the pair's maps."""
    grid_to_rect: tuple
    fb: float
    offset: float
    partner_center: tuple
    disparity: object
    confidence: object = None


class CDepthMap(ctypes.Structure):
    _fields_ = [("view", ctypes.c_uint32), ("camera_width", ctypes.c_int32), ("camera_height", ctypes.c_int32),
                ("num_partners", ctypes.c_uint32), ("fx", ctypes.c_double), ("fy", ctypes.c_double),
                ("cx", ctypes.c_double), ("cy", ctypes.c_double), ("rotation", ctypes.c_double * 9),
                ("center", ctypes.c_double * 3), ("partners", ctypes.c_uint32 * 16), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("depth", ctypes.c_void_p), ("confidence", ctypes.c_void_p),
                ("pair_mask", ctypes.c_void_p), ("support", ctypes.c_void_p), ("color", ctypes.c_void_p)]


@dataclasses.dataclass
class DepthMapFile:
    """The contents of a .r3dd file (mvs/depth_map.hpp:16-45): view, grid camera
    (intrinsics of camera_width x camera_height, row-major rotation, centre),
    partner views and the five images of width x height."""
    view: int
    camera_width: int
    camera_height: int
    fx: float
    fy: float
    cx: float
    cy: float
    rotation: tuple
    center: tuple
    partners: tuple
    depth: object       # float32 (h, w)
    confidence: object  # float32 (h, w)
    pair_mask: object   # uint16 (h, w)
    support: object     # uint8 (h, w)
    color: object       # uint8 (h, w, 3)


STAGES = ("pyramid", "prepare", "ranges", "layout", "aggregate", "select", "blend", "filter")


class CMatchInfo(ctypes.Structure):
    _fields_ = [("disparity_min", ctypes.c_int32), ("num_disparities", ctypes.c_int32),
                ("levels", ctypes.c_int32), ("tiles_level0", ctypes.c_int32), ("cells", ctypes.c_uint64),
                ("tile_pixels", ctypes.c_uint64), ("seconds", ctypes.c_double),
                ("aggregate_launches", ctypes.c_uint64), ("aggregate_bytes", ctypes.c_uint64),
                ("stage_seconds", ctypes.c_double * len(STAGES))]


class CMatchDebug(ctypes.Structure):
    _fields_ = [("labels", CImage), ("lr_mask", CImage)]


class CServerStats(ctypes.Structure):
    """r3dg_server_stats: counters of the running request loop."""
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "completed", "failed", "steps", "requeued", "pending",
                                                "active", "in_use", "view_bytes")] + [
        (k, ctypes.c_double) for k in ("admission_scale", "loop_seconds", "wait_seconds")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CPair(ctypes.Structure):
    """r3dg_pair: one pair of r3dg_match_batch."""
    _fields_ = [("left", CImage), ("right", CImage), ("left_valid", CImage), ("right_valid", CImage),
                ("disparity_min", ctypes.c_int32), ("num_disparities", ctypes.c_int32), ("disparity", CImage),
                ("confidence", CImage), ("info", CMatchInfo), ("status", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


@dataclasses.dataclass
class StereoOptions:
    """r3d::StereoOptions (stereo_matcher.hpp:30-57) with the same defaults."""
    p1: int = 28
    p2: int = 200
    p2_adapt: int = 4
    invalid_cost: int = 31
    subpixel: SubpixelMethod = SubpixelMethod.equiangular
    uniqueness: float = 0.0
    lr_max_diff: int = 1
    lr_full_match: bool = True
    speckle_size: int = 100
    speckle_diff: float = 1.0
    median: bool = True
    levels: int = -1
    range_radius: int = 3
    range_margin: int = 4
    range_fill: int = 64
    tile_size: int = 512
    tile_margin: int = 32
    tile_blend: int = 32
    memory_budget: int = 2 << 30

    def to_c(self) -> COptions:
        return COptions(int(self.p1), int(self.p2), int(self.p2_adapt), int(self.invalid_cost), int(self.subpixel),
                        float(self.uniqueness), int(self.lr_max_diff), int(bool(self.lr_full_match)),
                        int(self.speckle_size), float(self.speckle_diff), int(bool(self.median)), int(self.levels),
                        int(self.range_radius), int(self.range_margin), int(self.range_fill), int(self.tile_size),
                        int(self.tile_margin), int(self.tile_blend), int(self.memory_budget))


@dataclasses.dataclass
class DisparityMap:
    """r3d::DisparityMap (stereo_matcher.hpp:61-68) plus the level-0 integer outputs."""
    disparity: object
    confidence: object
    disparity_min: int
    num_disparities: int
    levels: int
    cells: int
    tile_pixels: int = 0
    tiles_level0: int = 0
    labels: object = None
    lr_mask: object = None
    seconds: float = 0.0
    aggregate_launches: int = 0
    aggregate_bytes: int = 0
    stage_seconds: dict = dataclasses.field(default_factory=dict)


def align16(n: int) -> int:
    return (n + 15) // 16 * 16
