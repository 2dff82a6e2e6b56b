"""MI355X-native Census-SGM stereo matcher (drop-in for r3d::match_stereo)."""
from ._abi import (Camera, CameraModel, DepthMapFile, DisparityMap, GridCamera, MultiStereoOptions, PairGeometry,  # noqa: F401
                   StereoOptions, SubpixelMethod, align16)
