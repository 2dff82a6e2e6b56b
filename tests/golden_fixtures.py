"""Loader of the golden fixtures in tests/golden/ (see make_fixtures.py there)."""
import json
import os

import numpy as np

from regard3d_mvs_amd import StereoOptions, SubpixelMethod, native

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cases():
    with open(os.path.join(DIR, "manifest.json")) as f:
        return json.load(f)["cases"]


def load(case, name):
    e = case["files"][name]
    path = os.path.join(DIR, e["file"])
    if e["file"].endswith(".pfm"):
        return native.read_pfm(path)
    dt = np.dtype(e["dtype"]).newbyteorder("<")
    return np.fromfile(path, dtype=dt).astype(np.dtype(e["dtype"])).reshape(e["shape"])


def options(case):
    o = dict(case["options"])
    o["subpixel"] = SubpixelMethod(o["subpixel"])
    return StereoOptions(**o)


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape and a.dtype.itemsize == b.dtype.itemsize
    u = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize]
    bad = a.view(u) != b.view(u)
    assert int(bad.sum()) == 0, f"{int(bad.sum())} of {a.size} differ"
