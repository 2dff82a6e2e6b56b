"""Writes the golden fixtures of tests/golden/: small seeded pairs with the
oracle's outputs, as raw little-endian arrays (.u8/.u16/.i16/.u64) and PFM
maps (.pfm), described by manifest.json.

The reference ships no golden vectors (all of its test inputs are
procedural), so these fixtures are the oracle's outputs, committed so that
the oracle and the GPU library can be checked against fixed bytes: the
oracle must reproduce them (tests/test_golden.py, CPU) and so must the GPU
library (tests/test_gpu_golden.py).  Regenerate only when the oracle's
definition changes:

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from regard3d_mvs_amd import StereoOptions, SubpixelMethod, native, synth  # noqa: E402

SUFFIX = {np.dtype(np.uint8): "u8", np.dtype(np.uint16): "u16", np.dtype(np.int16): "i16",
          np.dtype(np.uint64): "u64"}


def save(case, name, a, files):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        path = f"{case}.{name}.pfm"
        native.write_pfm(os.path.join(HERE, path), a)
    else:
        path = f"{case}.{name}.{SUFFIX[a.dtype]}"
        a.astype(a.dtype.newbyteorder("<")).tofile(os.path.join(HERE, path))
    files[name] = {"file": path, "shape": list(a.shape), "dtype": str(a.dtype)}


def options_dict(o: StereoOptions) -> dict:
    return {k: (v.value if hasattr(v, "value") else v) for k, v in vars(o).items()}


def main():
    cases = []

    # 1. Full range, one level: Census, the aggregated volume and the match.
    p = synth.make_pair(64, 40, 32, 11)
    o = StereoOptions(levels=0)
    files = {}
    save("full_range", "left", p.left, files)
    save("full_range", "right", p.right, files)
    save("full_range", "census_left", O.census(p.left), files)
    save("full_range", "census_right", O.census(p.right), files)
    save("full_range", "volume", O.cost_volume(p.left, p.right, 0, 32, o), files)
    g = O.match_stereo(p.left, p.right, None, None, 0, 32, o, threads=1)
    for k in ("disparity", "confidence", "labels", "lr_mask"):
        save("full_range", k, getattr(g, k), files)
    cases.append({"name": "full_range", "dmin": 0, "nd": 32, "options": options_dict(o), "cells": g.cells,
                  "levels": g.levels, "files": files})

    # 2. Two coarser levels with masks on both images.
    p = synth.make_pair(160, 96, 48, 12)
    lv = np.full(p.left.shape, 255, np.uint8)
    lv[:, :9] = 0
    lv[40:50, 100:120] = 0
    rv = np.full(p.right.shape, 255, np.uint8)
    rv[:, -7:] = 0
    rv[10:20, 30:60] = 90
    o = StereoOptions(levels=2)
    files = {}
    for k, a in (("left", p.left), ("right", p.right), ("left_valid", lv), ("right_valid", rv)):
        save("masked_levels", k, a, files)
    g = O.match_stereo(p.left, p.right, lv, rv, 0, 48, o, threads=1)
    for k in ("disparity", "confidence", "labels", "lr_mask"):
        save("masked_levels", k, getattr(g, k), files)
    cases.append({"name": "masked_levels", "dmin": 0, "nd": 48, "options": options_dict(o), "cells": g.cells,
                  "levels": g.levels, "files": files})

    # 3. Per-pixel label ranges: the volume of detail::sgm_cost_volume.
    p = synth.make_pair(48, 24, 32, 13)
    rng = np.random.default_rng(13)
    num = (rng.integers(1, 3, p.left.shape) * 16).astype(np.uint16)
    lo = rng.integers(0, 48 - num + 1).astype(np.uint16)
    o = StereoOptions(levels=0, p2_adapt=0)
    files = {}
    for k, a in (("left", p.left), ("right", p.right), ("lo", lo), ("num", num)):
        save("ranges", k, a, files)
    save("ranges", "volume", O.cost_volume(p.left, p.right, 4, 48, o, ranges=(lo, num)), files)
    cases.append({"name": "ranges", "dmin": 4, "nd": 48, "options": options_dict(o), "files": files})

    # 4. Unequal widths, parabola fit, the diagonal left-right check.
    p = synth.make_pair(80, 32, 24, 14)
    right = np.ascontiguousarray(p.right[:, :70])
    o = StereoOptions(levels=0, subpixel=SubpixelMethod.parabola, lr_full_match=False, uniqueness=0.1)
    files = {}
    save("diagonal_lr", "left", p.left, files)
    save("diagonal_lr", "right", right, files)
    g = O.match_stereo(p.left, right, None, None, -2, 24, o, threads=1)
    for k in ("disparity", "confidence", "labels", "lr_mask"):
        save("diagonal_lr", k, getattr(g, k), files)
    cases.append({"name": "diagonal_lr", "dmin": -2, "nd": 24, "options": options_dict(o), "cells": g.cells,
                  "levels": g.levels, "files": files})

    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_fixtures.py (oracle outputs; regenerate only when the oracle "
                                "changes)", "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
