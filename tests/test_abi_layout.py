"""The ctypes structures of the Python binding have the sizes and field
offsets of the structs in include/r3dg.h, as the host C compiler lays them
out (no GPU needed)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from regard3d_mvs_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = [("r3dg_image", _abi.CImage), ("r3dg_options", _abi.COptions), ("r3dg_match_info", _abi.CMatchInfo),
           ("r3dg_match_debug", _abi.CMatchDebug), ("r3dg_pair", _abi.CPair),
           ("r3dg_server_stats", _abi.CServerStats), ("r3dg_camera", _abi.CCamera),
           ("r3dg_pair_geometry", _abi.CPairGeometry), ("r3dg_grid", _abi.CGrid),
           ("r3dg_multi_options", _abi.CMultiOptions), ("r3dg_depth_map", _abi.CDepthMap)]


def test_struct_layouts_match_the_header(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "r3dg.h"', "int main(void) {"]
    for cname, ct in STRUCTS:
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in ct._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        name, field, value = line.split()
        got[(name, field)] = int(value)
    for cname, ct in STRUCTS:
        assert got[(cname, "sizeof")] == ctypes.sizeof(ct), cname
        for f, _ in ct._fields_:
            assert got[(cname, f)] == getattr(ct, f).offset, (cname, f)
