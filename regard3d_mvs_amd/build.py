"""Builds libr3dg.so (gfx950 HIP kernels + host driver + C ABI) in-tree.

    python -m regard3d_mvs_amd.build

Objects go to regard3d_mvs_amd/_build/, the library to
regard3d_mvs_amd/libr3dg.so.  Flags: -ffp-contract=off keeps the float steps
(sub-pixel fit, blending) separate IEEE operations like the reference's x86
build; packed FP32 instructions are disabled for every kernel.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libr3dg.so")
API_LIB = os.path.join(HERE, "libr3d_stereo.so")          # r3d:: C++ API over the C ABI
API_TEST = os.path.join(ROOT, "tests", "cpp", "_build", "test_stereo_api")
SOURCES = ["kernels_image.hip", "kernels_sgm.hip", "kernels_rows.hip", "static_scratch.hip", "matcher.cpp"]
ARCH = os.environ.get("R3DG_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc()] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_api()
    if verbose:
        print(f"built {LIB}, {API_LIB}, {API_TEST}")
    return LIB


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _build_api() -> None:
    """libr3d_stereo.so (host C++ only, over libr3dg.so) and its C++ test program."""
    inc = os.path.join(ROOT, "include")
    cxx = ["g++", "-O2", "-std=c++20", "-fPIC", "-Wall", "-Wextra", "-I", inc]
    tmp = API_LIB + ".tmp"
    _run(cxx + ["-shared", "-o", tmp, os.path.join(CSRC, "api", "r3d_stereo.cpp"), "-L", HERE, "-lr3dg",
                "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, API_LIB)
    tdir = os.path.dirname(API_TEST)
    os.makedirs(tdir, exist_ok=True)
    cpp = os.path.join(ROOT, "tests", "cpp")
    oracle_src = [os.path.join(ROOT, "oracle", "oracle.cpp")]
    _run(cxx + ["-ffp-contract=off", "-pthread", "-o", API_TEST + ".tmp", os.path.join(cpp, "test_stereo_api.cpp"),
                os.path.join(cpp, "reference_shim.cpp")] + oracle_src +
         ["-L", HERE, "-lr3d_stereo", "-lr3dg", f"-Wl,-rpath,{HERE}"])
    os.replace(API_TEST + ".tmp", API_TEST)


def build_sanitized() -> str:
    """Host AddressSanitizer + UndefinedBehaviorSanitizer builds of
    libr3d_stereo.so and the C++ API test program (the HIP library itself is
    the regular one: no GPU sanitizer is used).  Returns the test program."""
    inc = os.path.join(ROOT, "include")
    san = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    cxx = ["g++", "-std=c++20", "-fPIC", "-Wall", "-Wextra", "-I", inc] + san
    tdir = os.path.dirname(API_TEST)
    os.makedirs(tdir, exist_ok=True)
    api = os.path.join(tdir, "libr3d_stereo_asan.so")
    _run(cxx + ["-shared", "-o", api, os.path.join(CSRC, "api", "r3d_stereo.cpp"), "-L", HERE, "-lr3dg",
                f"-Wl,-rpath,{HERE}"])
    cpp = os.path.join(ROOT, "tests", "cpp")
    out = API_TEST + "_asan"
    _run(cxx + ["-ffp-contract=off", "-pthread", "-o", out, os.path.join(cpp, "test_stereo_api.cpp"),
                os.path.join(cpp, "reference_shim.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp"),
                api, "-L", HERE, "-lr3dg", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN"])
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
