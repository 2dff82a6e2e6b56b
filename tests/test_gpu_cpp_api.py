"""The r3d C++ drop-in API (libr3d_stereo.so over the C ABI) on the GPU,
exercised by a C++ program written like the reference's doctest tests."""
import os
import subprocess

import pytest

from regard3d_mvs_amd import build

pytestmark = pytest.mark.gpu


def test_cpp_api_program():
    if not os.path.exists(build.API_TEST):
        pytest.fail(f"{build.API_TEST} is not built; run `python -m regard3d_mvs_amd.build`")
    r = subprocess.run([build.API_TEST], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
