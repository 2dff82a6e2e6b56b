"""The oracle under host AddressSanitizer + UndefinedBehaviorSanitizer
(make -C oracle sanitize: every entry point on small inputs, odd widths,
ranges, masks, levels, both left-right checks, filters, rectification and
combine_pairs) runs without a report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.slow
def test_oracle_under_asan_ubsan():
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "sanitize_check: ok" in r.stdout
