"""The C ABI library loads and exports every symbol include/r3dg.h declares
(no GPU needed: no compute call is made)."""
import ctypes
import os

import pytest

from regard3d_mvs_amd import native


def test_header_declares_the_boundary():
    syms = native.declared_symbols()
    for s in ("r3dg_ctx_create", "r3dg_ctx_destroy", "r3dg_census", "r3dg_cost_volume", "r3dg_match",
              "r3dg_last_error", "r3dg_ranges_from_coarse", "r3dg_remove_speckles", "r3dg_median_filter_3x3"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(native.LIB_PATH):
        pytest.fail("libr3dg.so is not built; run `python -m regard3d_mvs_amd.build`")
    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [s for s in native.declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_host_only_entry_points():
    lib = native.lib()
    assert lib.r3dg_abi_version() == 3
    assert native.automatic_levels(6048, 4032, 512) == 4
    assert native.automatic_levels(640, 480, 64) == 1
    from regard3d_mvs_amd._abi import COptions, StereoOptions
    o = COptions()
    lib.r3dg_default_options(ctypes.byref(o))
    d = StereoOptions().to_c()
    for name, _ in COptions._fields_:
        assert getattr(o, name) == getattr(d, name), name


def test_null_context_is_an_error():
    lib = native.lib()
    img = native.CImage(None, 0, 0, 0, 0, 0)
    assert lib.r3dg_census(None, img, img) == 1
    buf = ctypes.create_string_buffer(256)
    assert lib.r3dg_global_error(buf, ctypes.c_size_t(256)) == 1
    assert b"null context" in buf.value
