"""The C-ABI library without a GPU: it loads, exports every function that
include/*.h declares, and its argument checks and device probe fail
cleanly (no compute call is made here)."""
from __future__ import annotations

import ctypes
import subprocess

import pytest

from conftest import gpu_available


def test_library_exports_every_header_symbol(shk_lib):
    from shiokaze_amd import _lib
    declared = set(_lib.header_symbols())
    assert len(declared) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert declared <= exported, sorted(declared - exported)
    for name in declared:
        getattr(shk_lib, name)  # ctypes resolves every declared symbol


def test_every_declared_symbol_has_a_signature():
    from shiokaze_amd import _lib
    assert set(_lib.header_symbols()) <= set(_lib._SIGNATURES), sorted(set(_lib.header_symbols())
                                                                       - set(_lib._SIGNATURES))


def test_version_and_build_info(shk_lib):
    assert shk_lib.shk_abi_version() >= 1
    info = shk_lib.shk_build_info().decode()
    assert "gfx950" in info


def test_argument_checks(shk_lib):
    from shiokaze_amd import _lib
    ctx = ctypes.c_void_p()
    assert shk_lib.shk_create(0, 8, 8, 0.1, 0, ctypes.byref(ctx)) == _lib.SHK_BAD_ARG
    assert shk_lib.shk_create(8, 8, 8, -1.0, 0, ctypes.byref(ctx)) == _lib.SHK_BAD_ARG
    assert shk_lib.shk_create(8, 8, 8, 0.1, 0, None) == _lib.SHK_BAD_ARG
    assert shk_lib.shk_configure(None, b"Residual", b"1e-6") == _lib.SHK_BAD_ARG
    assert shk_lib.shk_project(None, 1.0, *([None] * 8), 0.0, None) == _lib.SHK_BAD_ARG
    assert shk_lib.shk_unet_create(0, 0, None) == _lib.SHK_BAD_ARG
    assert shk_lib.shk_unet_num_params(None, None) == _lib.SHK_BAD_ARG
    assert shk_lib.shk_export_scaled(None, *([None] * 6)) == _lib.SHK_BAD_ARG
    # unsupported channel counts are refused before any device work
    assert shk_lib.shk_nn_conv3d(8, 8, 8, ctypes.c_void_p(16), 12, ctypes.c_void_p(16), None, 16, 0,
                                 ctypes.c_void_p(16)) == _lib.SHK_BAD_ARG
    assert shk_lib.shk_nn_conv3d(8, 8, 8, ctypes.c_void_p(16), 16, ctypes.c_void_p(16), None, 80, 0,
                                 ctypes.c_void_p(16)) == _lib.SHK_BAD_ARG


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_no_device_is_reported(shk_lib):
    from shiokaze_amd import _lib
    ctx = ctypes.c_void_p()
    assert shk_lib.shk_create(8, 8, 8, 0.1, 0, ctypes.byref(ctx)) == _lib.SHK_NO_DEVICE
    net = ctypes.c_void_p()
    assert shk_lib.shk_unet_create(0, 0, ctypes.byref(net)) == _lib.SHK_NO_DEVICE


def test_python_mirror_fails_loudly_without_device():
    import shiokaze_amd
    if gpu_available():
        pytest.skip("checks the no-device path")
    with pytest.raises(shiokaze_amd.ShkError):
        shiokaze_amd.MacPressureSolver3((8, 8, 8), 0.125)
    with pytest.raises(shiokaze_amd.ShkError):
        shiokaze_amd.UNet()
