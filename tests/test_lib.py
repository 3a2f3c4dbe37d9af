"""libpxl.so without a GPU: it loads, exports every symbol include/pxl.h
declares, validates arguments before touching the device, and its host-side
camera maths equals the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "pxl.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pxl_\w+)\s*\(", text)))


def test_exports_every_declared_symbol(pxl_lib):
    syms = declared_symbols()
    assert len(syms) >= 19
    assert set(syms) == set(pxl_lib.EXPORTS)
    L = ctypes.CDLL(pxl_lib.LIB_PATH)
    for s in syms:
        assert hasattr(L, s), s


def test_version_and_error_string(pxl_lib):
    assert pxl_lib.lib().pxl_version() >> 16 == 1
    assert pxl_lib.lib().pxl_abi_sizes(None, 0) == 9
    assert isinstance(pxl_lib.lib().pxl_last_error(), bytes)


def test_struct_layouts_match_header(pxl_lib):
    """ctypes mirrors of the C structs have the C sizes."""
    assert ctypes.sizeof(pxl_lib.Camera) == 24 * 8
    assert ctypes.sizeof(pxl_lib.BBox) == 24
    assert ctypes.sizeof(pxl_lib.RaysC) == 17 * 8 + 8 + 8
    assert ctypes.sizeof(pxl_lib.GridC) == 24 + 8 + 6 * 8 + 24 + 16
    assert ctypes.sizeof(pxl_lib.TrainConfigC) == 8 + 4 * 8 + 4 * 4 + 8 + 4 * 2 + 4 * 8
    assert ctypes.sizeof(pxl_lib.LevelConfigC) == 40
    assert ctypes.sizeof(pxl_lib.LevelC) == 24
    assert ctypes.sizeof(pxl_lib.ScheduleC) == 16
    assert ctypes.sizeof(pxl_lib.OptimC) == 56


def test_argument_validation_without_device(pxl_lib):
    L = pxl_lib.lib()
    g = pxl_lib.GridC()
    g.W = g.H = g.D = 8
    g.sh_dim = 7  # unsupported
    g.stride = 32
    g.n_vox = 512
    g.params = 16  # never dereferenced: validation fails first
    g.step = 0.5
    r = pxl_lib.RaysC()
    for k in ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t"):
        setattr(r, k, 16)
    si = ctypes.c_void_p(16)
    assert L.pxl_render_fwd(ctypes.byref(g), ctypes.byref(r), si, 0, 4, None) == pxl_lib.PXL_EINVAL
    assert b"sh_dim" in L.pxl_last_error()
    g.sh_dim = 9
    g.stride = 30  # not a multiple of 4
    assert L.pxl_render_fwd(ctypes.byref(g), ctypes.byref(r), si, 0, 4, None) == pxl_lib.PXL_EINVAL
    g.stride = 32
    g.sparse = 1  # no occupancy / index
    assert L.pxl_render_fwd(ctypes.byref(g), ctypes.byref(r), si, 0, 4, None) == pxl_lib.PXL_EINVAL
    assert L.pxl_render_fwd(None, ctypes.byref(r), si, 0, 4, None) == pxl_lib.PXL_EINVAL
    # scratch too small -> ENOMEM
    io = ctypes.c_void_p(256)
    assert L.pxl_build_sparse_index(io, 64, 64, 64, io, None, io, None, 0, None) == pxl_lib.PXL_ENOMEM
    assert L.pxl_sparse_index_scratch_bytes(64 ** 3) >= (64 ** 3 // 4096) * 4
    assert L.pxl_adam_step(None, None, 0.1, 0.1, 1, None) == pxl_lib.PXL_EINVAL


def test_render_kernel_name(pxl_lib):
    L = pxl_lib.lib()
    g = pxl_lib.GridC()
    g.sh_dim, g.sparse, g.storage = 9, 1, pxl_lib.STORE_F32
    assert pxl_lib.get_option(pxl_lib.OPT_MARCHER) == pxl_lib.MARCHER_GROUP
    assert L.pxl_render_kernel_name(ctypes.byref(g), 2) == b"k_march8<9,sparse,fused,f32>"
    g.storage = pxl_lib.STORE_BF16
    assert L.pxl_render_kernel_name(ctypes.byref(g), 0) == b"k_march8<9,sparse,fwd,bf16>"
    g.sh_dim, g.sparse = 16, 0
    assert L.pxl_render_kernel_name(ctypes.byref(g), 1) == b"k_render<16,dense,fixed,bwd,bf16>"
    g.sh_dim, g.march_mode, g.storage = 9, pxl_lib.MARCH_REF_ADAPTIVE, pxl_lib.STORE_F32
    assert L.pxl_render_kernel_name(ctypes.byref(g), 2) == b"k_render<9,dense,adaptive,fused,f32>"
    g.march_mode = pxl_lib.MARCH_FIXED
    pxl_lib.set_option(pxl_lib.OPT_MARCHER, pxl_lib.MARCHER_WAVE)
    try:
        assert L.pxl_render_kernel_name(ctypes.byref(g), 2) == b"k_render<9,dense,fixed,fused,f32>"
        assert L.pxl_render_kernel_name(ctypes.byref(g), 3) == b"k_render<9,dense,fixed,fused,f32>"
    finally:
        pxl_lib.set_option(pxl_lib.OPT_MARCHER, pxl_lib.MARCHER_GROUP)
    assert L.pxl_set_option(pxl_lib.OPT_MARCHER, 7) == pxl_lib.PXL_EINVAL
    assert L.pxl_set_option(99, 0) == pxl_lib.PXL_EINVAL
    assert pxl_lib.get_option(pxl_lib.OPT_CLAIM_GRID) == 0
    assert L.pxl_set_option(pxl_lib.OPT_CLAIM_GRID, -1) == pxl_lib.PXL_EINVAL
    pxl_lib.set_option(pxl_lib.OPT_CLAIM_GRID, 3)
    try:
        assert pxl_lib.get_option(pxl_lib.OPT_CLAIM_GRID) == 3
    finally:
        pxl_lib.set_option(pxl_lib.OPT_CLAIM_GRID, 0)
    assert (L.pxl_render_kernel_name(ctypes.byref(g), 3) ==
            b"k_march8<9,dense,log,f32>+k_log_grad<9>")
    assert L.pxl_render_kernel_name(ctypes.byref(g), 4) is None
    assert L.pxl_render_kernel_name(None, 0) is None


def test_camera_equals_oracle(pxl_lib, orc):
    rng = np.random.default_rng(7)
    oc = np.array([64.0, 64.0, 64.0])
    for i in range(20):
        t = rng.normal(size=3)
        t = 4 * t / np.linalg.norm(t)
        if i == 0:
            t[2] = 0.0  # camera level with the target
        cam, w2c = pxl_lib.camera_from_pose(t, oc, 1.0, 274.75, 274.75, 100, 100)
        c2w_o, w2c_o, ctr_o = orc.camera_from_pose(t, oc, 1.0, 274.75, 274.75, 100, 100)
        np.testing.assert_allclose(np.array(cam.c2w).reshape(4, 4), c2w_o, rtol=0, atol=1e-12)
        np.testing.assert_allclose(np.array(w2c).reshape(4, 4), w2c_o, rtol=0, atol=1e-12)
        np.testing.assert_allclose(np.array(cam.centre), ctr_o, rtol=0, atol=1e-9)


def test_fails_loudly_without_library(tmp_path, monkeypatch):
    """The binding raises instead of falling back when libpxl.so is absent."""
    import importlib

    import pxl

    monkeypatch.setattr(pxl, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(pxl, "_lib", None)
    with pytest.raises(ImportError):
        pxl.lib()
    monkeypatch.undo()
    importlib.reload(pxl)
