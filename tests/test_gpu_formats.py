"""Scene directories in the reference's text format (objectScene::saveSelf /
initFromFile, SparseVolumeRadianceField/SceneUtils.cu:1246-1299, 703-790)
through pxl_grid_save_text / pxl_grid_load_text.

Bars: integers (occupancy, in-cage flags, index, record order) bit-exact;
each float equal to the C library's strtof of its "%.6lf" text, as the
reference's fscanf("%f") reads it."""
import ctypes
import os

import numpy as np
import pytest

import pxl
from helpers import c1_planes, device_sparse_grid, pack_bits, sparse_from_dense, to_records

pytestmark = pytest.mark.gpu

N, S, BAND = 20, 9, 2
P = 1 + 3 * S

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def _text_round(planes):
    """float32 value the reference reads back from its "%.6lf " text."""
    flat = np.asarray(planes, np.float32).reshape(-1)
    out = np.array([_libc.strtof(b"%.6f" % float(v), None) for v in flat], np.float32)
    return out.reshape(np.shape(planes))


@pytest.fixture(scope="module")
def planes():
    return c1_planes(N, S, n_blobs=4, seed=31)


def _bits(t):
    return t.cpu().numpy().view(np.uint32)


def test_dense_round_trip(dev, planes, tmp_path):
    import torch

    g = pxl.Grid(N, N, N, S, to_records(planes, dev))
    pxl.grid_save_text(g, tmp_path, BAND)
    names = sorted(os.listdir(tmp_path))
    assert "opacity_band2.txt" in names and "occupancy.txt" in names
    assert sum(n.startswith("red_") for n in names) == 9
    back = pxl.grid_load_text(tmp_path, N, N, N, BAND, sparse=False, device=dev)
    torch.cuda.synchronize()
    got = back.params.cpu().numpy()
    assert np.array_equal(got[:, :P].T.view(np.uint32), _text_round(planes).view(np.uint32))
    assert np.all(got[:, P:] == 0.0)


def test_caged_round_trip_and_reference_directory(dev, planes, tmp_path):
    """A caged grid saved and loaded back; then the same directory as the
    reference leaves it (the count written over the in-cage file, no count
    file): the loader rebuilds the in-cage flags from occupancy.txt."""
    import torch

    occ, cage, io, n, c2d, compact = sparse_from_dense(planes, N, 10.0, dev)
    g = device_sparse_grid(N, S, compact, occ, io, cage, c2d, dev)
    pxl.grid_save_text(g, tmp_path, BAND)
    expect = _text_round(compact).view(np.uint32)
    for variant in ("own", "reference"):
        if variant == "reference":
            (tmp_path / "sparse_inCageInfo_band2.txt").write_text("%d" % n)
            os.remove(tmp_path / "InCageNumsInfo.txt")
        back = pxl.grid_load_text(tmp_path, N, N, N, BAND, sparse=True, device=dev)
        torch.cuda.synchronize()
        assert back.n_vox == n, variant
        assert np.array_equal(_bits(back.occupancy)[:len(pack_bits(occ))], pack_bits(occ))
        assert np.array_equal(back.in_cage.cpu().numpy(), cage), variant
        assert np.array_equal(_bits(back.index_offsets), io), variant
        assert np.array_equal(_bits(back.compact_to_dense)[:n], c2d), variant
        got = back.params.cpu().numpy()
        assert np.array_equal(got[:, :P].T.view(np.uint32), expect), variant


def test_band_subsets(dev, planes, tmp_path):
    """hierarchy_mode saves only the last band's coefficients; density_only
    no SH files; a lower band loads into a wider grid with zeros above."""
    import torch

    g = pxl.Grid(N, N, N, S, to_records(planes, dev))
    (tmp_path / "h").mkdir()
    pxl.grid_save_text(g, tmp_path / "h", BAND, hierarchy_mode=True)
    reds = sorted(n for n in os.listdir(tmp_path / "h") if n.startswith("red_"))
    assert reds == ["red_%d.txt" % i for i in range(4, 9)]
    (tmp_path / "d").mkdir()
    pxl.grid_save_text(g, tmp_path / "d", BAND, density_only=True)
    assert not any(n.startswith(("red_", "green_", "blue_")) for n in os.listdir(tmp_path / "d"))
    (tmp_path / "b1").mkdir()
    pxl.grid_save_text(g, tmp_path / "b1", 1)
    back = pxl.grid_load_text(tmp_path / "b1", N, N, N, 1, sparse=False, sh_dim=S, device=dev)
    torch.cuda.synchronize()
    got = back.params.cpu().numpy()
    keep = [0] + [1 + ch * S + i for ch in range(3) for i in range(4)]
    assert np.array_equal(got[:, keep].T.view(np.uint32),
                          _text_round(planes[keep]).view(np.uint32))
    zero = [1 + ch * S + i for ch in range(3) for i in range(4, 9)]
    assert np.all(got[:, zero] == 0.0)
