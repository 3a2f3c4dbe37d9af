"""Parity of the coarse-to-fine upsampling (pxl_upsample_density,
pxl_upsample; upsampleScene / upsample_trilinear / initSelfByUpsample,
SparseVolumeRadianceField/SceneUtils.cu:250-657) with the oracle.

The GPU repeats the reference's float operations in its order, so records,
occupancy bits, in-cage flags and the caged index must match bit for bit."""
import sys

import numpy as np
import pytest

import oracle
import pxl
from helpers import c1_planes, device_sparse_grid, pack_bits, sparse_from_dense, to_records

pytestmark = pytest.mark.gpu

N, S = 40, 9
P = 1 + 3 * S


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def coarse():
    return c1_planes(N, S, n_blobs=6, seed=21)


def _bits(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_upsample_dense(dev, coarse, storage):
    import torch

    say(f"[upsample] dense {N}^3 -> {2 * N}^3, {storage} source")
    planes = coarse
    rec = to_records(planes, dev)
    if storage == "bf16":
        rec = pxl.records_to_16bit(rec, pxl.STORE_BF16)
        src = pxl.Grid(N, N, N, S, rec, storage=pxl.STORE_BF16)
        # the oracle reads the same bfloat16 values, widened
        wide = (rec.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
        planes = np.ascontiguousarray(wide[:, :P].T)
    else:
        src = pxl.Grid(N, N, N, S, rec)
    dst = pxl.upsample_grid(src, 6.0, to_sparse=False)
    torch.cuda.synchronize()
    ref = oracle.upsample(oracle.Grid(N, N, N, S, planes, 0.5), 2 * N, 2 * N, 2 * N, S, 6.0)
    got = dst.params.cpu().numpy()
    assert np.array_equal(got[:, :P].T.view(np.uint32), ref.view(np.uint32))
    assert np.all(got[:, P:] == 0.0)
    dens = pxl.upsample_density(src, 2 * N, 2 * N, 2 * N, 6.0)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(dens), ref[0].view(np.uint32))


def test_upsample_low_band(dev, coarse):
    import torch

    say("[upsample] low band into a 1-coefficient grid")
    src = pxl.Grid(N, N, N, S, to_records(coarse, dev))
    dst = pxl.upsample_grid(src, 6.0, to_sparse=False, low_band=True)
    torch.cuda.synchronize()
    assert dst.S == 1
    ref = oracle.upsample(oracle.Grid(N, N, N, S, coarse, 0.5), 2 * N, 2 * N, 2 * N, 1, 6.0,
                          low_band=True)
    got = dst.params.cpu().numpy()
    assert np.array_equal(got[:, :4].T.view(np.uint32), ref.view(np.uint32))


def test_upsample_caged(dev, coarse):
    """Caged source to caged destination: density, prune, index and records
    of the finer grid against the oracle's pipeline."""
    import torch

    tau = 8.0
    say(f"[upsample] caged {N}^3 (tau {tau}) -> caged {2 * N}^3")
    occ, cage, io, n, c2d, compact = sparse_from_dense(coarse, N, tau, dev)
    src = device_sparse_grid(N, S, compact, occ, io, cage, c2d, dev)
    dst = pxl.upsample_grid(src, tau, to_sparse=True)
    torch.cuda.synchronize()
    N2 = 2 * N
    grid_o = oracle.Grid(N, N, N, S, compact, 0.5, occupancy=occ, index_offsets=io)
    dens = oracle.upsample_density(grid_o, N2, N2, N2, tau)
    occ2, cage2, nocc2 = oracle.build_occupancy(dens, N2, N2, N2, tau)
    io2, n2 = oracle.build_index(cage2, N2, N2, N2)
    assert dst.n_vox == n2 and n2 > 1000
    assert np.array_equal(_bits(dst.occupancy)[:len(pack_bits(occ2))], pack_bits(occ2))
    assert np.array_equal(dst.in_cage.cpu().numpy(), cage2)
    assert np.array_equal(_bits(dst.index_offsets), io2)
    assert np.array_equal(_bits(dst.compact_to_dense)[:n2],
                          np.flatnonzero(cage2).astype(np.uint32))
    ref = oracle.upsample(grid_o, N2, N2, N2, S, tau, in_cage=cage2, index_offsets=io2, n=n2)
    got = dst.params.cpu().numpy()
    assert np.array_equal(got[:, :P].T.view(np.uint32), ref.view(np.uint32))
    live = (ref[0] != 0).mean()
    say(f"    {n2} caged records, {nocc2} occupied cells, nonzero density {live:.3f}")
    assert 0.0 < live < 1.0
