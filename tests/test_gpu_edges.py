"""Edge cases of the train step against the oracle, on every launch path
(fused kernel, logged two-kernel step, logged step with rays redone past a
short log, persistent single-wave blocks, chunk-parallel gradient,
brick-binned gradient with a full and a short log):

- ragged batch sizes: 1, 63, 65 and 257 rays (partial waves and blocks);
- rays that miss the grid (near > far: outputs zeroed and no sample,
  RenderUtils.cu:1919-1946) and rays of zero length (near == far: no
  sample, background only), mixed into a batch;
- a batch in which every ray misses, and an empty batch.

Bars: losses within 255 * 1e-5 and colours within 1e-5 absolute, per ray
(a ray dropped or taken twice fails here); gradients within
helpers.grad_close with a floor of 1e-2 of the plane maximum.  With a few
dozen rays a plane's maximum comes from one or two rays, and entries 1e-4 of
it and below can hinge on a pixel whose colour nearly equals its ground
truth: the loss gradient C - gt then carries the FP32 forward's rounding
(~1e-7) at a relative 1e-3.  The 4096-ray parity tests keep the 1e-3 floor.
"""
import sys

import numpy as np
import pytest

import oracle
import pxl
from helpers import (c1_planes, cameras, device_sparse_grid, from_records, grad_close,
                     oracle_rays, sparse_from_dense, to_records)

pytestmark = pytest.mark.gpu

N, S, IMG = 48, 9, 48
PATHS = ("fused", "logged", "redo", "persistent", "chunked", "binned", "binned_redo")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scene(dev):
    say("[edges] building 48^3 scene")
    planes = c1_planes(N, S, n_blobs=6, seed=11)
    cams = cameras(6, IMG, 4.0, 40.0, N)
    rays_np, _ = oracle_rays(cams, IMG, (0, N - 1, 0, N - 1, 0, N - 1), seed=2)
    occ, cage, io, n, c2d, compact = sparse_from_dense(planes, N, 12.0, dev)
    return dict(planes=planes, rays_np=rays_np, dev=dev,
                sparse=dict(occ=occ, cage=cage, io=io, c2d=c2d, compact=compact))


def _grids(scene, kind):
    dev = scene["dev"]
    if kind == "dense":
        grid = pxl.Grid(N, N, N, S, to_records(scene["planes"], dev), step=0.5)
        return grid, oracle.Grid(N, N, N, S, scene["planes"], 0.5)
    s = scene["sparse"]
    grid = device_sparse_grid(N, S, s["compact"], s["occ"], s["io"], s["cage"], s["c2d"], dev)
    return grid, oracle.Grid(N, N, N, S, s["compact"], 0.5, occupancy=s["occ"],
                             index_offsets=s["io"])


def _run(grid, rays, si, path, dev):
    import torch

    nr = si.numel()
    optim = pxl.Optim(grid)
    losses = torch.zeros(max(nr, 1), dtype=torch.float64, device=dev)
    if path == "fused":
        pxl.train_rays(grid, optim, rays, losses, si)
    else:
        cap = 8 if path in ("redo", "binned_redo") else 512
        binned = path.startswith("binned")
        ws = pxl.train_workspace(max(nr, 1), cap, dev, grid=grid if binned else None)
        if binned:
            pxl.train_rays_sched(grid, optim, rays, losses, si, ws, cap, binned=True)
        elif path == "persistent":
            pxl.train_rays_sched(grid, optim, rays, losses, si, ws, cap, waves_per_block=1,
                                 persistent=True)
        elif path == "chunked":
            pxl.train_rays_sched(grid, optim, rays, losses, si, ws, cap, chunked=True)
        else:
            pxl.train_rays_ws(grid, optim, rays, losses, si, ws, cap)
    torch.cuda.synchronize()
    return optim, losses[:nr].cpu().numpy()


def _compare(grid, grid_o, rays_np, batch, path, dev):
    import torch

    rays = pxl.Rays.from_arrays(rays_np, device=dev)
    si = torch.from_numpy(batch).to(dev)
    optim, losses = _run(grid, rays, si, path, dev)
    rays_o = oracle.Rays(rays_np)
    fwd = oracle.render_fwd(grid_o, rays_o, batch)
    l_ref, g_ref = oracle.render_bwd(grid_o, rays_o, batch, fwd, mask_coef=1.0,
                                     huber_delta=0.1)
    assert np.abs(losses - l_ref).max(initial=0.0) <= 255 * 1e-5
    for k_g, k_o in (("acc_r", "r"), ("acc_g", "g"), ("acc_b", "b")):
        err = np.abs(getattr(rays, k_g).cpu().numpy()[batch] - fwd[k_o][batch]).max(initial=0.0)
        assert err <= 1e-5, (k_g, err)
    worst = grad_close(from_records(optim.grad, grid.P), g_ref, floor=1e-2)
    say(f"    worst gradient error / tolerance = {worst:.3g}")
    assert worst <= 1.0
    return fwd


@pytest.mark.parametrize("kind", ["dense", "sparse"])
@pytest.mark.parametrize("path", PATHS)
def test_ragged_batches(scene, kind, path):
    grid, grid_o = _grids(scene, kind)
    rng = np.random.default_rng(5)
    nrays = len(scene["rays_np"]["ox"])
    for nr in (1, 63, 65, 257):
        say(f"[edges] {kind} {path}: {nr} rays")
        batch = rng.choice(nrays, nr, replace=False).astype(np.int32)
        _compare(grid, grid_o, scene["rays_np"], batch, path, scene["dev"])


@pytest.mark.parametrize("kind", ["dense", "sparse"])
@pytest.mark.parametrize("path", PATHS)
def test_misses_and_zero_length(scene, kind, path):
    say(f"[edges] {kind} {path}: misses and zero-length rays in a 200-ray batch")
    grid, grid_o = _grids(scene, kind)
    rays_np = {k: v.copy() for k, v in scene["rays_np"].items()}
    nrays = len(rays_np["ox"])
    batch = np.random.default_rng(6).choice(nrays, 200, replace=False).astype(np.int32)
    hit = batch[rays_np["near_t"][batch] < rays_np["far_t"][batch]]
    miss, zero = hit[:40], hit[40:60]
    rays_np["near_t"][miss] = rays_np["far_t"][miss] + 1.0
    rays_np["near_t"][zero] = rays_np["far_t"][zero]
    fwd = _compare(grid, grid_o, rays_np, batch, path, scene["dev"])
    assert np.all(fwd["r"][miss] == 0.0)
    assert np.all(fwd["r"][zero] == 1.0)  # white background, no sample


@pytest.mark.parametrize("path", PATHS)
def test_all_miss_and_empty_batches(scene, path):
    import torch

    say(f"[edges] {path}: every ray misses; empty batch")
    grid, grid_o = _grids(scene, "sparse")
    rays_np = {k: v.copy() for k, v in scene["rays_np"].items()}
    batch = np.arange(0, 130, dtype=np.int32)
    rays_np["near_t"][batch] = rays_np["far_t"][batch] + 1.0
    _compare(grid, grid_o, rays_np, batch, path, scene["dev"])
    dev = scene["dev"]
    rays = pxl.Rays.from_arrays(scene["rays_np"], device=dev)
    empty = torch.zeros(0, dtype=torch.int32, device=dev)
    optim, losses = _run(grid, rays, empty, path, dev)
    assert losses.size == 0
    assert float(optim.grad.abs().max()) == 0.0
    pxl.render_fwd(grid, rays, empty)
    pxl.batch_order(grid, rays, empty, torch.zeros(0, dtype=torch.int32, device=dev))
    torch.cuda.synchronize()


def test_trainer_create_reports_why(scene):
    """A trainer that cannot be built says why (the release of the partly
    built trainer keeps the error text): schedules without the logged step,
    the chunk-parallel gradient in persistent form."""
    import torch

    grid, _ = _grids(scene, "sparse")
    dev = scene["dev"]
    rays = pxl.Rays.from_arrays(scene["rays_np"], device=dev)
    order = torch.arange(256, dtype=torch.int32, device=dev)
    with pytest.raises(pxl.PxlError, match="need the logged step"):
        pxl.Trainer(grid, pxl.Optim(grid), rays, order, batch_size=128, log_samples=0,
                    autotune=True)
    with pytest.raises(pxl.PxlError, match="no persistent form"):
        pxl.Trainer(grid, pxl.Optim(grid), rays, order, batch_size=128, log_samples=64,
                    persistent=True, chunked=True)
    torch.cuda.synchronize()
