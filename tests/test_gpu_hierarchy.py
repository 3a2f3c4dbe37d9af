"""The coarse-to-fine loop (pxl.hierarchy.train, VolumeFieldRender.cu:456-577)
over two levels: a dense 32^3 SH-2 grid trained at level 1, upsampled to a
caged 64^3 grid (to_sparse_hierarchy -1), trained at level 0; 4 views of
32 x 32 random 8-bit targets per level.

- each level runs epochs * ceil(rays / batch) batches per SH band at the
  plan's batch size (trainer time step), on a grid of the level's size;
  with SHsForTargetResolutionOnly (the default configuration) level 1 trains
  bands 0, 1, 2 in turn and level 0 the whole band-2 grid (Render.cu:267-300);
- the loop (pxl_hierarchy_plan + pxl_trainer_learn_level in the C++ host
  layer) equals the same stages issued by hand (pxl.Trainer per level with
  set_band per band, pxl.upsample_grid between): same caged record count,
  per-level batch losses within 1e-6 relative, final parameters equal up to
  the order of float atomics (at most 1e-3 of them beyond 1e-3 absolute).
"""
import math
import sys

import numpy as np
import pytest

import pxl
from pxl import hierarchy

pytestmark = pytest.mark.gpu

BASE, S, VIEWS, IMG = 64, 9, 4, 32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _start_grid(dev):
    import torch

    from pxl import scenes

    n = BASE // 2
    dens = scenes.blob_density(n, 8, 21, device=dev).reshape(-1)
    rec = torch.zeros((n ** 3, pxl.record_stride(S)), dtype=torch.float32, device=dev)
    rec[:, 0] = dens
    g = torch.Generator(device=dev)
    g.manual_seed(22)
    rec[:, 1:1 + 3 * S].normal_(0.0, 0.3, generator=g)
    return pxl.Grid(n, n, n, S, rec, step=0.5)


def _make_rays(dev):
    import torch

    from pxl import scenes

    def make(level, grid):
        N = grid.W
        f, _, c, _ = scenes.view_intrinsics(IMG, 40.0)
        rays = pxl.Rays(VIEWS * IMG * IMG, device=dev)
        g8 = torch.Generator(device=dev)
        g8.manual_seed(23)
        for v, t in enumerate(scenes.fibonacci_sphere(VIEWS, 4.0, upper_only=True)):
            cam = pxl.camera_from_pose(t, [N / 2.0] * 3, 1.0, f, f, c, c)[0]
            ims = tuple(torch.randint(0, 256, (IMG * IMG,), generator=g8, device=dev,
                                      dtype=torch.int32).to(torch.uint8) for _ in range(4))
            pxl.rays_init(cam, (0, N - 1, 0, N - 1, 0, N - 1), IMG, IMG, rays,
                          start_index=v * IMG * IMG, img=ims)
        order = torch.arange(VIEWS * IMG * IMG, dtype=torch.int32, device=dev)
        return rays, order

    return make


KW = dict(learning_rate=0.1, sh_learning_rate=0.2, variance_coefficient=1e-5,
          sh_coefficient=1e-6, fused=True, log_samples=512)


def test_two_levels_match_hand_sequence(dev):
    import torch

    levels = hierarchy.plan(1, 0, 1, 256, to_sparse_hierarchy=-1)
    assert [(lv["epochs"], lv["batch_size"]) for lv in levels] == [(4, 1024), (5, 2048)]
    make = _make_rays(dev)
    seen = []

    def on_level(lv, grid, tr):
        t, _, _ = tr.state()
        n = VIEWS * IMG * IMG
        seen.append((lv["level"], grid.W, grid.sparse, t, tr.loss()[0]))
        assert t == lv["epochs"] * math.ceil(n / lv["batch_size"])

    say("[hierarchy] two levels through pxl.hierarchy.train")
    g_loop = hierarchy.train(_start_grid(dev), make, levels, 1.0, on_level=on_level, **KW)
    torch.cuda.synchronize()
    say(f"    levels: {seen}")
    assert [(s[0], s[1], bool(s[2])) for s in seen] == [(1, 32, False), (0, 64, True)]

    say("[hierarchy] the same stages by hand")
    grid = _start_grid(dev)
    hand = []
    for lv in levels:
        rays, order = make(lv["level"], grid)
        tr = pxl.Trainer(grid, pxl.Optim(grid), rays, order, batch_size=lv["batch_size"],
                         optimize_step=lv["level"], **KW)
        # Bands 0..2 in band mode above the last level, the whole grid at it.
        bands = [(0, True), (1, True), (2, True)] if lv["level"] > 0 else [(2, False)]
        for band, band_mode in bands:
            tr.set_band(band, band_mode=band_mode)
            for _ in range(lv["epochs"] * math.ceil(order.numel() / lv["batch_size"])):
                tr.step()
        hand.append(tr.loss()[0])
        tr.close()
        if lv["upsample"]:
            grid = pxl.upsample_grid(grid, 1.0, to_sparse=lv["upsample"]["to_sparse"])
    torch.cuda.synchronize()
    say(f"    batch MSE per level: loop {[s[4] for s in seen]}, by hand {hand}")
    assert g_loop.n_vox == grid.n_vox and g_loop.W == grid.W == BASE
    for a, b in zip([s[4] for s in seen], hand):
        assert abs(a - b) <= 1e-6 * abs(b)
    P = 1 + 3 * S
    diff = (g_loop.params[:, :P] - grid.params[:, :P]).abs()
    frac = float((diff > 1e-3).float().mean())
    say(f"    final records: {grid.n_vox}, max |diff| {float(diff.max()):.3g}, "
        f"fraction beyond 1e-3 {frac:.3g}")
    assert frac <= 1e-3
