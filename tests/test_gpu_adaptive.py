"""The reference's own marching rule, t += |step - w| (volumeRender_RGB_train
and renderRays_withGradients_RGBGPU, RenderUtils.cu:1582-1718, 2051-2256),
on the GPU against the FP64 oracle.

Scene: a dense 64^3 SH-2 grid (8 seeded blobs) and its caged copy (cells
whose corners exceed 12), 8 views of 64 x 64 on a radius-4 sphere, random
8-bit targets; 1024-ray batches.

- forward (eight rays per wave, k_adapt8): pixels within 1e-5, per-ray
  sample count and the hash of the (step, cell) sequence bit-exact, dense
  and caged;
- backward on every path: the one-sample fused kernel, the two-kernel
  logged step (log holding every ray, or 24 samples so that long rays take
  the redo launch), its chunk-parallel, persistent and brick-binned
  gradient schedules;
  clipped (Huber delta 0.1) and unclipped (10) residuals; losses within
  255 * 1e-5, gradients helpers.grad_close (record 0 of the caged grid 1e-4
  relative on the logged paths, 1e-3 on the single fused kernel);
- the logged forward with its claim grid capped at 3 blocks, so that slots
  march many batch entries each and claims cross every slice.
"""
import sys

import numpy as np
import pytest

import oracle
import pxl
from helpers import (c1_planes, cameras, device_sparse_grid, from_records, grad_close_alias,
                     oracle_rays, sparse_from_dense, to_records)

pytestmark = pytest.mark.gpu

N, S, IMG, NR = 64, 9, 64, 1024


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sc(dev):
    import torch

    say("[adaptive] building the 64^3 scene")
    planes = c1_planes(N, S, n_blobs=8, seed=61)
    cams = cameras(8, IMG, 4.0, 40.0, N)
    rays_np, _ = oracle_rays(cams, IMG, (0, N - 1, 0, N - 1, 0, N - 1), seed=62)
    batch = np.random.default_rng(63).choice(len(rays_np["ox"]), NR, replace=False)
    batch = batch.astype(np.int32)
    occ, cage, io, n, c2d, compact = sparse_from_dense(planes, N, 12.0, dev)
    return dict(planes=planes, rays_np=rays_np, batch=batch, dev=dev,
                si=torch.from_numpy(batch).to(dev), occ=occ, cage=cage, io=io, c2d=c2d,
                compact=compact)


def _grids(sc, kind):
    dev = sc["dev"]
    if kind == "dense":
        g = pxl.Grid(N, N, N, S, to_records(sc["planes"], dev), step=0.5,
                     march_mode=pxl.MARCH_REF_ADAPTIVE)
        go = oracle.Grid(N, N, N, S, sc["planes"], 0.5, adaptive=True)
    else:
        g = device_sparse_grid(N, S, sc["compact"], sc["occ"], sc["io"], sc["cage"], sc["c2d"],
                               dev, march=pxl.MARCH_REF_ADAPTIVE)
        go = oracle.Grid(N, N, N, S, sc["compact"], 0.5, occupancy=sc["occ"],
                         index_offsets=sc["io"], adaptive=True)
    return g, go


@pytest.mark.parametrize("kind", ["dense", "caged"])
def test_adaptive_forward(sc, kind):
    import torch

    grid, go = _grids(sc, kind)
    assert pxl.render_kernel_name(grid, "fwd").startswith("k_adapt8")
    rays = pxl.Rays.from_arrays(sc["rays_np"], device=sc["dev"])
    tr = torch.zeros(NR * 2, dtype=torch.int32, device=sc["dev"])
    pxl.render_fwd(grid, rays, sc["si"], trace=tr)
    torch.cuda.synchronize()
    ref = oracle.render_fwd(go, oracle.Rays(sc["rays_np"]), sc["batch"], trace=True)
    b = sc["batch"]
    err = max(np.abs(getattr(rays, kg).cpu().numpy()[b] - ref[ko][b]).max()
              for kg, ko in (("acc_r", "r"), ("acc_g", "g"), ("acc_b", "b"), ("acc_w", "w")))
    t = tr.cpu().numpy().view(np.uint32).reshape(-1, 2)
    say(f"[adaptive] {kind} forward: max pixel error {err:.3g}, samples per ray "
        f"{ref['trace'][:, 0].mean():.1f} (max {ref['trace'][:, 0].max()})")
    assert err <= 1e-5
    assert np.array_equal(t[:, 0], ref["trace"][:, 0]), "sample counts differ"
    assert np.array_equal(t[:, 1], ref["trace"][:, 1]), "cell sequences differ"


@pytest.mark.parametrize("kind", ["dense", "caged"])
@pytest.mark.parametrize("path", ["fused", "logged", "logged_24", "chunked", "persistent",
                                  "binned", "binned_24"])
@pytest.mark.parametrize("huber", [0.1, 10.0])
def test_adaptive_backward(sc, kind, path, huber):
    import torch

    dev = sc["dev"]
    grid, go = _grids(sc, kind)
    optim = pxl.Optim(grid)
    rays = pxl.Rays.from_arrays(sc["rays_np"], device=dev)
    losses = torch.zeros(NR, dtype=torch.float64, device=dev)
    if path == "fused":
        pxl.train_rays(grid, optim, rays, losses, sc["si"], huber_delta=huber)
    else:
        cap = 24 if path.endswith("_24") else 512
        binned = path.startswith("binned")
        ws = pxl.train_workspace(NR, cap, dev, grid=grid if binned else None)
        ws.fill_(0xAB)  # stale contents must not leak into the result
        pxl.train_rays_sched(grid, optim, rays, losses, sc["si"], ws, cap,
                             waves_per_block=1 if path == "persistent" else 4,
                             persistent=path == "persistent", chunked=path == "chunked",
                             binned=binned, huber_delta=huber)
        torch.cuda.synchronize()
        off = 256 + (NR * 32 + 255) // 256 * 256
        cnt = ws[off:off + 4 * NR].view(torch.int32).cpu().numpy()
        n_redo = int((cnt == -1).sum())
        assert int((cnt > 0).sum()) > 0 and (n_redo > 0) == (cap == 24), n_redo
    torch.cuda.synchronize()
    ro = oracle.Rays(sc["rays_np"])
    fwd = oracle.render_fwd(go, ro, sc["batch"])
    l_ref, g_ref = oracle.render_bwd(go, ro, sc["batch"], fwd, 1.0, huber)
    dl = np.abs(losses.cpu().numpy() - l_ref).max()
    rel0 = 1e-3 if path == "fused" else 1e-4
    worst = grad_close_alias(from_records(optim.grad, grid.P), g_ref, rel0=rel0)
    say(f"[adaptive] {kind} {path} huber {huber}: loss error {dl:.3g}, gradient / tolerance "
        f"{worst:.3g}")
    assert dl <= 255 * 1e-5
    assert worst <= 1.0


@pytest.mark.parametrize("kind", ["dense", "caged"])
@pytest.mark.parametrize("path", ["logged", "logged_24", "persistent"])
def test_adaptive_backward_small_claim_grid(sc, kind, path):
    """The logged forward's slots claim batch entries: with the claim grid
    capped at 3 blocks (12 or 3 waves for 1024 entries) every slot marches
    many entries and the waves move across all 8 slices of the batch; the
    result is the oracle's as with the resident grid."""
    import torch

    dev = sc["dev"]
    grid, go = _grids(sc, kind)
    optim = pxl.Optim(grid)
    rays = pxl.Rays.from_arrays(sc["rays_np"], device=dev)
    losses = torch.zeros(NR, dtype=torch.float64, device=dev)
    cap = 24 if path == "logged_24" else 512
    ws = pxl.train_workspace(NR, cap, dev)
    ws.fill_(0xAB)
    pxl.set_option(pxl.OPT_CLAIM_GRID, 3)
    try:
        pxl.train_rays_sched(grid, optim, rays, losses, sc["si"], ws, cap,
                             waves_per_block=1 if path == "persistent" else 4,
                             persistent=path == "persistent", huber_delta=0.1)
        torch.cuda.synchronize()
    finally:
        pxl.set_option(pxl.OPT_CLAIM_GRID, 0)
    off = 256 + (NR * 32 + 255) // 256 * 256
    cnt = ws[off:off + 4 * NR].view(torch.int32).cpu().numpy()
    assert int((cnt > 0).sum()) > 0 and (int((cnt == -1).sum()) > 0) == (cap == 24)
    ro = oracle.Rays(sc["rays_np"])
    fwd = oracle.render_fwd(go, ro, sc["batch"])
    l_ref, g_ref = oracle.render_bwd(go, ro, sc["batch"], fwd, 1.0, 0.1)
    dl = np.abs(losses.cpu().numpy() - l_ref).max()
    worst = grad_close_alias(from_records(optim.grad, grid.P), g_ref, rel0=1e-4)
    say(f"[adaptive] {kind} {path}, claim grid 3: loss error {dl:.3g}, gradient / tolerance "
        f"{worst:.3g}")
    assert dl <= 255 * 1e-5
    assert worst <= 1.0
