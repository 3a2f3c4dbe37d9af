"""Parity at the bench's full size: the 256^3 SH-2 caged grid at 10 % cell
occupancy (64 seeded blobs, 1.66 M records) and a 65 536-ray batch from
four 800 x 800 views of the upper hemisphere, random 8-bit targets.

- occupancy bits, in-cage flags and caged index bit-exact against the oracle
  at 256^3;
- 128 rays of the batch against the FP64 oracle on the full grid (colours
  1e-5, losses 255 * 1e-5, gradients helpers.grad_close), fused and logged;
- the whole batch against the FP64 oracle (threaded on the host): the
  fused kernel, the logged step, the persistent one-wave schedule, the
  chunk-parallel gradient and the brick-binned gradient give the oracle's
  per-ray colours and losses, and
  gradients within helpers.grad_close or, for entries whose terms cancel
  (one entry here is 9.9e-7 from four rays' terms of 1e-3 in magnitude),
  within 1e-6 of the entry's magnitude sum (helpers.grad_close_cond: FP32
  arithmetic resolves such an entry to a few rounding units of that sum);
  record 0, where every out-of-cage
  corner lands (~10^5 largely cancelling terms), within 1e-4 relative on the
  logged paths, which sum it per ray in FP64 and fold it in a fixed order;
  the single fused kernel (no workspace) adds it with one float atomic per
  ray and is held to 1e-3 there;
- the whole batch's gradient is the sum of its two halves' (no cross-ray
  interference, record 0 included, 1e-4);
- 128 rays of the batch under the reference's adaptive rule (fused, logged
  and brick-binned) against the oracle;
- TV over the full caged grid and one Adam step over all records against
  the oracle.
"""
import sys

import numpy as np
import pytest

import oracle
import pxl
from helpers import from_records, grad_close, grad_close_alias, grad_close_cond, pack_bits

pytestmark = pytest.mark.gpu

N, S, P = 256, 9, 28
VIEWS, IMG, BATCH = 4, 800, 65536


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c2(dev):
    return build_c2(dev)


def build_c2(dev):
    """The 256^3 caged scene, its rays and the 65 536-ray batch."""
    import torch

    from pxl import scenes

    say("[c2] building the 256^3 caged scene")
    dens = scenes.blob_density(N, 64, 11, device=dev).reshape(-1).contiguous()
    ncell = (N - 1) ** 3
    cellmax = torch.nn.functional.max_pool3d(dens.view(1, 1, N, N, N), 2, stride=1).reshape(-1)
    tau = float(torch.kthvalue(cellmax.float().cpu(), int(round(0.9 * ncell))).values)
    del cellmax
    occ = torch.zeros((ncell + 31) // 32, dtype=torch.int32, device=dev)
    cage = torch.zeros(N ** 3, dtype=torch.uint8, device=dev)
    nocc = torch.zeros(1, dtype=torch.int64, device=dev)
    pxl.build_occupancy(dens, N, N, N, tau, occ, cage, nocc)
    io = torch.zeros(N ** 3, dtype=torch.int32, device=dev)
    c2d = torch.zeros(N ** 3, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    scratch = torch.zeros(pxl.sparse_index_scratch_bytes(N ** 3) // 4 + 1, dtype=torch.int32,
                          device=dev)
    pxl.build_sparse_index(cage, N, N, N, io, c2d, cnt, scratch)
    torch.cuda.synchronize()
    n_vox = int(cnt.item())
    c2d = c2d[:n_vox].clone()
    gen = torch.Generator(device=dev)
    gen.manual_seed(12)
    stride = pxl.record_stride(S)
    dense_rec = torch.zeros((N ** 3, stride), dtype=torch.float32, device=dev)
    dense_rec[:, 0] = dens
    dense_rec[:, 1:P].normal_(0.0, 0.3, generator=gen)
    rec = torch.zeros((n_vox, stride), dtype=torch.float32, device=dev)
    pxl.grid_compact(dense_rec, c2d, n_vox, rec)
    del dense_rec
    bricks = pxl.build_brick_occupancy(occ, N, N, N)
    grid_kw = dict(step=0.5, occupancy=occ, index_offsets=io, in_cage=cage,
                   compact_to_dense=c2d, brick_occupancy=bricks)
    say(f"[c2] tau {tau:.3f}, {int(nocc.item()) / ncell:.4f} of cells occupied, {n_vox} records")

    f, _, c, _ = scenes.view_intrinsics(IMG, 40.0)
    cams = [pxl.camera_from_pose(t, [N / 2.0] * 3, 1.0, f, f, c, c)[0]
            for t in scenes.fibonacci_sphere(VIEWS, 4.0, upper_only=True)]
    npix = IMG * IMG
    rays = pxl.Rays(VIEWS * npix, device=dev)
    g8 = torch.Generator(device=dev)
    g8.manual_seed(13)
    for v, cam in enumerate(cams):
        ims = tuple(torch.randint(0, 256, (npix,), generator=g8, device=dev,
                                  dtype=torch.int32).to(torch.uint8) for _ in range(4))
        pxl.rays_init(cam, (0, N - 1, 0, N - 1, 0, N - 1), IMG, IMG, rays, start_index=v * npix,
                      img=ims)
    batch = np.sort(np.random.default_rng(14).choice(VIEWS * npix, BATCH, replace=False))
    torch.cuda.synchronize()
    say("[c2] rays built")
    return dict(dens=dens, tau=tau, occ=occ, cage=cage, io=io, n_vox=n_vox, rec=rec,
                grid_kw=grid_kw, rays=rays, batch=batch.astype(np.int32), dev=dev)


@pytest.fixture(scope="module")
def c2_oracle(c2):
    return build_c2_oracle(c2)


def build_c2_oracle(c2):
    """The oracle's occupancy and index of the same density, its grid over
    the same records, and the rays as host arrays."""
    occ_o, cage_o, n_o = oracle.build_occupancy(c2["dens"].cpu().numpy(), N, N, N, c2["tau"])
    io_o, cnt_o = oracle.build_index(cage_o, N, N, N)
    planes = np.ascontiguousarray(c2["rec"][:, :P].T.cpu().numpy())
    grid_o = oracle.Grid(N, N, N, S, planes, 0.5, occupancy=occ_o, index_offsets=io_o)
    r = c2["rays"]
    rays_np = {k: getattr(r, k).cpu().numpy() for k in
               ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t", "gt_r", "gt_g", "gt_b",
                "gt_mask")}
    return dict(occ=occ_o, cage=cage_o, n_occ=n_o, io=io_o, cnt=cnt_o, grid=grid_o,
                rays=oracle.Rays(rays_np))


def _grid(c2):
    return pxl.Grid(N, N, N, S, c2["rec"], **c2["grid_kw"])


def _train(c2, si, path):
    import torch

    grid = _grid(c2)
    # "logged_g32": the bench's layout, gradient rows of 32 floats beside
    # moments of 28.
    optim = (pxl.Optim(grid, stride=28, grad_stride=32) if path == "logged_g32"
             else pxl.Optim(grid))
    n = si.numel()
    losses = torch.zeros(n, dtype=torch.float64, device=c2["dev"])
    if path == "fused":
        pxl.train_rays(grid, optim, c2["rays"], losses, si)
    else:
        ws = pxl.train_workspace(n, 512, c2["dev"], grid=grid if path == "binned" else None)
        if path == "persistent":
            pxl.train_rays_sched(grid, optim, c2["rays"], losses, si, ws, 512, waves_per_block=1,
                                 persistent=True)
        elif path == "binned":
            pxl.train_rays_sched(grid, optim, c2["rays"], losses, si, ws, 512, waves_per_block=1,
                                 binned=True, grad_waves_per_block=4)
        elif path == "chunked":
            pxl.train_rays_sched(grid, optim, c2["rays"], losses, si, ws, 512, chunked=True)
        else:
            pxl.train_rays_ws(grid, optim, c2["rays"], losses, si, ws, 512)
    torch.cuda.synchronize()
    r = c2["rays"]
    cols = np.stack([getattr(r, k).cpu().numpy()[si.cpu().numpy()]
                     for k in ("acc_r", "acc_g", "acc_b", "acc_w")])
    return optim, losses.cpu().numpy(), cols


def test_c2_occupancy_and_index(c2, c2_oracle):
    o = c2_oracle
    say(f"[c2] occupancy and index vs oracle: {o['n_occ']} occupied cells, {o['cnt']} records")
    assert o["cnt"] == c2["n_vox"]
    assert np.array_equal(c2["occ"].cpu().numpy().view(np.uint32), pack_bits(o["occ"]))
    assert np.array_equal(c2["cage"].cpu().numpy(), o["cage"])
    assert np.array_equal(c2["io"].cpu().numpy().view(np.uint32), o["io"])


@pytest.mark.parametrize("path", ["fused", "logged"])
def test_c2_sampled_rays(c2, c2_oracle, path):
    import torch

    sub = c2["batch"][np.random.default_rng(15).choice(BATCH, 128, replace=False)]
    say(f"[c2] 128 rays of the batch vs oracle, {path}")
    optim, losses, cols = _train(c2, torch.from_numpy(sub).to(c2["dev"]), path)
    fwd = oracle.render_fwd(c2_oracle["grid"], c2_oracle["rays"], sub)
    l_ref, g_ref = oracle.render_bwd(c2_oracle["grid"], c2_oracle["rays"], sub, fwd,
                                     mask_coef=1.0, huber_delta=0.1)
    err = max(np.abs(cols[i] - fwd[k][sub]).max() for i, k in enumerate("rgbw"))
    say(f"    max colour / opacity error {err:.3g}, max loss error {np.abs(losses - l_ref).max():.3g}")
    assert err <= 1e-5
    assert np.abs(losses - l_ref).max() <= 255 * 1e-5
    worst = grad_close(from_records(optim.grad, P), g_ref)
    say(f"    worst gradient error / tolerance = {worst:.3g}")
    assert worst <= 1.0


@pytest.mark.parametrize("path", ["fused", "logged", "binned"])
def test_c2_sampled_rays_adaptive(c2, c2_oracle, path):
    """128 rays of the batch under the reference's adaptive marching rule
    (t += |step - w|) against the oracle on the full 256^3 grid."""
    import torch

    sub = c2["batch"][np.random.default_rng(16).choice(BATCH, 128, replace=False)]
    say(f"[c2] 128 rays, adaptive rule, {path}")
    grid = pxl.Grid(N, N, N, S, c2["rec"], march_mode=pxl.MARCH_REF_ADAPTIVE, **c2["grid_kw"])
    optim = pxl.Optim(grid)
    losses = torch.zeros(128, dtype=torch.float64, device=c2["dev"])
    si = torch.from_numpy(sub).to(c2["dev"])
    if path == "fused":
        pxl.train_rays(grid, optim, c2["rays"], losses, si)
    elif path == "binned":
        ws = pxl.train_workspace(128, 512, c2["dev"], grid=grid)
        pxl.train_rays_sched(grid, optim, c2["rays"], losses, si, ws, 512, binned=True)
    else:
        ws = pxl.train_workspace(128, 512, c2["dev"])
        pxl.train_rays_ws(grid, optim, c2["rays"], losses, si, ws, 512)
    torch.cuda.synchronize()
    g_o = c2_oracle["grid"]
    grid_o = oracle.Grid(N, N, N, S, g_o.planes, 0.5, occupancy=g_o.occupancy,
                         index_offsets=g_o.index_offsets, adaptive=True)
    fwd = oracle.render_fwd(grid_o, c2_oracle["rays"], sub)
    l_ref, g_ref = oracle.render_bwd(grid_o, c2_oracle["rays"], sub, fwd, mask_coef=1.0,
                                     huber_delta=0.1)
    r = c2["rays"]
    err = max(np.abs(getattr(r, k).cpu().numpy()[sub] - fwd[o][sub]).max()
              for k, o in (("acc_r", "r"), ("acc_g", "g"), ("acc_b", "b"), ("acc_w", "w")))
    dl = np.abs(losses.cpu().numpy() - l_ref).max()
    worst = grad_close(from_records(optim.grad, P), g_ref)
    say(f"    max colour / opacity error {err:.3g}, loss error {dl:.3g}, gradient / tolerance "
        f"{worst:.3g}")
    assert err <= 1e-5 and dl <= 255 * 1e-5
    assert worst <= 1.0


@pytest.fixture(scope="module")
def c2_full_ref(c2, c2_oracle):
    """The oracle's forward and backward of the whole 65 536-ray batch."""
    import os
    import time

    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    oracle.set_threads(threads)
    try:
        t0 = time.time()
        batch = c2["batch"]
        fwd = oracle.render_fwd(c2_oracle["grid"], c2_oracle["rays"], batch)
        l_ref, g_ref, g_abs = oracle.render_bwd_cond(c2_oracle["grid"], c2_oracle["rays"], batch,
                                                     fwd, mask_coef=1.0, huber_delta=0.1)
        say(f"[c2] oracle of the whole batch on {threads} threads: {time.time() - t0:.1f} s")
    finally:
        oracle.set_threads(1)
    cols = np.stack([fwd[k][batch] for k in "rgbw"])
    return cols, l_ref, g_ref, g_abs


@pytest.mark.parametrize("path", ["fused", "logged", "logged_g32", "persistent", "chunked",
                                  "binned"])
def test_c2_full_batch_vs_oracle(c2, c2_full_ref, path):
    import torch

    si = torch.from_numpy(c2["batch"]).to(c2["dev"])
    cols_ref, l_ref, g_ref, g_abs = c2_full_ref
    say(f"[c2] full batch vs oracle: {path}")
    o, l, c = _train(c2, si, path)
    dc, dl = np.abs(c - cols_ref).max(), np.abs(l - l_ref).max()
    g = from_records(o.grad, P)
    del o
    rel0 = 1e-3 if path == "fused" else 1e-4
    worst, n_cond = grad_close_cond(g, g_ref, g_abs, rel0=rel0)
    e0 = max(abs(g[p][0] - g_ref[p][0]) / max(1e-30, abs(g_ref[p][0])) for p in range(P))
    say(f"    colour error {dc:.3g}, loss error {dl:.3g}, gradient / tolerance {worst:.3g} "
        f"({n_cond} of {int((g_ref != 0).sum())} nonzero entries held by the magnitude term), "
        f"record 0 relative error {e0:.3g} (bar {rel0:g})")
    assert dc <= 1e-5 and dl <= 255 * 1e-5
    assert worst <= 1.0


def test_c2_gradient_additive(c2):
    import torch

    say("[c2] full batch gradient = sum of its halves'")
    si = torch.from_numpy(c2["batch"]).to(c2["dev"])
    o_full, _, _ = _train(c2, si, "logged")
    g_full = from_records(o_full.grad, P)
    del o_full
    h = BATCH // 2
    o_a, _, _ = _train(c2, si[:h].clone(), "logged")
    g_sum = from_records(o_a.grad, P)
    del o_a
    o_b, _, _ = _train(c2, si[h:].clone(), "logged")
    g_sum += from_records(o_b.grad, P)
    worst = grad_close_alias(g_full, g_sum, rel0=1e-4)
    say(f"    worst difference / tolerance = {worst:.3g}")
    assert worst <= 1.0


def test_c2_tv_and_adam(c2, c2_oracle):
    import torch

    say("[c2] TV over the full caged grid and one Adam step, vs oracle")
    grid = _grid(c2)
    optim = pxl.Optim(grid)
    pxl.add_variance_loss(grid, optim, 1e-5, 1e-6, pxl.TV_DISPARITY, 6, optimize_step=0)
    torch.cuda.synchronize()
    g_tv = from_records(optim.grad, P)
    ref = oracle.tv_disparity(c2_oracle["grid"], 1e-5, 1e-6, optimize_step=0)
    worst = grad_close(g_tv, ref)
    say(f"    worst TV error / tolerance = {worst:.3g}")
    assert worst <= 1.0
    # Adam on every record from the TV gradient (t = 1).
    p_o = c2_oracle["grid"].planes.copy()
    gr = np.ascontiguousarray(g_tv.astype(np.float32))
    m = np.zeros_like(p_o)
    v = np.zeros_like(p_o)
    oracle.adam(p_o, gr, m, v, 0.3, 1.0, 1)
    rec0 = c2["rec"].clone()
    pxl.adam_step(grid, optim, 0.3, 1.0, 1)
    torch.cuda.synchronize()
    got = from_records(grid.params, P)
    c2["rec"].copy_(rec0)  # the other tests use the unstepped grid
    err = np.abs(got - p_o) / (np.abs(p_o) + 1e-30)
    say(f"    Adam: worst relative parameter error {err[np.abs(p_o) > 1e-6].max():.3g}")
    np.testing.assert_allclose(got, p_o, rtol=1e-6, atol=1e-7)
    assert float(optim.grad.abs().max()) == 0.0
