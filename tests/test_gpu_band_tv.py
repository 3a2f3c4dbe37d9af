"""GPU parity of the SH band strategy, the variance-loss types, the NaN path
of the backward and the trainer's side-stream overlap, against the oracle.

Scene: a dense 64^3 SH-2 grid (8 seeded blobs) and its caged copy (cells
whose corners exceed 12), 8 views of 64 x 64 on a radius-4 sphere, random
8-bit targets; 1024-ray batches.

- bands (RenderUtils.cu:859-1000, 1117-1132): colour from the coefficients
  below sh_active, gradients of [sh_first, sh_active), on the fused, logged
  and separate paths and both marchers;
- variance losses (RenderUtils.cu:2442-2851): type 0 (6 neighbours) and type
  2 on dense and caged grids, band-limited, against the oracle's serial sum;
- Adam and the optimiser's band initialisation (Optimizer.cu:11-64,
  359-415);
- the C++ level loop (pxl_trainer_learn_level) equals the same bands,
  epochs and batches issued stage by stage;
- a NaN parameter: rays through it get loss 255 and add no gradient, every
  other ray equals the oracle (RenderUtils.cu:2158-2162); the batch loss
  stays finite, as in the reference;
- the trainer with the variance loss beside the render equals the stage
  calls on grids the logged step does not serve (the adaptive rule, SH
  degree 3) and on a caged grid with an interval of 3, overlap on and off.
Gradient bar: helpers.grad_close (1e-4 relative, floor 1e-3 of the plane
maximum), pixels 1e-5.
"""
import math
import sys

import numpy as np
import pytest

import oracle
import pxl
from helpers import (c1_planes, cameras, device_sparse_grid, from_records, grad_close,
                     grad_close_alias, oracle_rays, sparse_from_dense, to_records)

pytestmark = pytest.mark.gpu

N, S, IMG, NR = 64, 9, 64, 1024


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sc(dev):
    import torch

    say("[band] building the 64^3 scene")
    planes = c1_planes(N, S, n_blobs=8, seed=31)
    cams = cameras(8, IMG, 4.0, 40.0, N)
    bbox = (0, N - 1, 0, N - 1, 0, N - 1)
    rays_np, _ = oracle_rays(cams, IMG, bbox, seed=32)
    batch = np.random.default_rng(33).choice(len(rays_np["ox"]), NR, replace=False)
    batch = batch.astype(np.int32)
    occ, cage, io, n, c2d, compact = sparse_from_dense(planes, N, 12.0, dev)
    return dict(planes=planes, rays_np=rays_np, batch=batch, dev=dev,
                si=torch.from_numpy(batch).to(dev), occ=occ, cage=cage, io=io, n=n, c2d=c2d,
                compact=compact)


def _grids(sc, kind, sh_active=0, march=pxl.MARCH_FIXED):
    dev = sc["dev"]
    adapt = march == pxl.MARCH_REF_ADAPTIVE
    if kind == "dense":
        g = pxl.Grid(N, N, N, S, to_records(sc["planes"], dev), step=0.5, sh_active=sh_active,
                     march_mode=march)
        go = oracle.Grid(N, N, N, S, sc["planes"], 0.5, sh_active=sh_active, adaptive=adapt)
    else:
        g = device_sparse_grid(N, S, sc["compact"], sc["occ"], sc["io"], sc["cage"], sc["c2d"],
                               dev, march=march)
        g.sh_active = sh_active
        go = oracle.Grid(N, N, N, S, sc["compact"], 0.5, occupancy=sc["occ"],
                         index_offsets=sc["io"], sh_active=sh_active, in_cage=sc["cage"],
                         adaptive=adapt)
    return g, go


def _rays(sc):
    return pxl.Rays.from_arrays(sc["rays_np"], device=sc["dev"], white_bg=True)


@pytest.fixture(params=["group", "wave"])
def marcher(request):
    pxl.set_option(pxl.OPT_MARCHER,
                   pxl.MARCHER_WAVE if request.param == "wave" else pxl.MARCHER_GROUP)
    yield request.param
    pxl.set_option(pxl.OPT_MARCHER, pxl.MARCHER_GROUP)


def _train(sc, grid, optim, path, huber=0.1):
    import torch

    rays = _rays(sc)
    losses = torch.zeros(NR, dtype=torch.float64, device=sc["dev"])
    if path == "logged":
        ws = pxl.train_workspace(NR, 512, sc["dev"])
        pxl.train_rays_ws(grid, optim, rays, losses, sc["si"], ws, 512, huber_delta=huber)
    elif path == "fused":
        pxl.train_rays(grid, optim, rays, losses, sc["si"], huber_delta=huber)
    else:
        pxl.render_fwd(grid, rays, sc["si"])
        pxl.render_bwd_rgb(grid, optim, rays, losses, sc["si"], huber_delta=huber)
    torch.cuda.synchronize()
    return rays, losses.cpu().numpy()


@pytest.mark.parametrize("kind", ["dense", "caged"])
@pytest.mark.parametrize("path", ["fused", "logged", "separate"])
def test_band_backward(sc, kind, path, marcher):
    """Band 1 of an SH-2 grid in band mode: colour from k < 4, gradients of
    the density and k in [1, 4)."""
    say(f"[band] {kind}, {path}, {marcher} marcher: sh_active 4, sh_first 1")
    grid, go = _grids(sc, kind, sh_active=4)
    go.sh_first = 1
    optim = pxl.Optim(grid, sh_first=1)
    rays, losses = _train(sc, grid, optim, path)
    ro = oracle.Rays(sc["rays_np"])
    fwd = oracle.render_fwd(go, ro, sc["batch"])
    if path == "separate":
        fwd = {k_o: getattr(rays, k_g).cpu().numpy().astype(np.float64)
               for k_g, k_o in (("acc_r", "r"), ("acc_g", "g"), ("acc_b", "b"), ("acc_w", "w"))}
    else:
        for k_g, k_o in (("acc_r", "r"), ("acc_g", "g"), ("acc_b", "b"), ("acc_w", "w")):
            err = np.abs(getattr(rays, k_g).cpu().numpy()[sc["batch"]] - fwd[k_o][sc["batch"]])
            assert err.max() <= 1e-5, (k_g, err.max())
    l_ref, g_ref = oracle.render_bwd(go, ro, sc["batch"], fwd, 1.0, 0.1)
    assert np.abs(losses - l_ref).max() <= 255 * 1e-5
    g = from_records(optim.grad, 1 + 3 * S)
    live = np.zeros(1 + 3 * S, bool)
    live[0] = True
    for ch in range(3):
        live[1 + ch * S + 1:1 + ch * S + 4] = True
    assert np.abs(g[~live]).max() == 0.0
    worst = grad_close(g, g_ref, floor=1e-3 if path != "separate" else 2e-3)
    say(f"    worst gradient error / tolerance = {worst:.3g}")
    assert worst <= 1.0


@pytest.mark.parametrize("kind", ["dense", "caged"])
@pytest.mark.parametrize("tv_type", [pxl.TV_NEIGHBOUR, pxl.TV_BOTH])
@pytest.mark.parametrize("band", [None, (4, 9)])
def test_variance_types(sc, kind, tv_type, band):
    import torch

    say(f"[tv] type {tv_type}, {kind}, band {band}")
    grid, go = _grids(sc, kind, sh_active=0 if band is None else band[1])
    optim = pxl.Optim(grid, sh_first=0 if band is None else band[0])
    go.sh_first = optim.sh_first
    pxl.add_variance_loss(grid, optim, 1e-2, 1e-3, tv_type, 6, optimize_step=0)
    torch.cuda.synchronize()
    if tv_type == pxl.TV_NEIGHBOUR:
        ref = oracle.tv_neighbour6(go, 1e-2, 1e-3)
    else:
        ref = oracle.tv_both(go, 1e-2, 1e-3, 0)
    worst = grad_close(from_records(optim.grad, grid.P), ref)
    say(f"    worst TV error / tolerance = {worst:.3g}")
    assert worst <= 1.0


def test_neighbour6_caged_record0(sc):
    """The caged 6-neighbour term when dense vertex 0 is in the cage: every
    interior vertex outside the cage then adds into record 0 (the
    reference's test at the record number, RenderUtils.cu:2669)."""
    import torch

    dev = sc["dev"]
    cage = sc["cage"].copy()
    cage[0] = 1
    io, n = oracle.build_index(cage, N, N, N)
    c2d = np.flatnonzero(cage).astype(np.uint32)
    compact = np.ascontiguousarray(sc["planes"][:, c2d])
    grid = device_sparse_grid(N, S, compact, sc["occ"], io, cage, c2d, dev)
    go = oracle.Grid(N, N, N, S, compact, 0.5, occupancy=sc["occ"], index_offsets=io,
                     in_cage=cage)
    optim = pxl.Optim(grid)
    pxl.add_variance_loss(grid, optim, 1e-2, 1e-3, pxl.TV_NEIGHBOUR, 6)
    torch.cuda.synchronize()
    ref = oracle.tv_neighbour6(go, 1e-2, 1e-3)
    assert np.abs(ref[:, 0]).max() > 0.0
    worst = grad_close(from_records(optim.grad, grid.P), ref)
    say(f"[tv] caged record 0: worst TV error / tolerance = {worst:.3g}")
    assert worst <= 1.0


def test_adam_band_and_init(dev):
    import torch

    rng = np.random.default_rng(41)
    P, n = 1 + 3 * S, 20_000
    planes = rng.normal(size=(P, n)).astype(np.float32)
    planes[0] = np.abs(planes[0])
    grid = pxl.Grid(n, 1, 1, S, to_records(planes, dev), sh_active=9)
    optim = pxl.Optim(grid, sh_first=4)
    m = np.zeros_like(planes)
    v = np.zeros_like(planes)
    p_o = planes.copy()
    for t in range(1, 4):
        g = (rng.normal(size=(P, n)) * 1e-2).astype(np.float32)
        optim.grad[:, :P] = torch.from_numpy(g.T.copy()).to(dev)
        pxl.adam_step(grid, optim, 0.3, 1.0, t)
        oracle.adam(p_o, g, m, v, 0.3, 1.0, t, band=(4, 9))
    torch.cuda.synchronize()
    for got, exp in ((grid.params, p_o), (optim.m, m), (optim.v, v)):
        np.testing.assert_allclose(from_records(got, P), exp, rtol=1e-6, atol=1e-7)
    # Band initialisation: band 2 keeps the density state, clears k in [4, 9).
    stride = optim.stride
    for t_ in (optim.grad, optim.m, optim.v):
        t_.fill_(1.0)
    pxl.optim_init_band(grid, optim, 2, band_mode=True, density_initialized=True)
    torch.cuda.synchronize()
    cleared = np.zeros(stride, bool)
    for ch in range(3):
        cleared[1 + ch * S + 4:1 + ch * S + 9] = True
    for t_ in (optim.grad, optim.m, optim.v):
        a = t_.cpu().numpy()
        assert np.all(a[:, cleared] == 0.0) and np.all(a[:, ~cleared] == 1.0)
    pxl.optim_init_band(grid, optim, 0, band_mode=True, density_initialized=True)
    torch.cuda.synchronize()
    assert float(optim.m[:, 0].abs().max()) == 0.0 and float(optim.m[:, 1 + S].abs().max()) == 0.0
    pxl.optim_init_band(grid, optim, 2, band_mode=False)
    torch.cuda.synchronize()
    assert float(optim.v.abs().max()) == 0.0


def test_init_band_two_strides(dev):
    """pxl_optim_init_band with gradient records (36 floats) wider than the
    moment records (28): each tensor is cleared at its own stride, the band's
    columns only in band mode, whole records (padding included) otherwise."""
    import torch

    S, n = 9, 5000
    P = 1 + 3 * S
    planes = np.random.default_rng(5).normal(size=(P, n)).astype(np.float32)
    grid = pxl.Grid(n, 1, 1, S, to_records(planes, dev), sh_active=9)
    optim = pxl.Optim(grid, stride=28, grad_stride=36)
    for t_ in (optim.grad, optim.m, optim.v):
        t_.fill_(1.0)
    pxl.optim_init_band(grid, optim, 1, band_mode=True, density_initialized=True)
    torch.cuda.synchronize()
    for t_ in (optim.grad, optim.m, optim.v):
        a = t_.cpu().numpy()
        cleared = np.zeros(a.shape[1], bool)
        for ch in range(3):
            cleared[1 + ch * S + 1:1 + ch * S + 4] = True
        assert np.all(a[:, cleared] == 0.0) and np.all(a[:, ~cleared] == 1.0)
    pxl.optim_init_band(grid, optim, 1, band_mode=False)
    torch.cuda.synchronize()
    for t_ in (optim.grad, optim.m, optim.v):
        assert float(t_.abs().max()) == 0.0


def test_learn_level_matches_stage_calls(sc):
    """pxl_trainer_learn_level with SHsForTargetResolutionOnly above the
    minimum hierarchy: bands 0, 1, 2 in turn, each from a cleared band state
    with learning rates and time step restarted (Render.cu:264-315), 1 epoch
    of 2 batches each; against the same stages issued one by one."""
    import torch

    dev = sc["dev"]
    B = 512
    order = sc["si"]
    kw = dict(learning_rate=0.1, sh_learning_rate=0.2, variance_coefficient=1e-5,
              sh_coefficient=1e-6, variance_interval=3)
    g0 = pxl.Grid(N, N, N, S, to_records(sc["planes"], dev), step=0.5)
    o0 = pxl.Optim(g0)
    tr = pxl.Trainer(g0, o0, _rays(sc), order, batch_size=B, fused=True, log_samples=512, **kw)
    tr.learn_level(1, 2, sh_for_target_only=True, current_hierarchy=1, minimum_hierarchy=0)
    t_, lr_, _ = tr.state()
    assert t_ == 2
    tr.close()
    g1 = pxl.Grid(N, N, N, S, to_records(sc["planes"], dev), step=0.5)
    o1 = pxl.Optim(g1)
    rays = _rays(sc)
    losses = torch.zeros(B, dtype=torch.float64, device=dev)
    for band in range(3):
        pxl.optim_init_band(g1, o1, band, band_mode=True, density_initialized=True)
        g1.sh_active = (band + 1) ** 2
        o1.sh_first = band * band
        lr, sl = 0.1, 0.2
        for t in range(2):
            pxl.train_rays(g1, o1, rays, losses, order, start=t * B, length=B)
            pxl.add_variance_loss(g1, o1, 1e-5, 1e-6, pxl.TV_DISPARITY, 6, 0)
            pxl.adam_step(g1, o1, lr, sl, t + 1)
            d = math.pow(2.0, 1.0 / 2 / 1.0)
            lr, sl = lr / d, sl / d
    torch.cuda.synchronize()
    assert abs(lr_ - lr) < 1e-15
    diff = (g0.params - g1.params).abs()
    say(f"[level] bands 0..2: max |param diff| {float(diff.max()):.3g}")
    assert float((diff > 1e-4).float().mean()) <= 1e-4, float(diff.max())
    # The frozen lower bands stayed as they were after their own band.
    assert float(diff.max()) < 1e-2


@pytest.mark.parametrize("kind", ["dense", "caged"])
@pytest.mark.parametrize("path", ["fused", "logged"])
def test_nan_parameter(sc, kind, path):
    """A NaN density in one record: rays through it get loss 255 and no
    gradient (RenderUtils.cu:2158-2162); every other ray's loss and the
    gradient equal the oracle's, which applies the same rule."""
    import torch

    grid, go = _grids(sc, kind)
    planes = go.planes.copy()
    # A record the batch touches: a corner of the cell at the middle of the
    # first ray of the batch that is there in an occupied cell.
    r_nan = None
    rn = sc["rays_np"]
    for ray in sc["batch"]:
        t = 0.5 * (float(rn["near_t"][ray]) + float(rn["far_t"][ray]))
        p = np.floor([rn[o][ray] + t * rn[d][ray] for o, d in (("ox", "dx"), ("oy", "dy"),
                                                                ("oz", "dz"))]).astype(int)
        if np.any(p < 0) or np.any(p > N - 2):
            continue
        cell = p[2] * (N - 1) ** 2 + p[1] * (N - 1) + p[0]
        if kind == "caged" and not sc["occ"][cell]:
            continue
        v = (p[2] + 1) * N * N + (p[1] + 1) * N + (p[0] + 1)
        r_nan = int(sc["io"][v]) if kind == "caged" else v
        if r_nan != 0:
            break
    assert r_nan
    planes[0, r_nan] = np.nan
    grid.params[r_nan, 0] = float("nan")
    go = oracle.Grid(N, N, N, S, planes, 0.5, occupancy=go.occupancy,
                     index_offsets=go.index_offsets)
    optim = pxl.Optim(grid)
    _, losses = _train(sc, grid, optim, path)
    ro = oracle.Rays(sc["rays_np"])
    fwd = oracle.render_fwd(go, ro, sc["batch"])
    l_ref, g_ref = oracle.render_bwd(go, ro, sc["batch"], fwd, 1.0, 0.1)
    hit = l_ref == 255.0
    say(f"[nan] {kind} {path}: {int(hit.sum())} rays through the NaN record")
    assert hit.sum() > 0 and np.array_equal(losses == 255.0, hit)
    assert np.abs(losses[~hit] - l_ref[~hit]).max() <= 255 * 1e-5
    g = from_records(optim.grad, grid.P)
    assert np.isfinite(g).all()
    # Record 0 of a caged grid: the single fused kernel adds each ray's sum
    # with a float atomic (1e-3 relative), the logged step folds FP64 sums.
    assert grad_close_alias(g, g_ref, rel0=1e-3 if path == "fused" else 1e-4) <= 1.0
    out = torch.zeros(2, dtype=torch.float64, device=sc["dev"])
    pxl.loss_reduce(torch.from_numpy(losses).to(sc["dev"]), out)
    torch.cuda.synchronize()
    assert np.isfinite(out[0].item())
    nan_l = torch.tensor([1.0, float("nan")], dtype=torch.float64, device=sc["dev"])
    pxl.loss_reduce(nan_l, out)
    torch.cuda.synchronize()
    assert math.isnan(out[0].item())


@pytest.mark.parametrize("case", ["adaptive", "sh16", "caged", "caged-inline", "caged-fused",
                                  "adaptive-fused"])
def test_trainer_overlap_matches_stage_calls(sc, case):
    """The trainer with the logged step and the variance loss on its second
    stream equals the stage calls (render, variance loss, Adam per batch):
    on the adaptive rule and SH degree 3 (whose render writes gradients in
    the forward kernel, so the variance loss must not run beside it) and on
    a caged grid with an interval of 3 (side stream on, and off); and with
    the variance loss fused into Adam (caged, adaptive), read back through
    pxl_trainer_sync."""
    import torch

    dev = sc["dev"]
    B = 512
    order = sc["si"]
    if case == "sh16":
        S16 = 16
        planes = c1_planes(N, S16, n_blobs=8, seed=31)
        mk = lambda: pxl.Grid(N, N, N, S16, to_records(planes, dev), step=0.5)
    elif case.startswith("adaptive"):
        mk = lambda: _grids(sc, "dense", march=pxl.MARCH_REF_ADAPTIVE)[0]
    else:
        mk = lambda: _grids(sc, "caged")[0]
    steps = 4
    g0, g1 = mk(), mk()
    o0, o1 = pxl.Optim(g0), pxl.Optim(g1)
    kw = dict(learning_rate=0.1, sh_learning_rate=0.2, variance_coefficient=1e-3,
              sh_coefficient=1e-4, variance_interval=3)
    tr = pxl.Trainer(g0, o0, _rays(sc), order, batch_size=B, fused=True, log_samples=512,
                     tv_inline=case == "caged-inline", tv_separate=not case.endswith("fused"),
                     **kw)
    for _ in range(steps):
        tr.step()
    tr.sync()
    rays = _rays(sc)
    losses = torch.zeros(B, dtype=torch.float64, device=dev)
    ws = pxl.train_workspace(B, 512, dev)
    lr, sl = 0.1, 0.2
    bc = (order.numel() + B - 1) // B
    for t in range(steps):
        b = t % bc
        pxl.train_rays_ws(g1, o1, rays, losses, order, ws, 512, start=b * B, length=B)
        if not g1.sparse or b % 3 == 0:
            pxl.add_variance_loss(g1, o1, 1e-3, 1e-4, pxl.TV_DISPARITY, 6, 0)
        pxl.adam_step(g1, o1, lr, sl, t + 1)
        d = math.pow(2.0, 1.0 / bc / (t // bc + 1.0))
        lr, sl = lr / d, sl / d
    torch.cuda.synchronize()
    mse, _ = tr.loss()
    assert np.isfinite(mse)
    diff = (g0.params - g1.params).abs()
    say(f"[trainer] {case}: max |param diff| {float(diff.max()):.3g}")
    assert float((diff > 1e-4).float().mean()) <= 1e-4, float(diff.max())
    tr.close()
