"""Parity of the HIP path (libpxl.so on the GPU) with the CPU oracle.

Bars: integer and index work bit-exact (rays, occupancy bits, in-cage
flags, sparse index, per-ray sample count and cell-sequence hash); rendered
pixels within 1e-5 absolute of the FP64 oracle; gradients within 1e-4
relative (helpers.grad_close: |g - ref| <= 1e-4 (|ref| + 1e-3 max|ref|) per
plane, the floor covering entries that cancel to ~0 under FP32 atomics).
"""
import sys

import os

import numpy as np
import pytest

import oracle
import pxl
from helpers import (brick_bits, c1_planes, cameras, device_sparse_grid, from_records,
                     grad_close, oracle_rays, pack_bits, sparse_from_dense, to_records)

pytestmark = pytest.mark.gpu

N1, S1, IMG1 = 128, 9, 200


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c1(dev):
    """Small plain-method workload: 128^3 dense SH-2, 20 views 200x200 on a
    radius-4 sphere, fov 40 deg; a 4096-ray batch drawn with seed 3."""
    import torch

    say("[c1] building scene")
    planes = c1_planes(N1, S1)
    cams = cameras(20, IMG1, 4.0, 40.0, N1)
    bbox = (0, N1 - 1, 0, N1 - 1, 0, N1 - 1)
    rays_np, imgs = oracle_rays(cams, IMG1, bbox, seed=9)
    nrays = len(rays_np["ox"])
    batch = np.random.default_rng(3).choice(nrays, 4096, replace=False).astype(np.int32)
    return dict(planes=planes, cams=cams, bbox=bbox, rays_np=rays_np, imgs=imgs, batch=batch,
                rec=to_records(planes, dev), dev=dev,
                si=torch.from_numpy(batch).to(dev))


@pytest.fixture(params=["group", "wave"])
def marcher(request):
    """Fixed-rule marcher under test: eight samples per wave (default) or
    one sample per wave (library option PXL_OPT_MARCHER)."""
    pxl.set_option(pxl.OPT_MARCHER,
                   pxl.MARCHER_WAVE if request.param == "wave" else pxl.MARCHER_GROUP)
    yield request.param
    pxl.set_option(pxl.OPT_MARCHER, pxl.MARCHER_GROUP)


def device_rays(rays_np, dev, white_bg=True):
    return pxl.Rays.from_arrays(rays_np, device=dev, white_bg=white_bg)


def test_rays_bit_exact(c1):
    import torch

    say("[rays] 20 views 200x200 vs oracle")
    dev = c1["dev"]
    n = len(c1["rays_np"]["ox"])
    rays = pxl.Rays(n, device=dev)
    for v, (cam, c2w, ctr, f, c) in enumerate(c1["cams"]):
        img = tuple(torch.from_numpy(x).to(dev) for x in c1["imgs"][v])
        pxl.rays_init(cam, c1["bbox"], IMG1, IMG1, rays, start_index=v * IMG1 * IMG1, img=img)
    torch.cuda.synchronize()
    for k in ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t"):
        got = getattr(rays, k).cpu().numpy().view(np.uint32)
        exp = c1["rays_np"][k].view(np.uint32)
        assert np.array_equal(got, exp), (k, int((got != exp).sum()))
    for k in ("gt_r", "gt_g", "gt_b", "gt_mask"):
        assert np.array_equal(getattr(rays, k).cpu().numpy(), c1["rays_np"][k]), k


@pytest.mark.parametrize("case", ["tiny", "edges", "c1"])
def test_occupancy_and_index_bit_exact(dev, c1, case):
    import torch

    say(f"[occupancy] {case}")
    if case == "tiny":
        D, H, W = 9, 10, 11
        dens = np.random.default_rng(4).uniform(0, 10, (D, H, W)).astype(np.float32)
        tau = 8.0
    elif case == "edges":
        # occupied cells touching every face (exercises the face rules)
        D = H = W = 17
        dens = np.zeros((D, H, W), np.float32)
        dens[-1, :, :] = 9
        dens[:, -1, :] = 9
        dens[:, :, -1] = 9
        dens[3, 4, 5] = 9
        tau = 8.0
    else:
        D = H = W = N1
        dens = c1["planes"][0].reshape(N1, N1, N1)
        tau = 8.0
    occ_o, cage_o, n_o = oracle.build_occupancy(dens, W, H, D, tau)
    io_o, cnt_o = oracle.build_index(cage_o, W, H, D)
    ncell = (W - 1) * (H - 1) * (D - 1)
    d_dens = torch.from_numpy(dens.reshape(-1)).to(dev)
    occ = torch.zeros((ncell + 31) // 32, dtype=torch.int32, device=dev)
    cage = torch.zeros(W * H * D, dtype=torch.uint8, device=dev)
    nocc = torch.zeros(1, dtype=torch.int64, device=dev)
    pxl.build_occupancy(d_dens, W, H, D, tau, occ, cage, nocc)
    io = torch.zeros(W * H * D, dtype=torch.int32, device=dev)
    c2d = torch.zeros(max(1, W * H * D), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    scratch = torch.zeros(pxl.sparse_index_scratch_bytes(W * H * D) // 4 + 1, dtype=torch.int32,
                          device=dev)
    pxl.build_sparse_index(cage, W, H, D, io, c2d, cnt, scratch)
    torch.cuda.synchronize()
    assert int(nocc.item()) == n_o
    assert np.array_equal(occ.cpu().numpy().view(np.uint32), pack_bits(occ_o))
    assert np.array_equal(cage.cpu().numpy(), cage_o)
    assert int(cnt.item()) == cnt_o
    assert np.array_equal(io.cpu().numpy().view(np.uint32), io_o)
    assert np.array_equal(c2d.cpu().numpy()[:cnt_o].view(np.uint32),
                          np.flatnonzero(cage_o).astype(np.uint32))
    br = pxl.build_brick_occupancy(occ, W, H, D)
    torch.cuda.synchronize()
    exp = brick_bits(occ_o, W, H, D)
    got = br.cpu().numpy().view(np.uint32)
    assert len(got) == 2 * len(exp) == pxl.lib().pxl_brick_words(W, H, D)
    assert np.array_equal(got[:len(exp)], exp)
    # The second half: the map dilated by one brick.
    assert np.array_equal(got[len(exp):], brick_bits(occ_o, W, H, D, dilated=True))


def _fwd_compare(grid_o, rays_np, batch, got, trace_gpu, white=True):
    rays_o = oracle.Rays(rays_np, white_bg=white)
    ref = oracle.render_fwd(grid_o, rays_o, batch, trace=True)
    idx = batch
    worst = 0.0
    for k_g, k_o in (("acc_r", "r"), ("acc_g", "g"), ("acc_b", "b"), ("acc_w", "w")):
        err = np.abs(got[k_g][idx].astype(np.float64) - ref[k_o][idx]).max()
        assert err <= 1e-5, (k_g, err)
        worst = max(worst, err)
    say(f"    max pixel / opacity error {worst:.2e}")
    dref = ref["density"][idx]
    derr = np.abs(got["acc_density"][idx] - dref) / np.maximum(1.0, np.abs(dref))
    assert derr.max() <= 1e-5, derr.max()
    tr = trace_gpu.reshape(-1, 2)
    assert np.array_equal(tr[:, 0], ref["trace"][:, 0]), "sample counts differ"
    assert np.array_equal(tr[:, 1], ref["trace"][:, 1]), "cell sequences differ"
    return ref


def _gpu_fwd(grid, rays, si, dev):
    import torch

    tr = torch.zeros(si.numel() * 2, dtype=torch.int32, device=dev)
    pxl.render_fwd(grid, rays, si, trace=tr)
    torch.cuda.synchronize()
    got = {k: getattr(rays, k).cpu().numpy() for k in pxl._RAY_OUT}
    return got, tr.cpu().numpy().view(np.uint32)


def test_render_fwd_dense_c1(c1, marcher):
    say(f"[fwd] dense 128^3 SH-2, 4096 rays, {marcher} marcher")
    dev = c1["dev"]
    grid = pxl.Grid(N1, N1, N1, S1, c1["rec"], step=0.5)
    rays = device_rays(c1["rays_np"], dev)
    got, tr = _gpu_fwd(grid, rays, c1["si"], dev)
    grid_o = oracle.Grid(N1, N1, N1, S1, c1["planes"], 0.5)
    ref = _fwd_compare(grid_o, c1["rays_np"], c1["batch"], got, tr)
    assert ref["trace"][:, 0].max() > 300  # long rays are exercised


@pytest.fixture(scope="module")
def sparse1(c1):
    planes = c1["planes"]
    occ, cage, io, n, c2d, compact = sparse_from_dense(planes, N1, 12.0, c1["dev"])
    return dict(occ=occ, cage=cage, io=io, n=n, c2d=c2d, compact=compact)


@pytest.mark.parametrize("bricks", [True, False])
def test_render_fwd_sparse(c1, sparse1, bricks, marcher):
    s = sparse1
    say(f"[fwd] sparse 128^3, {s['n']} caged vertices, brick map {bricks}, {marcher} marcher")
    dev = c1["dev"]
    grid = device_sparse_grid(N1, S1, s["compact"], s["occ"], s["io"], s["cage"], s["c2d"], dev,
                              bricks=bricks)
    rays = device_rays(c1["rays_np"], dev)
    got, tr = _gpu_fwd(grid, rays, c1["si"], dev)
    grid_o = oracle.Grid(N1, N1, N1, S1, s["compact"], 0.5, occupancy=s["occ"],
                         index_offsets=s["io"])
    _fwd_compare(grid_o, c1["rays_np"], c1["batch"], got, tr)


@pytest.mark.parametrize("step", [0.3, 1.0, 2.5])
def test_render_fwd_sparse_steps(c1, sparse1, step):
    """Other step sizes on the caged grid with the brick map: the coarse
    probe's stride (positions per 8 cells: 26, 8, 3) changes with the step,
    the evaluated samples must not."""
    s = sparse1
    say(f"[fwd] sparse 128^3, step {step}, brick map on")
    dev = c1["dev"]
    grid = device_sparse_grid(N1, S1, s["compact"], s["occ"], s["io"], s["cage"], s["c2d"], dev,
                              step=step)
    rays = device_rays(c1["rays_np"], dev)
    got, tr = _gpu_fwd(grid, rays, c1["si"], dev)
    grid_o = oracle.Grid(N1, N1, N1, S1, s["compact"], step, occupancy=s["occ"],
                         index_offsets=s["io"])
    _fwd_compare(grid_o, c1["rays_np"], c1["batch"], got, tr)


def test_render_fwd_adaptive(c1):
    say("[fwd] dense 128^3 adaptive step, 1024 rays")
    dev = c1["dev"]
    grid = pxl.Grid(N1, N1, N1, S1, c1["rec"], step=0.5, march_mode=pxl.MARCH_REF_ADAPTIVE)
    rays = device_rays(c1["rays_np"], dev)
    si = c1["si"][:1024]
    got, tr = _gpu_fwd(grid, rays, si, dev)
    grid_o = oracle.Grid(N1, N1, N1, S1, c1["planes"], 0.5, adaptive=True)
    _fwd_compare(grid_o, c1["rays_np"], c1["batch"][:1024], got, tr)


# Backward entry points: the separate forward + backward kernels, the fused
# marcher, and the two-kernel logged step with a log that holds every ray
# (512 samples, several 64-sample chunks per ray) or only short rays (24
# samples: the longer rays go through the fused marcher's redo launch).
# The logged step also runs with optimiser records of 28 floats and with
# gradient records of 36 floats beside moments of 28 (the gradient walk
# flushes rows at the gradient's own stride).
BWD_PATHS = ["separate", "fused", "logged", "logged_24", "logged_o28", "logged_g36"]


def _bwd_case(c1, grid, grid_o, nr, path, huber=0.1):
    import torch

    fused = path != "separate"
    dev = c1["dev"]
    rays = device_rays(c1["rays_np"], dev)
    si = c1["si"][:nr]
    batch = c1["batch"][:nr]
    if path in ("logged_o28", "logged_g36"):
        optim = pxl.Optim(grid, stride=28, grad_stride=36 if path == "logged_g36" else None)
        optim.grad.fill_(0.0)
    else:
        optim = pxl.Optim(grid)
    losses = torch.zeros(nr, dtype=torch.float64, device=dev)
    rays_o = oracle.Rays(c1["rays_np"])
    if path.startswith("logged"):
        cap = 24 if path == "logged_24" else 512
        ws = pxl.train_workspace(nr, cap, dev)
        ws.fill_(0xAB)  # stale contents must not leak into the result
        pxl.train_rays_ws(grid, optim, rays, losses, si, ws, cap, mask_coef=1.0,
                          huber_delta=huber)
        torch.cuda.synchronize()
        fwd = oracle.render_fwd(grid_o, rays_o, batch)
        if (pxl.get_option(pxl.OPT_MARCHER) == pxl.MARCHER_GROUP
                and grid.march_mode == pxl.MARCH_FIXED):
            # Per-ray counts follow 256 bytes of claim counters and the FP64
            # totals (32 bytes per ray, 256-byte aligned): both the logged
            # and the redo route must be taken.
            off = 256 + (nr * 32 + 255) // 256 * 256
            cnt = ws[off:off + 4 * nr].view(torch.int32).cpu().numpy()
            n_logged, n_redo = int((cnt > 0).sum()), int((cnt == -1).sum())
            say(f"    logged rays {n_logged}, redone rays {n_redo}")
            assert n_logged > 0 and set(np.unique(cnt)) <= set(range(-1, cap + 1))
            assert (n_redo > 0) == (cap == 24)
    elif fused:
        pxl.train_rays(grid, optim, rays, losses, si, mask_coef=1.0, huber_delta=huber)
        torch.cuda.synchronize()
        fwd = oracle.render_fwd(grid_o, rays_o, batch)
    else:
        # The separate kernels exchange C and A through the FP32 ray planes
        # (RenderUtils.cu:1709-1716), as in the reference; the oracle reads
        # the same FP32 values.
        pxl.render_fwd(grid, rays, si)
        torch.cuda.synchronize()
        fwd = {k_o: getattr(rays, k_g).cpu().numpy().astype(np.float64)
               for k_g, k_o in (("acc_r", "r"), ("acc_g", "g"), ("acc_b", "b"), ("acc_w", "w"))}
        pxl.render_bwd_rgb(grid, optim, rays, losses, si, mask_coef=1.0, huber_delta=huber)
        torch.cuda.synchronize()
    l_ref, g_ref = oracle.render_bwd(grid_o, rays_o, batch, fwd, mask_coef=1.0,
                                     huber_delta=huber)
    l_got = losses.cpu().numpy()
    assert np.abs(l_got - l_ref).max() <= 255 * 1e-5
    g = from_records(optim.grad, grid.P)
    rel = max(float(np.abs(g[p] - g_ref[p]).max() / np.abs(g_ref[p]).max())
              for p in range(g_ref.shape[0]) if np.abs(g_ref[p]).max() > 0)
    say(f"    max loss error {np.abs(l_got - l_ref).max():.2e}, max gradient error / plane "
        f"max {rel:.2e}")
    # The separate backward reads C and A rounded to FP32 by the forward, so
    # the suffix C - C_i of late samples carries +-ulp(C)/2 (~6e-8) on
    # either side of the comparison: its floor is 2e-3 of the plane max.
    worst = grad_close(g, g_ref, floor=1e-3 if fused else 2e-3)
    say(f"    worst gradient error / tolerance = {worst:.3g}")
    assert worst <= 1.0
    return optim


@pytest.mark.parametrize("path", BWD_PATHS)
def test_backward_dense(c1, path, marcher):
    say(f"[bwd] dense 128^3, 1024 rays, {path}, {marcher} marcher")
    grid = pxl.Grid(N1, N1, N1, S1, c1["rec"], step=0.5)
    grid_o = oracle.Grid(N1, N1, N1, S1, c1["planes"], 0.5)
    _bwd_case(c1, grid, grid_o, 1024, path)


@pytest.mark.parametrize("path", BWD_PATHS)
def test_backward_sparse(c1, sparse1, path, marcher):
    s = sparse1
    say(f"[bwd] sparse 128^3, 1024 rays, {path}, {marcher} marcher")
    grid = device_sparse_grid(N1, S1, s["compact"], s["occ"], s["io"], s["cage"], s["c2d"],
                              c1["dev"])
    grid_o = oracle.Grid(N1, N1, N1, S1, s["compact"], 0.5, occupancy=s["occ"],
                         index_offsets=s["io"])
    _bwd_case(c1, grid, grid_o, 1024, path)


def _order_keys_np(rays_np, ids, N, step, mode):
    """The batch-order key of each ray id, restated in numpy."""
    ox, oy, oz = (rays_np[k][ids].astype(np.float64) for k in ("ox", "oy", "oz"))
    dx, dy, dz = (rays_np[k][ids].astype(np.float32).astype(np.float64) for k in ("dx", "dy", "dz"))
    tn, tf = rays_np["near_t"][ids].astype(np.float64), rays_np["far_t"][ids].astype(np.float64)
    t0 = np.maximum(tn, 0.0)

    def q(p):
        return np.clip(p / (N - 1) * 1023.0, 0.0, 1023.0).astype(np.uint64)

    def spread(v):
        out = np.zeros_like(v)
        for bit in range(10):
            out |= ((v >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit)
        return out

    key = spread(q(ox + t0 * dx)) | (spread(q(oy + t0 * dy)) << np.uint64(1)) | \
        (spread(q(oz + t0 * dz)) << np.uint64(2))
    if mode == pxl.ORDER_LENGTH_MORTON:
        n = (tf - tn) / step
        steps = np.where(n < 1.0, 0, np.minimum(n, 4.0e9)).astype(np.uint64)
        bits = np.array([int(v).bit_length() for v in steps], dtype=np.uint64)
        key |= (np.uint64(15) - np.minimum(bits, np.uint64(15))) << np.uint64(30)
    key[tn > tf] = np.uint64(1) << np.uint64(40)
    return key


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_order(c1, mode):
    """pxl_batch_order returns a permutation of the batch whose keys, restated
    in numpy, are non-decreasing (bit-exact keys)."""
    import torch

    say(f"[order] batch order mode {mode}, 4096 rays")
    dev = c1["dev"]
    rays = device_rays(c1["rays_np"], dev)
    grid = pxl.Grid(N1, N1, N1, S1, c1["rec"], step=0.5)
    si = c1["si"]
    out = torch.empty_like(si)
    pxl.batch_order(grid, rays, si, out, mode=mode)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert sorted(got.tolist()) == sorted(si.cpu().numpy().tolist())
    keys = _order_keys_np(c1["rays_np"], got, N1, 0.5, mode)
    assert np.all(keys[1:] >= keys[:-1])


@pytest.mark.parametrize("sched", [(1, False, 0, 0, 0), (2, True, 1, 0, 4),
                                   (4, True, 2, 0, 0), (4, False, 1, 0, 1),
                                   (1, False, 0, 0, 4), (4, False, 0, 1, 0),
                                   (1, False, 2, 1, 2), (1, False, 0, 2, 1),
                                   (4, True, 1, 2, 4), (2, False, 2, 2, 4)])
def test_logged_schedules(c1, sparse1, sched):
    """The two-kernel step under each schedule and batch order matches the
    oracle (sparse 128^3, 1024 rays, log of 64 samples: some rays redone;
    the chunk-parallel gradient with a log of 256: rays of up to 4 chunks;
    the brick-binned gradient with a log of 64)."""
    import torch

    wpb, persist, mode, gmode, gwpb = sched
    chunked, binned = gmode == pxl.GRAD_CHUNKED, gmode == pxl.GRAD_BINNED
    say(f"[sched] waves/block {wpb} (gradient {gwpb or wpb}), persistent {persist}, "
        f"order {mode}, gradient {pxl.GRAD_NAMES[gmode]}")
    s = sparse1
    dev = c1["dev"]
    grid = device_sparse_grid(N1, S1, s["compact"], s["occ"], s["io"], s["cage"], s["c2d"], dev)
    grid_o = oracle.Grid(N1, N1, N1, S1, s["compact"], 0.5, occupancy=s["occ"],
                         index_offsets=s["io"])
    nr = 1024
    rays = device_rays(c1["rays_np"], dev)
    si = c1["si"][:nr]
    if mode:
        ordered = torch.empty_like(si)
        pxl.batch_order(grid, rays, si, ordered, mode=mode)
        si = ordered
    torch.cuda.synchronize()
    batch = si.cpu().numpy().astype(np.int32)
    optim = pxl.Optim(grid)
    losses = torch.zeros(nr, dtype=torch.float64, device=dev)
    cap = 256 if chunked else 64
    ws = pxl.train_workspace(nr, cap, dev, grid=grid if binned else None)
    pxl.train_rays_sched(grid, optim, rays, losses, si, ws, cap, waves_per_block=wpb,
                         chunked=chunked, binned=binned, persistent=persist,
                         grad_waves_per_block=gwpb)
    torch.cuda.synchronize()
    rays_o = oracle.Rays(c1["rays_np"])
    fwd = oracle.render_fwd(grid_o, rays_o, batch)
    l_ref, g_ref = oracle.render_bwd(grid_o, rays_o, batch, fwd, mask_coef=1.0, huber_delta=0.1)
    assert np.abs(losses.cpu().numpy() - l_ref).max() <= 255 * 1e-5
    assert np.abs(rays.acc_r.cpu().numpy()[batch] - fwd["r"][batch]).max() <= 1e-5
    worst = grad_close(from_records(optim.grad, grid.P), g_ref, floor=1e-3)
    say(f"    worst gradient error / tolerance = {worst:.3g}")
    assert worst <= 1.0


def test_backward_adaptive_unclipped(c1):
    say("[bwd] dense adaptive, 256 rays, no clipping")
    grid = pxl.Grid(N1, N1, N1, S1, c1["rec"], step=0.5, march_mode=pxl.MARCH_REF_ADAPTIVE)
    grid_o = oracle.Grid(N1, N1, N1, S1, c1["planes"], 0.5, adaptive=True)
    _bwd_case(c1, grid, grid_o, 256, "fused", huber=10.0)
    # The logged step serves only the fixed rule: it runs as the fused one.
    _bwd_case(c1, grid, grid_o, 256, "logged", huber=10.0)


@pytest.mark.parametrize("kind", ["dense", "sparse"])
def test_tv_disparity(c1, sparse1, kind):
    import torch

    say(f"[tv] disparity {kind}")
    dev = c1["dev"]
    if kind == "dense":
        grid = pxl.Grid(N1, N1, N1, S1, c1["rec"], step=0.5)
        grid_o = oracle.Grid(N1, N1, N1, S1, c1["planes"], 0.5)
    else:
        s = sparse1
        grid = device_sparse_grid(N1, S1, s["compact"], s["occ"], s["io"], s["cage"], s["c2d"],
                                  dev)
        grid_o = oracle.Grid(N1, N1, N1, S1, s["compact"], 0.5, occupancy=s["occ"],
                             index_offsets=s["io"])
    optim = pxl.Optim(grid)
    pxl.add_variance_loss(grid, optim, 1.0, 0.5, pxl.TV_DISPARITY, 6, optimize_step=0)
    torch.cuda.synchronize()
    ref = oracle.tv_disparity(grid_o, 1.0, 0.5, optimize_step=0)
    worst = grad_close(from_records(optim.grad, grid.P), ref)
    say(f"    worst TV error / tolerance = {worst:.3g}")
    assert worst <= 1.0


@pytest.mark.parametrize("kind", ["f32", "bf16", "wide", "dense", "dense-wide"])
@pytest.mark.parametrize("form", ["short", "64"])
def test_tv_disparity_forms(c1, sparse1, kind, form, request):
    """TV in both address forms: the short one (caged: 24-bit record
    offsets; dense: 32-bit offsets through a buffer resource) and the
    64-bit form that grids beyond 2^24 records or 4 GiB of records take
    (forced here by the PXL_OPT_TV_ADDR64 option); on caged FP32 and bfloat16 records,
    on records wider than 32 parameters (SH degree 3: 49 parameters, one
    wave per record), and on dense grids (interior vertices take their own
    path, boundary vertices the general gather)."""
    import torch

    from helpers import to_16bit_planes

    pxl.set_option(pxl.OPT_TV_ADDR64, 1 if form == "64" else 0)
    request.addfinalizer(lambda: pxl.set_option(pxl.OPT_TV_ADDR64, 0))
    dev = c1["dev"]
    if kind.startswith("dense"):
        N, S = (40, 16) if kind == "dense-wide" else (N1, S1)
        planes = c1_planes(N, S, n_blobs=6, seed=5) if kind == "dense-wide" else c1["planes"]
        say(f"[tv] dense {N}^3 SH {S}, {form} offsets")
        grid = pxl.Grid(N, N, N, S, to_records(planes, dev), step=0.5)
        optim = pxl.Optim(grid)
        pxl.add_variance_loss(grid, optim, 1.0, 0.5, pxl.TV_DISPARITY, 6, optimize_step=0)
        torch.cuda.synchronize()
        ref = oracle.tv_disparity(oracle.Grid(N, N, N, S, planes, 0.5), 1.0, 0.5,
                                  optimize_step=0)
        worst = grad_close(from_records(optim.grad, grid.P), ref)
        say(f"    worst TV error / tolerance = {worst:.3g}")
        assert worst <= 1.0
        return
    if kind == "wide":
        N, S = 40, 16
        planes = c1_planes(N, S, n_blobs=6, seed=5)
        occ, cage, io, n, c2d, compact = sparse_from_dense(planes, N, 30.0, dev)
    else:
        N, S = N1, S1
        s = sparse1
        occ, cage, io, n, c2d, compact = (s["occ"], s["cage"], s["io"], s["n"], s["c2d"],
                                          s["compact"])
    say(f"[tv] caged table, {kind} records, {form}-bit offsets, {n} records")
    storage = pxl.STORE_BF16 if kind == "bf16" else pxl.STORE_F32
    rec = to_records(compact, dev)
    if storage != pxl.STORE_F32:
        rec = pxl.records_to_16bit(rec, storage)
        compact = to_16bit_planes(compact, storage)
    grid = pxl.Grid(N, N, N, S, rec, step=0.5, storage=storage,
                    occupancy=torch.from_numpy(pack_bits(occ).view(np.int32)).to(dev),
                    index_offsets=torch.from_numpy(io.view(np.int32)).to(dev),
                    in_cage=torch.from_numpy(cage).to(dev),
                    compact_to_dense=torch.from_numpy(c2d.view(np.int32)).to(dev))
    optim = pxl.Optim(grid)
    pxl.add_variance_loss(grid, optim, 1.0, 0.5, pxl.TV_DISPARITY, 6, optimize_step=0)
    torch.cuda.synchronize()
    grid_o = oracle.Grid(N, N, N, S, compact, 0.5, occupancy=occ, index_offsets=io)
    ref = oracle.tv_disparity(grid_o, 1.0, 0.5, optimize_step=0)
    worst = grad_close(from_records(optim.grad, grid.P), ref)
    say(f"    worst TV error / tolerance = {worst:.3g}")
    assert worst <= 1.0


def test_tv_neighbour6_dense(dev):
    import torch

    say("[tv] neighbour-6 dense 40^3")
    N = 40
    planes = c1_planes(N, 4, n_blobs=4, seed=21)
    grid = pxl.Grid(N, N, N, 4, to_records(planes, dev))
    optim = pxl.Optim(grid)
    pxl.add_variance_loss(grid, optim, 1e-2, 1e-3, pxl.TV_NEIGHBOUR, 6)
    torch.cuda.synchronize()
    ref = oracle.tv_neighbour6(oracle.Grid(N, N, N, 4, planes, 0.5), 1e-2, 1e-3)
    assert grad_close(from_records(optim.grad, grid.P), ref) <= 1.0


@pytest.mark.parametrize("gstride", [None, 36])
def test_adam_matches_oracle(dev, gstride):
    """Adam against the oracle; with a gradient stride of 36 the gradient
    records are wider than the parameter and moment records (28)."""
    import torch

    say(f"[adam] 5 steps, 100k records, gradient stride {gstride or 'default'}")
    rng = np.random.default_rng(11)
    P, n = 28, 100_000
    planes = rng.normal(size=(P, n)).astype(np.float32)
    planes[0] = np.abs(planes[0])
    alr = rng.uniform(0.5, 1.0, n).astype(np.float32)
    grid = pxl.Grid(n, 1, 1, 9, to_records(planes, dev))  # dense: n_vox = W*H*D
    optim = pxl.Optim(grid, adaptive_lr=torch.from_numpy(alr).to(dev), grad_stride=gstride)
    if gstride:
        optim.grad[:, P:] = 7.0  # padding of the gradient records is neither read nor cleared
    m = np.zeros_like(planes)
    v = np.zeros_like(planes)
    p_o = planes.copy()
    for t in range(1, 6):
        g = (rng.normal(size=(P, n)) * (10.0 ** rng.integers(-4, 1))).astype(np.float32)
        optim.grad[:, :P] = torch.from_numpy(g.T.copy()).to(dev)
        pxl.adam_step(grid, optim, 0.3, 1.0, t)
        go = g.copy()
        oracle.adam(p_o, go, m, v, 0.3, 1.0, t, adaptive_lr=alr)
    torch.cuda.synchronize()
    for name, got, exp in (("params", grid.params, p_o), ("m", optim.m, m), ("v", optim.v, v)):
        g = from_records(got, P)
        err = np.abs(g - exp) / (np.abs(exp) + 1e-30)
        i = np.unravel_index(np.argmax(err), err.shape)
        say(f"    {name}: worst rel err {err[i]:.3g} at {i}: gpu {g[i]!r} oracle {exp[i]!r}")
        np.testing.assert_allclose(g, exp, rtol=1e-6, atol=1e-7 if name == "params" else 1e-20)
    assert float(optim.grad[:, :P].abs().max()) == 0.0
    if gstride:
        assert bool((optim.grad[:, P:] == 7.0).all())


def test_loss_reduce(dev):
    import torch

    say("[loss] FP64 reduce")
    x = np.random.default_rng(2).uniform(0, 255, 65536)
    out = torch.zeros(2, dtype=torch.float64, device=dev)
    pxl.loss_reduce(torch.from_numpy(x).to(dev), out)
    torch.cuda.synchronize()
    mse = oracle.mse(x)
    assert abs(out[0].item() - mse) <= 1e-12 * mse
    assert abs(out[1].item() - 10 * np.log10(65025 / mse)) < 1e-10


def test_import_export_compact(dev):
    import torch

    say("[layout] import / export / compact")
    rng = np.random.default_rng(8)
    S, n = 9, 5000
    P = 1 + 3 * S
    planes = rng.normal(size=(P, n)).astype(np.float32)
    d_planes = [torch.from_numpy(planes[p].copy()).to(dev) for p in range(P)]
    rec = torch.full((n, 32), 7.0, dtype=torch.float32, device=dev)
    pxl.grid_import_planes(d_planes, rec)
    back = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(P)]
    pxl.grid_export_planes(rec, back)
    c2d = np.sort(rng.choice(n, 1234, replace=False)).astype(np.int32)
    out = torch.zeros((1234, 32), dtype=torch.float32, device=dev)
    pxl.grid_compact(rec, torch.from_numpy(c2d).to(dev), 1234, out)
    torch.cuda.synchronize()
    r = rec.cpu().numpy()
    assert np.array_equal(r[:, :P], planes.T) and np.all(r[:, P:] == 0)
    for p in range(P):
        assert np.array_equal(back[p].cpu().numpy(), planes[p])
    assert np.array_equal(out.cpu().numpy(), r[c2d])


def test_train_step_c1(c1):
    """One full train step (fused render + backward, TV, Adam) against the
    oracle's composition of the same four stages."""
    import torch

    say("[step] full train step, 128^3 dense, 4096 rays")
    dev = c1["dev"]
    P = 1 + 3 * S1
    rec = c1["rec"].clone()
    grid = pxl.Grid(N1, N1, N1, S1, rec, step=0.5)
    optim = pxl.Optim(grid)
    rays = device_rays(c1["rays_np"], dev)
    losses = torch.zeros(4096, dtype=torch.float64, device=dev)
    pxl.train_rays(grid, optim, rays, losses, c1["si"], mask_coef=1.0, huber_delta=0.1)
    pxl.add_variance_loss(grid, optim, 1e-5, 1e-5, pxl.TV_DISPARITY, 6, 0)
    pxl.adam_step(grid, optim, 0.1, 0.1, 1)
    torch.cuda.synchronize()

    grid_o = oracle.Grid(N1, N1, N1, S1, c1["planes"], 0.5)
    rays_o = oracle.Rays(c1["rays_np"])
    fwd = oracle.render_fwd(grid_o, rays_o, c1["batch"])
    _, g = oracle.render_bwd(grid_o, rays_o, c1["batch"], fwd, 1.0, 0.1)
    oracle.tv_disparity(grid_o, 1e-5, 1e-5, 0, grad=g)
    p_o = c1["planes"].copy()
    g32 = g.astype(np.float32)
    m = np.zeros_like(p_o)
    v = np.zeros_like(p_o)
    oracle.adam(p_o, g32, m, v, 0.1, 0.1, 1)
    got = from_records(rec, P)
    diff = np.abs(got - p_o)
    # Adam's first step moves each parameter by lr * g / (|g| + 1e-6): a
    # gradient that cancels to ~1e-6 can legitimately move differently.
    frac_off = (diff > 1e-4 * 0.1).mean()
    say(f"    max |dparam| = {diff.max():.3g}, fraction beyond 1e-5 = {frac_off:.3g}")
    assert frac_off <= 1e-4
    assert diff.max() <= 0.2 + 1e-6


@pytest.mark.parametrize("log_samples,autotune,tv_mode",
                         [(0, False, "fused"), (512, False, "fused"), (512, True, "overlap"),
                          (512, False, "overlap"), (512, False, "inline"),
                          (512, False, "binned"), (512, True, "fused_o28g32")])
def test_trainer_matches_stage_calls(c1, log_samples, autotune, tv_mode):
    """The C++ batch loop (pxl_trainer_step) issues the same stages as the
    Python-driven sequence: dense grid (TV every batch), 3 batches, LR decay
    lr /= 2^(1/(batchCount (epoch+1))) per batch (Render.cu:499-500); with
    log_samples the trainer runs the two-kernel logged step.  TV modes: fused
    with Adam (the trainer alternates two record buffers: after 3 batches the
    current records are in its own, pxl_trainer_sync brings them back), or
    separate, on the second stream beside the forward (overlap) or after the
    render (inline); "binned": fused TV + Adam with the brick-binned
    gradient schedule; "fused_o28g32": the bench's optimiser layout on the
    trainer's side (moments of 28 floats, gradient rows of 32), autotuned,
    against the stage calls on records of 32."""
    import math

    import torch

    say(f"[trainer] 3 batches vs stage calls, log_samples={log_samples}, autotune={autotune}, "
        f"tv {tv_mode}")
    dev = c1["dev"]
    B = 1024
    order = c1["si"]
    n = order.numel()
    grids, optims = [], []
    for _ in range(2):
        rec = c1["rec"].clone()
        g = pxl.Grid(N1, N1, N1, S1, rec, step=0.5)
        grids.append(g)
        if tv_mode == "fused_o28g32" and not optims:
            optims.append(pxl.Optim(g, stride=28, grad_stride=32))
        else:
            optims.append(pxl.Optim(g))
    rays_a = device_rays(c1["rays_np"], dev)
    rays_b = device_rays(c1["rays_np"], dev)
    tr = pxl.Trainer(grids[0], optims[0], rays_a, order, batch_size=B, learning_rate=0.1,
                     sh_learning_rate=0.2, variance_coefficient=1e-5, sh_coefficient=1e-6,
                     variance_interval=3, fused=True, log_samples=log_samples,
                     autotune=autotune, tv_inline=tv_mode == "inline",
                     tv_separate=tv_mode not in ("fused", "binned", "fused_o28g32"),
                     binned=tv_mode == "binned")
    for _ in range(3):
        tr.step()
    current = tr.grid_c().params
    assert (current != grids[0].params.data_ptr()) == (tv_mode in ("fused", "binned",
                                                                    "fused_o28g32"))
    tr.sync()
    assert tr.grid_c().params == grids[0].params.data_ptr()
    losses = torch.zeros(B, dtype=torch.float64, device=dev)
    lr, sl = 0.1, 0.2
    bc = (n + B - 1) // B
    for t in range(3):
        pxl.train_rays(grids[1], optims[1], rays_b, losses, order, start=t * B, length=B)
        pxl.add_variance_loss(grids[1], optims[1], 1e-5, 1e-6, pxl.TV_DISPARITY, 6, 0)
        pxl.adam_step(grids[1], optims[1], lr, sl, t + 1)
        d = math.pow(2.0, 1.0 / bc / 1.0)
        lr, sl = lr / d, sl / d
    torch.cuda.synchronize()
    if autotune:
        order_, wpb_, pers_, grad_, gwpb_ = tr.schedule()
        say(f"    autotuned: order {order_}, waves/block {wpb_}, persistent {pers_}, "
            f"gradient {pxl.GRAD_NAMES[grad_]}")
        assert order_ in (0, 1, 2) and wpb_ in (1, 2, 4)
    t_, lr_, sl_ = tr.state()
    assert t_ == 3 and abs(lr_ - lr) < 1e-15 and abs(sl_ - sl) < 1e-15
    mse, psnr = tr.loss()
    assert np.isfinite(mse)
    diff = (grids[0].params - grids[1].params).abs()
    # identical kernels; only the order of float atomics differs
    assert float((diff > 1e-4).float().mean()) <= 1e-4, float(diff.max())
    tr.close()


# ------------------------------------------------------- 16-bit storage


@pytest.mark.parametrize("storage", [pxl.STORE_BF16, pxl.STORE_F16])
def test_render_16bit_storage(c1, storage, marcher):
    """16-bit grids: the GPU result matches the FP64 oracle run on the
    widened 16-bit values to the same bars as FP32; the storage error
    itself (vs the FP32 grid) is reported."""
    import torch

    from helpers import to_16bit_planes

    name = "bf16" if storage == pxl.STORE_BF16 else "f16"
    say(f"[{name}] dense 128^3 forward + fused backward, 1024 rays")
    dev = c1["dev"]
    P = 1 + 3 * S1
    rec16 = pxl.records_to_16bit(c1["rec"], storage)
    grid = pxl.Grid(N1, N1, N1, S1, rec16, step=0.5, storage=storage)
    planes16 = to_16bit_planes(c1["planes"], storage)
    got16 = rec16[:, :P].T.contiguous().cpu().numpy()
    if storage == pxl.STORE_BF16:
        from helpers import widen_bf16
        got16 = widen_bf16(got16.view(np.uint16))
    else:
        got16 = got16.view(np.float16).astype(np.float32)
    assert np.array_equal(got16, planes16), "16-bit conversion differs from RNE"
    rays = device_rays(c1["rays_np"], dev)
    si = c1["si"][:1024]
    batch = c1["batch"][:1024]
    got, tr = _gpu_fwd(grid, rays, si, dev)
    grid_o = oracle.Grid(N1, N1, N1, S1, planes16, 0.5)
    ref = _fwd_compare(grid_o, c1["rays_np"], batch, got, tr)
    ref32 = oracle.render_fwd(oracle.Grid(N1, N1, N1, S1, c1["planes"], 0.5),
                              oracle.Rays(c1["rays_np"]), batch)
    serr = max(np.abs(got[k][batch] - ref32[o][batch]).max()
               for k, o in (("acc_r", "r"), ("acc_g", "g"), ("acc_b", "b")))
    say(f"    pixel error vs FP32 grid (storage rounding): {serr:.3g}")
    optim = pxl.Optim(grid)
    losses = torch.zeros(1024, dtype=torch.float64, device=dev)
    pxl.train_rays(grid, optim, rays, losses, si)
    torch.cuda.synchronize()
    _, g_ref = oracle.render_bwd(grid_o, oracle.Rays(c1["rays_np"]), batch,
                                 oracle.render_fwd(grid_o, oracle.Rays(c1["rays_np"]), batch),
                                 1.0, 0.1)
    worst = grad_close(from_records(optim.grad, P), g_ref)
    say(f"    worst gradient error / tolerance = {worst:.3g}")
    assert worst <= 1.0
    # TV on the 16-bit grid
    optim.grad.zero_()
    pxl.add_variance_loss(grid, optim, 1.0, 0.5, pxl.TV_DISPARITY, 6, 0)
    torch.cuda.synchronize()
    assert grad_close(from_records(optim.grad, P), oracle.tv_disparity(grid_o, 1.0, 0.5)) <= 1.0


def test_adam_bf16_master(dev):
    """Adam on a bfloat16 grid with an FP32 master copy: the master follows
    the oracle's FP32 Adam; the grid holds its round-to-nearest-even copy."""
    import torch

    from helpers import round_bf16

    say("[bf16] Adam with master copy")
    rng = np.random.default_rng(12)
    P, n = 28, 20_000
    planes = rng.normal(size=(P, n)).astype(np.float32)
    planes[0] = np.abs(planes[0])
    rec32 = to_records(planes, dev)
    rec16 = pxl.records_to_16bit(rec32, pxl.STORE_BF16)
    grid = pxl.Grid(n, 1, 1, 9, rec16, storage=pxl.STORE_BF16)
    optim = pxl.Optim(grid, master_from=rec32)
    p_o, m, v = planes.copy(), np.zeros_like(planes), np.zeros_like(planes)
    for t in range(1, 4):
        g = rng.normal(size=(P, n)).astype(np.float32)
        optim.grad[:, :P] = torch.from_numpy(g.T.copy()).to(dev)
        pxl.adam_step(grid, optim, 0.1, 0.2, t)
        oracle.adam(p_o, g.copy(), m, v, 0.1, 0.2, t)
    torch.cuda.synchronize()
    master = from_records(optim.master, P).astype(np.float32)
    np.testing.assert_allclose(master, p_o, rtol=1e-6, atol=1e-7)
    stored = rec16[:, :P].T.contiguous().cpu().numpy().view(np.uint16)
    assert np.array_equal(stored, round_bf16(master))


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_adam_16bit_without_master(dev, storage):
    """Adam directly on 16-bit records (no FP32 master): each step is the
    oracle's FP32 Adam on the widened stored values, rounded to nearest
    even; stored values within one 16-bit step of that (a value at a rounding
    boundary may land on either side)."""
    import torch

    from helpers import round_bf16

    say(f"[{storage}] Adam without master copy")
    st = pxl.STORE_BF16 if storage == "bf16" else pxl.STORE_F16
    rng = np.random.default_rng(13)
    P, n = 28, 20_000
    planes = rng.normal(size=(P, n)).astype(np.float32)
    planes[0] = np.abs(planes[0]) + 0.5
    rec16 = pxl.records_to_16bit(to_records(planes, dev), st)
    grid = pxl.Grid(n, 1, 1, 9, rec16, storage=st)
    optim = pxl.Optim(grid)
    m, v = np.zeros_like(planes), np.zeros_like(planes)

    def stored():
        return rec16[:, :P].T.contiguous().cpu().numpy().view(np.uint16)

    def widen(h):
        if storage == "bf16":
            return (h.astype(np.uint32) << 16).view(np.float32)
        return h.view(np.float16).astype(np.float32)

    def narrow(x):
        if storage == "bf16":
            return round_bf16(x)
        return x.astype(np.float16).view(np.uint16)

    for t in range(1, 4):
        torch.cuda.synchronize()
        p_o = widen(stored())
        g = rng.normal(size=(P, n)).astype(np.float32)
        optim.grad[:, :P] = torch.from_numpy(g.T.copy()).to(dev)
        pxl.adam_step(grid, optim, 0.1, 0.2, t)
        oracle.adam(p_o, g.copy(), m, v, 0.1, 0.2, t)
        torch.cuda.synchronize()
        got, exp = stored().astype(np.int32), narrow(p_o).astype(np.int32)
        d = np.abs(got - exp)
        assert d.max() <= 1 and (d > 0).mean() < 1e-3, (t, d.max(), (d > 0).mean())
    # the moments are FP32 and follow the oracle's
    mom = from_records(optim.m, P).astype(np.float32)
    np.testing.assert_allclose(mom, m, rtol=1e-6, atol=1e-7)


# ------------------------------------------------------------ eval image


@pytest.mark.parametrize("kind", ["dense", "sparse"])
def test_render_image(c1, sparse1, kind):
    """pxl_render_image against the oracle's renderImageGPU restatement:
    float colour within 1e-5, disparity (a weighted sum of depths up to ~220)
    within 1e-6 relative + 1e-3, bytes and mask within the truncations of the
    oracle's value -/+ 1e-5."""
    import math

    import torch

    say(f"[image] eval render {kind}, 96x96")
    dev = c1["dev"]
    img = 96
    f = img / 2.0 / math.tan(math.radians(20.0))
    cam, _ = pxl.camera_from_pose((2.2, -3.1, 1.7), (N1 / 2.0,) * 3, 1.0, f, f, img / 2.0,
                                  img / 2.0)
    if kind == "dense":
        grid = pxl.Grid(N1, N1, N1, S1, c1["rec"], step=0.5)
        grid_o = oracle.Grid(N1, N1, N1, S1, c1["planes"], 0.5)
    else:
        s = sparse1
        grid = device_sparse_grid(N1, S1, s["compact"], s["occ"], s["io"], s["cage"], s["c2d"],
                                  dev)
        grid_o = oracle.Grid(N1, N1, N1, S1, s["compact"], 0.5, occupancy=s["occ"],
                             index_offsets=s["io"])
    out = pxl.render_image(grid, cam, img, img, white_bg=True, device=dev)
    torch.cuda.synchronize()
    c2w = np.array(cam.c2w[:], np.float64)
    centre = np.array(cam.centre[:], np.float64)
    ref = oracle.render_image(grid_o, c2w, centre, f, f, img / 2.0, img / 2.0, img, img)
    rgb = out["rgb"].cpu().numpy()
    assert np.abs(rgb - ref["rgb"]).max() <= 1e-5
    disp = out["disparity"].cpu().numpy()
    assert np.all(np.abs(disp - ref["disparity"]) <= 1e-6 * np.abs(ref["disparity"]) + 1e-3)
    # Bytes are truncations of 255 * value (+ 0.5 for colour): a value within
    # the 1e-5 pixel bar of an integer step may land on either side, so each
    # byte must lie between the truncations of the oracle value -/+ 1e-5.  The
    # accumulated weight A is 1 - (white - black background float colour).
    black = oracle.render_image(grid_o, c2w, centre, f, f, img / 2.0, img / 2.0, img, img,
                                white_bg=False)
    acc = 1.0 - (ref["rgb"][0] - black["rgb"][0])
    tol = 1e-5

    def byte_range(v, half):
        lo = np.clip(np.floor((v - tol) * 255.0 + half), 0, 255)
        hi = np.clip(np.floor((v + tol) * 255.0 + half), 0, 255)
        return lo, hi

    lo, hi = byte_range(acc, 0.0)
    m = out["mask"].cpu().numpy().astype(np.float64)
    assert np.all((m >= lo) & (m <= hi)), int(((m < lo) | (m > hi)).sum())
    lo, hi = byte_range(ref["rgb"], 0.5)  # over white, as the bytes always are
    b8 = out["rgb8"].cpu().numpy().astype(np.float64)
    assert np.all((b8 >= lo) & (b8 <= hi)), int(((b8 < lo) | (b8 > hi)).sum())
    lo, hi = byte_range(acc, 0.0)  # the derived weight reproduces the oracle's own mask
    assert np.all((ref["mask"] >= lo) & (ref["mask"] <= hi))
    hit = ref["rays"][:, 0] <= ref["rays"][:, 1]
    say(f"    pixels hitting the grid {hit.mean():.3f}, max colour error "
        f"{np.abs(rgb - ref['rgb']).max():.2e}")
    assert hit.any() and (~hit).any()
