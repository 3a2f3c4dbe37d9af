"""The fused TV + Adam pass (pxl_tv_adam_step) against the two separate
calls it replaces, pxl_add_variance_loss (type 1, addVarianceLossGPU,
RenderUtils.cu:2469-2547) then pxl_adam_step (updateRGBSceneGPU_ADAM,
Optimizer.cu:359-415), on the same inputs: stepped records, gradient, both
moments and the FP32 master equal bit for bit, over two consecutive steps
(the second from the first's output, as the trainer swaps its buffers).

Cases: dense grids with odd sizes (the z-marched interior, in chunks of 128
plus a tail on a tall grid, and the boundary gather in the same launch), a
grid too thin for the z-march, the
64-bit address forms (library option), caged grids (prepared table + the
record-0 alias term), S = 16 (64 lanes per record), BF16 with an FP32 master,
F16 without one, an SH band (sh_first 1, sh_active 4), optimize_step 2 (no
SH terms), per-vertex adaptive LR, optimiser records of 28 floats (no
padding streamed) and of the grid's stride, gradient records of another
stride than the moments (32 or 36 floats beside 28 or 32).  The separate calls are pinned
to the FP64 oracle by the TV and Adam parity tests.
"""
import sys

import numpy as np
import pytest

import pxl
from helpers import c1_planes, device_sparse_grid, sparse_from_dense

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _state(grid, optim, seed, dev):
    """Random gradient and moments (v >= 0) of the optimiser's records."""
    import torch

    gen = torch.Generator(device="cpu").manual_seed(seed)
    n, w = optim.m.shape
    optim.grad.copy_(torch.randn(n, optim.grad_stride, generator=gen).mul_(0.05))
    optim.m.copy_(torch.randn(n, w, generator=gen).mul_(0.01))
    optim.v.copy_(torch.rand(n, w, generator=gen).mul_(1e-3))


def _clone_optim(grid, optim, stride, grad_stride=None):
    o2 = pxl.Optim(grid, adaptive_lr=optim.adaptive_lr, sh_first=optim.sh_first, stride=stride,
                   grad_stride=grad_stride)
    w = min(o2.grad_stride, optim.grad_stride)
    o2.grad[:, :w].copy_(optim.grad[:, :w])
    o2.m.copy_(optim.m)
    o2.v.copy_(optim.v)
    if optim.master is not None:
        o2.master = optim.master.clone()
    return o2


def _same(name, a, b):
    import torch

    a = a.contiguous().view(torch.int32 if a.element_size() == 4 else torch.int16)
    b = b.contiguous().view(torch.int32 if b.element_size() == 4 else torch.int16)
    bad = int((a != b).sum())
    assert bad == 0, f"{name}: {bad} elements differ"


def _run(grid, optim, ostride, lam=(0.03, 0.002), opt_step=0, fus_gstride=None):
    """Two steps, separate vs fused; compares everything after each.  The
    fused pass may keep its gradient at another record stride
    (`fus_gstride`) than the separate calls."""
    import torch

    g_sep = grid
    o_sep = optim
    g_fus = pxl.Grid(grid.W, grid.H, grid.D, grid.S, grid.params.clone(),
                     occupancy=grid.occupancy, index_offsets=grid.index_offsets,
                     in_cage=grid.in_cage, compact_to_dense=grid.compact_to_dense,
                     brick_occupancy=grid.brick_occupancy, storage=grid.storage,
                     sh_active=grid.sh_active)
    o_fus = _clone_optim(g_fus, optim, ostride, fus_gstride)
    spare = torch.empty_like(grid.params)
    first = grid.params.clone()
    for t in (1, 2):
        pxl.add_variance_loss(g_sep, o_sep, lam[0], lam[1], pxl.TV_DISPARITY,
                              optimize_step=opt_step)
        pxl.adam_step(g_sep, o_sep, 0.3, 0.01, t)
        pxl.tv_adam_step(g_fus, o_fus, lam[0], lam[1], 0.3, 0.01, t, spare,
                         optimize_step=opt_step)
        torch.cuda.synchronize()
        g_fus.params, spare = spare, g_fus.params
        _same(f"records (t={t})", g_fus.params, g_sep.params)
        P = grid.P
        _same(f"grad (t={t})", o_fus.grad[:, :P], o_sep.grad[:, :P])
        _same(f"m (t={t})", o_fus.m[:, :P], o_sep.m[:, :P])
        _same(f"v (t={t})", o_fus.v[:, :P], o_sep.v[:, :P])
        if optim.master is not None:
            _same(f"master (t={t})", o_fus.master[:, :P], o_sep.master[:, :P])
    # The steps moved the records (the comparison is not between no-ops).
    assert not torch.equal(first, g_sep.params)


def _dense(dev, W, H, D, S, storage=pxl.STORE_F32, seed=3, alr=False, sh_active=0,
           master=True):
    import torch

    gen = torch.Generator(device="cpu").manual_seed(seed)
    P = 1 + 3 * S
    stride = pxl.record_stride(S, storage)
    rec32 = torch.zeros(W * H * D, stride)
    rec32[:, :P] = torch.randn(W * H * D, P, generator=gen)
    rec32[:, 0] = rec32[:, 0].abs() * 3.0
    rec32[:, P:] = torch.randn(W * H * D, stride - P, generator=gen)  # padding carried as is
    rec32 = rec32.to(dev)
    if storage == pxl.STORE_F32:
        g = pxl.Grid(W, H, D, S, rec32, sh_active=sh_active)
        o_master = None
    else:
        g = pxl.Grid(W, H, D, S, pxl.records_to_16bit(rec32, storage), storage=storage,
                     sh_active=sh_active)
        o_master = rec32[:, :pxl.record_stride(S)] if master else None
    a = (torch.rand(W * H * D, generator=gen) + 0.5).to(dev) if alr else None
    o = pxl.Optim(g, adaptive_lr=a, master_from=o_master)
    return g, o


DENSE = {
    # case: (W, H, D, S, storage, options)
    "f32_odd": (21, 19, 33, 9, pxl.STORE_F32, {}),
    "f32_odd_o28": (21, 19, 33, 9, pxl.STORE_F32, {"ostride": 28}),
    # Gradient records of 32 floats beside moments of 28: in the separate
    # calls (TV scatter + Adam) or in the fused pass.
    "f32_odd_o28_g32sep": (21, 19, 33, 9, pxl.STORE_F32, {"ostride": 28, "gsep": 32}),
    "f32_odd_o28_g32fus": (21, 19, 33, 9, pxl.STORE_F32, {"ostride": 28, "gfus": 32}),
    "bf16_g36fus": (21, 19, 33, 9, pxl.STORE_BF16, {"gfus": 36}),
    "f32_alr_step2": (21, 19, 33, 9, pxl.STORE_F32, {"alr": True, "opt_step": 2}),
    "f32_thin": (2, 7, 9, 9, pxl.STORE_F32, {}),
    "f32_tall": (5, 4, 300, 9, pxl.STORE_F32, {}),
    # The narrowest grid marched along z: every column on a face, its 3 x 3
    # block the whole xy slab; the last z chunk holds 2 vertices.
    "f32_3x3": (3, 3, 130, 9, pxl.STORE_F32, {}),
    "f32_addr64": (21, 19, 33, 9, pxl.STORE_F32, {"addr64": True}),
    "f32_band": (21, 19, 33, 9, pxl.STORE_F32, {"sh_active": 4, "sh_first": 1}),
    "bf16_master": (21, 19, 33, 9, pxl.STORE_BF16, {}),
    "f16_plain": (21, 19, 33, 9, pxl.STORE_F16, {"master": False}),
    "s16_f32": (13, 11, 29, 16, pxl.STORE_F32, {}),
    "s16_bf16": (13, 11, 29, 16, pxl.STORE_BF16, {}),
}


@pytest.mark.parametrize("case", list(DENSE))
def test_tv_adam_dense(dev, case):
    import torch

    W, H, D, S, storage, opt = DENSE[case]
    grid, optim = _dense(dev, W, H, D, S, storage=storage, alr=opt.get("alr", False),
                         sh_active=opt.get("sh_active", 0), master=opt.get("master", True))
    optim.sh_first = opt.get("sh_first", 0)
    if "ostride" in opt or "gsep" in opt:
        optim = pxl.Optim(grid, adaptive_lr=optim.adaptive_lr, sh_first=optim.sh_first,
                          stride=opt.get("ostride", optim.stride), grad_stride=opt.get("gsep"),
                          master_from=None if optim.master is None else optim.master)
    _state(grid, optim, 5, dev)
    say(f"[tv+adam] dense {W}x{H}x{D} S={S} {case}")
    if opt.get("addr64"):
        pxl.set_option(pxl.OPT_TV_ADDR64, 1)
    try:
        _run(grid, optim, optim.stride, opt_step=opt.get("opt_step", 0),
             fus_gstride=opt.get("gfus"))
    finally:
        pxl.set_option(pxl.OPT_TV_ADDR64, 0)
    torch.cuda.synchronize()


@pytest.mark.parametrize("case", ["f32", "f32_o28", "f32_o28_g32", "f32_addr64", "bf16_master",
                                  "s16_f32"])
def test_tv_adam_caged(dev, case):
    import torch

    N = 40
    S = 16 if case.startswith("s16") else 9
    planes = c1_planes(N, S, n_blobs=6, seed=71)
    occ, cage, io, n, c2d, compact = sparse_from_dense(planes, N, 12.0, dev)
    grid = device_sparse_grid(N, S, compact, occ, io, cage, c2d, dev)
    optim_master = None
    if case == "bf16_master":
        rec32 = grid.params
        grid = pxl.Grid(N, N, N, S, pxl.records_to_16bit(rec32, pxl.STORE_BF16),
                        occupancy=grid.occupancy, index_offsets=grid.index_offsets,
                        in_cage=grid.in_cage, compact_to_dense=grid.compact_to_dense,
                        brick_occupancy=grid.brick_occupancy, storage=pxl.STORE_BF16)
        optim_master = rec32
    ostride = 28 if case.startswith("f32_o28") else None
    optim = pxl.Optim(grid, master_from=optim_master, stride=ostride)
    _state(grid, optim, 9, dev)
    say(f"[tv+adam] caged {N}^3 S={S}, {grid.n_vox} records, {case}")
    if case == "f32_addr64":
        pxl.set_option(pxl.OPT_TV_ADDR64, 1)
    try:
        _run(grid, optim, optim.stride, fus_gstride=32 if case.endswith("_g32") else None)
    finally:
        pxl.set_option(pxl.OPT_TV_ADDR64, 0)
    torch.cuda.synchronize()


def test_tv_adam_rejects_overlap(dev):
    import torch

    grid, optim = _dense(dev, 8, 8, 8, 9)
    with pytest.raises(pxl.PxlError, match="overlaps"):
        pxl.tv_adam_step(grid, optim, 0.1, 0.1, 0.1, 0.1, 1, grid.params)
    torch.cuda.synchronize()
