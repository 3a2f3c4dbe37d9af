"""pxl_train_rays_noalloc: the logged train step with its sample log in the
library's per-GPU static array, through the C ABI, against
pxl_train_rays_sched with a caller workspace (same schedule, same kernels):
per-ray colours and losses identical, gradients within helpers.grad_close
(the order of float atomics differs between runs) with record 0 within
1e-4 relative.  Two calls in a row, the second smaller (the array keeps the
first call's contents), and a persistent schedule (its claim counters live
in the array); a batch beyond the array is rejected."""
import ctypes
import sys

import numpy as np
import pytest

import pxl
from helpers import (c1_planes, cameras, device_sparse_grid, from_records, grad_close_alias,
                     oracle_rays, sparse_from_dense)

pytestmark = pytest.mark.gpu

N, S, IMG = 64, 9, 64


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(grid, rays, si, start, length, cap, sched, noalloc, dev):
    import torch

    optim = pxl.Optim(grid)
    losses = torch.zeros(length, dtype=torch.float64, device=dev)
    g, o, r = grid.c(), optim.c(), rays.c()
    sc = pxl.ScheduleC(*sched)
    if noalloc:
        rc = pxl.lib().pxl_train_rays_noalloc(ctypes.byref(g), ctypes.byref(o), ctypes.byref(r),
                                              ctypes.c_void_p(losses.data_ptr()),
                                              ctypes.c_void_p(si.data_ptr()), start, length, 1.0,
                                              0.1, cap, ctypes.byref(sc), None)
    else:
        ws = pxl.train_workspace(length, cap, dev)
        rc = pxl.lib().pxl_train_rays_sched(ctypes.byref(g), ctypes.byref(o), ctypes.byref(r),
                                            ctypes.c_void_p(losses.data_ptr()),
                                            ctypes.c_void_p(si.data_ptr()), start, length, 1.0,
                                            0.1, ctypes.c_void_p(ws.data_ptr()), ws.numel(), cap,
                                            ctypes.byref(sc), None)
    assert rc == pxl.PXL_OK, pxl.lib().pxl_last_error()
    torch.cuda.synchronize()
    ids = si[start:start + length].cpu().numpy()
    cols = np.stack([getattr(rays, k).cpu().numpy()[ids] for k in ("acc_r", "acc_g", "acc_b",
                                                                    "acc_w")])
    return from_records(optim.grad, grid.P), losses.cpu().numpy(), cols


def test_noalloc_matches_workspace_entry(dev):
    import torch

    say("[noalloc] 64^3 caged grid, 1024 then 300 rays")
    planes = c1_planes(N, S, n_blobs=8, seed=51)
    occ, cage, io, n, c2d, compact = sparse_from_dense(planes, N, 12.0, dev)
    grid = device_sparse_grid(N, S, compact, occ, io, cage, c2d, dev)
    cams = cameras(4, IMG, 4.0, 40.0, N)
    rays_np, _ = oracle_rays(cams, IMG, (0, N - 1, 0, N - 1, 0, N - 1), seed=52)
    rays = pxl.Rays.from_arrays(rays_np, device=dev)
    batch = np.random.default_rng(53).choice(len(rays_np["ox"]), 2048, replace=False)
    si = torch.from_numpy(batch.astype(np.int32)).to(dev)
    for sched in ((4, 0, 0, 0), (1, 1, 0, 0), (4, 0, 1, 4)):
        for start, length, cap in ((0, 1024, 256), (1500, 300, 96)):
            a = _run(grid, rays, si, start, length, cap, sched, True, dev)
            b = _run(grid, rays, si, start, length, cap, sched, False, dev)
            worst = grad_close_alias(a[0], b[0], rel0=1e-4)
            say(f"    schedule {sched}, {length} rays x {cap} samples: gradient / tolerance "
                f"{worst:.3g}")
            assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
            assert worst <= 1.0
    # Beyond the array: rejected before any launch.
    g, o, r = grid.c(), pxl.Optim(grid).c(), rays.c()
    rc = pxl.lib().pxl_train_rays_noalloc(ctypes.byref(g), ctypes.byref(o), ctypes.byref(r),
                                          ctypes.c_void_p(16), ctypes.c_void_p(si.data_ptr()), 0,
                                          65536, 1.0, 0.1, 4096, None, None)
    assert rc == pxl.PXL_EINVAL and b"PXL_NOALLOC_SCRATCH_BYTES" in pxl.lib().pxl_last_error()
