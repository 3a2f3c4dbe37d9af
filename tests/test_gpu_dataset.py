"""Parity of the training-ray dataset build (mask-area rays, mask bounds,
shuffle, permutation; _rays::initialSelf with MaskAreaOnly + Render.cu:
219-232) with the oracle's ray generator and numpy.

Bars: mask bounds and permutations bit-exact; box rays bit-exact against the
same pixels of the oracle's full-view rays."""
import ctypes
import sys

import numpy as np
import pytest

import oracle
import pxl
from helpers import cameras

pytestmark = pytest.mark.gpu

N, IMG = 64, 96
RAY_KEYS = ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t", "gt_r", "gt_g", "gt_b",
            "gt_mask")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mask(rng, box):
    """A mask whose nonzero pixels span exactly `box` (rows r0..r1, cols
    c0..c1): its four extreme pixels set, random pixels inside."""
    r0, r1, c0, c1 = box
    m = np.zeros((IMG, IMG), np.uint8)
    inner = rng.uniform(size=(r1 - r0 + 1, c1 - c0 + 1)) < 0.3
    m[r0:r1 + 1, c0:c1 + 1] = np.where(inner, rng.integers(1, 256, inner.shape), 0)
    m[r0, c0 + (c1 - c0) // 2] = 200
    m[r1, c0] = 7
    m[r0 + (r1 - r0) // 2, c0] = 1
    m[r1, c1] = 255
    return m.reshape(-1)


def _bounds_np(m):
    rows, cols = np.nonzero(m.reshape(IMG, IMG))
    if len(rows) == 0:
        return [IMG - 1, 0, IMG - 1, 0]
    return [rows.min(), rows.max(), cols.min(), cols.max()]


def test_mask_bounds(dev):
    import torch

    say("[dataset] mask bounds")
    rng = np.random.default_rng(4)
    cases = [_mask(rng, (10, 70, 5, 90)), _mask(rng, (0, IMG - 1, 0, IMG - 1)),
             np.zeros(IMG * IMG, np.uint8)]
    one = np.zeros(IMG * IMG, np.uint8)
    one[IMG * IMG - 1] = 3
    cases.append(one)
    for m in cases:
        got = pxl.mask_bounds(torch.from_numpy(m).to(dev), IMG, IMG)
        torch.cuda.synchronize()
        assert got.cpu().tolist() == [int(v) for v in _bounds_np(m)]


@pytest.fixture(scope="module")
def views(dev):
    """Four views with random images and masks of different boxes."""
    rng = np.random.default_rng(11)
    cams = cameras(4, IMG, 4.0, 40.0, N)
    bbox = (0, N - 1, 0, N - 1, 0, N - 1)
    boxes = [(10, 70, 5, 90), (0, IMG - 1, 0, IMG - 1), (30, 31, 40, 60), (5, 80, 44, 45)]
    out = []
    for (cam, c2w, ctr, f, c), box in zip(cams, boxes):
        ims = [rng.integers(0, 256, IMG * IMG, dtype=np.uint8) for _ in range(3)]
        ims.append(_mask(rng, box))
        full = oracle.rays_init(c2w, ctr, f, f, c, c, bbox, IMG, IMG, img=ims)
        out.append(dict(cam=cam, ims=ims, box=box, full=full))
    return dict(views=out, bbox=bbox)


def test_dataset_build_and_shuffle(dev, views):
    """Mask-area rays of the views in a given view order, concatenated, then
    shuffled with ShuffleRank and gathered with pxl_rays_permute."""
    import torch

    vs = views["views"]
    order = [2, 0, 3, 1]
    total = 0
    for v in vs:
        r0, r1, c0, c1 = v["box"]
        total += (r1 - r0 + 1) * (c1 - c0 + 1)
    say(f"[dataset] {len(vs)} views, {total} mask-area rays")
    rays = pxl.Rays(total, device=dev)
    exp = {k: [] for k in RAY_KEYS}
    start = 0
    for i in order:
        v = vs[i]
        ims = [torch.from_numpy(x).to(dev) for x in v["ims"]]
        bnd = pxl.mask_bounds(ims[3], IMG, IMG).cpu().tolist()
        assert bnd == list(v["box"])
        pxl.rays_init_box(v["cam"], views["bbox"], IMG, IMG, bnd, rays, start_index=start,
                          img=ims)
        r0, r1, c0, c1 = bnd
        pix = (np.arange(r0, r1 + 1)[:, None] * IMG + np.arange(c0, c1 + 1)[None, :]).reshape(-1)
        for k in RAY_KEYS:
            exp[k].append(v["full"][k][pix])
        start += len(pix)
    torch.cuda.synchronize()
    exp = {k: np.concatenate(x) for k, x in exp.items()}
    for k in RAY_KEYS:
        got = getattr(rays, k).cpu().numpy()
        assert np.array_equal(got.view(np.uint8), exp[k].view(np.uint8)), k

    libc = ctypes.CDLL(None)
    libc.srand(3)
    perm = pxl.shuffle_rank(np.zeros(total, np.int32), total)
    shuffled = pxl.Rays(total, device=dev)
    pxl.rays_permute(rays, shuffled, torch.from_numpy(perm).to(dev))
    torch.cuda.synchronize()
    for k in RAY_KEYS:
        got = getattr(shuffled, k).cpu().numpy()
        assert np.array_equal(got.view(np.uint8), exp[k][perm].view(np.uint8)), k


def test_rays_permute_out_of_range(dev, views):
    import torch

    say("[dataset] permutation with indices outside the set")
    full = views["views"][0]["full"]
    src = pxl.Rays.from_arrays(full, device=dev)
    n = src.n
    perm = np.array([5, -1, n, n - 1, 0], np.int32)
    dst = pxl.Rays(len(perm), device=dev)
    pxl.rays_permute(src, dst, torch.from_numpy(perm).to(dev))
    torch.cuda.synchronize()
    for k in RAY_KEYS:
        got = getattr(dst, k).cpu().numpy()
        assert got[0] == full[k][5] and got[3] == full[k][n - 1] and got[4] == full[k][0], k
        assert got[1] == 0 and got[2] == 0, k
