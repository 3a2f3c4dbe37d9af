"""The benchmark's CPU baseline (oracle/refprec.c: the train step in the
reference's arithmetic types) against the FP64 oracle on the same inputs.

Scene: a dense 24^3 SH-2 grid of 6 blobs (raw density rescaled into
[10, 14]) and its caged copy (cells whose corners exceed 12), 4 views of
24 x 24 on a radius-3 sphere, random 8-bit targets and masks; 512 rays, 4
OpenMP threads (float atomics, TV parity passes).  Bars, about ten times the
measured distance of the float interpolation and activations from FP64
(3.5e-7, 5e-5, 2.9e-6): pixels and opacity within 2e-6, losses within 5e-4
(of 255), gradients within 3e-5 of each plane's largest magnitude; the
variance-loss gradient within 4e-5 of its largest (float sums of raw
values); Adam within 1e-6 relative.  The fixed marching rule and the
reference's adaptive one; the mask / logit term, the Huber clip and the SH
band sums all enter these comparisons.
"""
import numpy as np
import pytest

import oracle
from helpers import c1_planes, cameras, oracle_rays, sparse_from_dense

N, S, IMG = 24, 9, 24
P = 1 + 3 * S


@pytest.fixture(scope="module")
def scene():
    planes = c1_planes(N, S, n_blobs=6, seed=91)
    # Raw density rescaled into [10, 14] (sigma = softplus(raw - 13) from
    # 0.05 to 1.3): a semi-transparent volume, where every sample's density
    # gradient matters (the blob scene itself is opaque at the first sample
    # and its density gradients cancel to round-off).
    d = planes[0]
    planes[0] = 10.0 + 4.0 * (d - d.min()) / (d.max() - d.min())
    cams = cameras(4, IMG, 3.0, 40.0, N)
    rays_np, _ = oracle_rays(cams, IMG, (0, N - 1, 0, N - 1, 0, N - 1), seed=92)
    batch = np.random.default_rng(93).choice(len(rays_np["ox"]), 512, replace=False)
    occ, cage, io, n, c2d, compact = sparse_from_dense(planes, N, 12.0, None)
    return dict(planes=planes, rays=oracle.Rays(rays_np), batch=batch.astype(np.int32),
                occ=occ, io=io, compact=compact)


def _grid(sc, kind, adaptive=False):
    if kind == "dense":
        return oracle.Grid(N, N, N, S, sc["planes"], 0.5, adaptive=adaptive)
    return oracle.Grid(N, N, N, S, sc["compact"], 0.5, occupancy=sc["occ"],
                       index_offsets=sc["io"], adaptive=adaptive)


@pytest.mark.parametrize("kind", ["dense", "caged"])
@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("huber", [0.1, 10.0])
def test_render_train_vs_oracle(scene, kind, adaptive, huber):
    g = _grid(scene, kind, adaptive)
    b = scene["batch"]
    grad = np.zeros((P, g.n_vox), np.float32)
    fwd, losses = oracle.rp_render_train(g, scene["rays"], b, grad, 1.0, huber, threads=4)
    ref = oracle.render_fwd(g, scene["rays"], b)
    l_ref, g_ref = oracle.render_bwd(g, scene["rays"], b, ref, 1.0, huber)
    for i, k in enumerate(("r", "g", "b", "w")):
        err = np.abs(fwd[:, i] - ref[k][b]).max()
        assert err <= 2e-6, (k, err)
    assert np.abs(losses - l_ref).max() <= 5e-4
    scale = np.abs(g_ref).max(axis=1, keepdims=True)
    worst = (np.abs(grad - g_ref) / np.maximum(scale, 1e-30)).max()
    assert worst <= 3e-5, worst
    assert np.abs(g_ref).max() > 0.0


@pytest.mark.parametrize("kind", ["dense", "caged"])
@pytest.mark.parametrize("opt_step", [0, 2])
def test_tv_vs_oracle(scene, kind, opt_step):
    g = _grid(scene, kind)
    grad = np.zeros((P, g.n_vox), np.float32)
    oracle.rp_tv_disparity(g, 1e-2, 1e-3, opt_step, grad, threads=4)
    ref = np.zeros((P, g.n_vox), np.float64)
    oracle.tv_disparity(g, 1e-2, 1e-3, opt_step, grad=ref)
    # The reference sums 8 raw values near 12 in float (about 1e-6 absolute
    # in their mean): 4e-5 of the largest gradient.
    assert np.abs(grad - ref).max() <= 4e-5 * np.abs(ref).max()
    if opt_step > 1:
        assert np.abs(grad[1:]).max() == 0.0


def test_adam_vs_oracle(scene):
    rng = np.random.default_rng(5)
    p0 = rng.normal(size=(P, 4096)).astype(np.float32)
    g0 = rng.normal(size=(P, 4096)).astype(np.float32)
    m0 = rng.normal(size=(P, 4096)).astype(np.float32) * 0.1
    v0 = rng.random(size=(P, 4096)).astype(np.float32) * 0.01
    a = [x.copy() for x in (p0, g0, m0, v0)]
    b = [x.copy() for x in (p0, g0, m0, v0)]
    oracle.rp_adam(*a, 0.3, 1.0, 3, threads=4)
    oracle.adam(*b, 0.3, 1.0, 3)
    for x, y in zip(a, b):
        assert np.allclose(x, y, rtol=1e-6, atol=1e-7)
    assert np.abs(a[1]).max() == 0.0 and not np.array_equal(a[0], p0)
