"""The oracle's optional host threading (used by the CPU baseline): per-ray
outputs and Adam identical to the single-thread run, gradients and TV equal
up to the order of FP64 additions."""
import numpy as np

from helpers import c1_planes, cameras, oracle_rays


def test_threads_match_single_thread(orc):
    N, S = 32, 9
    planes = c1_planes(N, S, n_blobs=4, seed=5)
    occ, cage, _ = orc.build_occupancy(planes[0], N, N, N, 8.0)
    io, n = orc.build_index(cage, N, N, N)
    compact = np.ascontiguousarray(planes[:, np.flatnonzero(cage)])
    g = orc.Grid(N, N, N, S, compact, 0.5, occupancy=occ, index_offsets=io)
    cams = cameras(2, 24, 4.0, 40.0, N)
    rays_np, _ = oracle_rays(cams, 24, (0, N - 1, 0, N - 1, 0, N - 1), seed=2)
    rays = orc.Rays(rays_np)
    si = np.arange(len(rays_np["ox"]), dtype=np.int32)
    rng = np.random.default_rng(3)
    adam_in = [rng.normal(size=compact.shape).astype(np.float32) for _ in range(4)]
    adam_in[3] = np.abs(adam_in[3])
    out = {}
    for th in (1, 4):
        orc.set_threads(th)
        try:
            f = orc.render_fwd(g, rays, si, trace=True)
            losses, grad = orc.render_bwd(g, rays, si, f, 1.0, 0.1)
            tv = orc.tv_disparity(g, 1e-2, 1e-3, 0)
            a = [x.copy() for x in adam_in]
            orc.adam(*a, 0.1, 0.2, 3)
            out[th] = (f, losses, grad, tv, a)
        finally:
            orc.set_threads(1)
    (f1, l1, g1, tv1, a1), (f4, l4, g4, tv4, a4) = out[1], out[4]
    for k in ("r", "g", "b", "w", "density", "trace"):
        assert np.array_equal(f1[k], f4[k]), k
    assert np.array_equal(l1, l4)
    for x1, x4 in zip(a1, a4):
        assert np.array_equal(x1, x4)
    for r1, r4 in ((g1, g4), (tv1, tv4)):
        assert np.abs(r1).max() > 0
        np.testing.assert_allclose(r4, r1, rtol=1e-9, atol=1e-12 * np.abs(r1).max())
