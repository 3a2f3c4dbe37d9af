"""CPU pins of the oracle's upsampling (upsampleScene / upsample_trilinear,
SparseVolumeRadianceField/SceneUtils.cu:250-485): closed forms on affine and
constant fields, the upper-face clamp, the threshold rule, the caged source
(unoccupied cells, corners outside the cage read record 0), band masks."""
import numpy as np

import oracle


def _affine(W, H, D, S, coef=(0.5, 1.25, -2.0, 3.5)):
    """Plane q holds c0 + (q+1) * (c1 x + c2 y + c3 z): distinct slopes per
    axis, so a swapped axis or a wrong weight shows."""
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    base = coef[1] * x + coef[2] * y + coef[3] * z
    P = 1 + 3 * S
    planes = np.stack([coef[0] + (q + 1) * base for q in range(P)]).reshape(P, -1)
    return planes.astype(np.float32)


def _expected_affine(W, H, D, S, x2, y2, z2, coef=(0.5, 1.25, -2.0, 3.5)):
    # the sample position, each coordinate capped at the last vertex (the
    # clamped corner of an upper face repeats the last vertex)
    px = np.minimum(x2 / 2.0, W - 1)
    py = np.minimum(y2 / 2.0, H - 1)
    pz = np.minimum(z2 / 2.0, D - 1)
    base = coef[1] * px + coef[2] * py + coef[3] * pz
    return np.stack([coef[0] + (q + 1) * base for q in range(1 + 3 * S)])


def test_upsample_dense_affine_and_clamp():
    W, H, D, S = 5, 6, 7, 4
    g = oracle.Grid(W, H, D, S, _affine(W, H, D, S), 0.5)
    W2, H2, D2 = 2 * W, 2 * H, 2 * D
    out = oracle.upsample(g, W2, H2, D2, S, threshold=-1e30)
    z2, y2, x2 = np.meshgrid(np.arange(D2), np.arange(H2), np.arange(W2), indexing="ij")
    exp = _expected_affine(W, H, D, S, x2.reshape(-1), y2.reshape(-1), z2.reshape(-1))
    # interior: trilinear reproduces the affine field (float rounding only)
    np.testing.assert_allclose(out, exp, rtol=2e-6, atol=2e-5)
    # even vertices copy their source vertex exactly
    even = ((x2 % 2 == 0) & (y2 % 2 == 0) & (z2 % 2 == 0)).reshape(-1)
    src_idx = ((z2 // 2) * H * W + (y2 // 2) * W + x2 // 2).reshape(-1)
    assert np.array_equal(out[:, even], g.planes[:, src_idx[even]])


def test_upsample_threshold_and_bands():
    W = H = D = 4
    S = 4
    rng = np.random.default_rng(5)
    planes = rng.normal(size=(1 + 3 * S, W * H * D)).astype(np.float32)
    planes[0] = rng.uniform(0, 10, W * H * D)
    g = oracle.Grid(W, H, D, S, planes, 0.5)
    full = oracle.upsample(g, 8, 8, 8, S, threshold=-1e30)
    cut = oracle.upsample(g, 8, 8, 8, S, threshold=5.0)
    # density below the threshold becomes 0, the rest is unchanged
    assert np.array_equal(cut[0], np.where(full[0] < 5.0, 0.0, full[0]).astype(np.float32))
    assert np.array_equal(cut[1:], full[1:])
    dens = oracle.upsample_density(g, 8, 8, 8, 5.0)
    assert np.array_equal(dens, cut[0])
    # low band: coefficient 0 of each channel only, into a 1-coefficient grid
    low = oracle.upsample(g, 8, 8, 8, 1, threshold=5.0, low_band=True)
    assert np.array_equal(low[0], cut[0])
    for ch in range(3):
        assert np.array_equal(low[1 + ch], cut[1 + ch * S])
    # low band into a full-width grid leaves the higher coefficients zero
    low4 = oracle.upsample(g, 8, 8, 8, S, threshold=5.0, low_band=True)
    for ch in range(3):
        assert np.array_equal(low4[1 + ch * S], cut[1 + ch * S])
        assert np.all(low4[2 + ch * S:1 + (ch + 1) * S] == 0.0)


def _caged_copy(planes, W, H, D, occ, cage):
    io, n = oracle.build_index(cage, W, H, D)
    keep = np.flatnonzero(cage)
    return oracle.Grid(W, H, D, planes.shape[0] // 3, planes[:, keep], 0.5, occupancy=occ,
                       index_offsets=io)


def test_upsample_caged_source():
    W, H, D, S = 5, 5, 5, 1
    rng = np.random.default_rng(6)
    planes = rng.uniform(1, 2, size=(4, W * H * D)).astype(np.float32)
    dense = oracle.Grid(W, H, D, S, planes, 0.5)
    ref = oracle.upsample(dense, 10, 10, 10, S, threshold=-1e30)
    ncell = (W - 1) * (H - 1) * (D - 1)
    # every vertex caged and every cell occupied: identical to the dense source
    full = _caged_copy(planes, W, H, D, np.ones(ncell, np.uint8), np.ones(W * H * D, np.uint8))
    assert np.array_equal(oracle.upsample(full, 10, 10, 10, S, -1e30), ref)
    # one unoccupied cell (1, 2, 3) on the upper z face: the samples inside
    # it, and those past the last vertex plane that clamp into it, are zero;
    # the rest unchanged
    occ = np.ones(ncell, np.uint8)
    cell = 3 * (W - 1) * (H - 1) + 2 * (W - 1) + 1
    occ[cell] = 0
    holed = _caged_copy(planes, W, H, D, occ, np.ones(W * H * D, np.uint8))
    out = oracle.upsample(holed, 10, 10, 10, S, -1e30)
    z2, y2, x2 = np.meshgrid(np.arange(10), np.arange(10), np.arange(10), indexing="ij")
    cx = np.minimum(x2 // 2, W - 2)
    cy = np.minimum(y2 // 2, H - 2)
    cz = np.minimum(z2 // 2, D - 2)
    inside = ((cx == 1) & (cy == 2) & (cz == 3)).reshape(-1)
    assert inside.sum() == 16  # z2 = 6, 7 inside; z2 = 8, 9 clamped
    assert np.all(out[:, inside] == 0.0)
    assert np.array_equal(out[:, ~inside], ref[:, ~inside])
    # vertex (2, 2, 2) outside the cage: its even destination vertex reads record 0
    cage = np.ones(W * H * D, np.uint8)
    v = 2 * H * W + 2 * W + 2
    cage[v] = 0
    alias = _caged_copy(planes, W, H, D, np.ones(ncell, np.uint8), cage)
    out = oracle.upsample(alias, 10, 10, 10, S, -1e30)
    d2 = 4 * 100 + 4 * 10 + 4
    assert np.array_equal(out[:, d2], planes[:, 0])


def test_upsample_caged_destination():
    """A caged destination holds exactly the dense result at its caged
    vertices, record by record in dense order."""
    W, H, D, S = 4, 5, 6, 4
    rng = np.random.default_rng(8)
    planes = rng.normal(size=(1 + 3 * S, W * H * D)).astype(np.float32)
    g = oracle.Grid(W, H, D, S, planes, 0.5)
    W2, H2, D2 = 2 * W, 2 * H, 2 * D
    ref = oracle.upsample(g, W2, H2, D2, S, threshold=0.3)
    cage = (rng.uniform(size=W2 * H2 * D2) < 0.3).astype(np.uint8)
    io, n = oracle.build_index(cage, W2, H2, D2)
    out = oracle.upsample(g, W2, H2, D2, S, threshold=0.3, in_cage=cage, index_offsets=io, n=n)
    assert np.array_equal(out, ref[:, np.flatnonzero(cage)])
