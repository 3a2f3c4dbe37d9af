"""Pins of the oracle parts added for the SH band strategy, the 6-neighbour
variance loss on caged grids, VarianceType 2 and the proposal's masked
backward (all on the CPU):

- masked backward: central differences of the masked forward, as the plain
  backward is pinned (test_oracle.py::test_backward_finite_differences);
- band limit of the colour (trilerpRGB, RenderUtils.cu:859-1000): the same
  as zeroing the coefficients from sh_active on; band range of the gradient
  (RenderUtils.cu:1117-1132): zero outside [sh_first, sh_active), equal to
  the full gradient of that grid inside;
- band-limited variance losses and Adam (RenderUtils.cu:2480, 2594;
  Optimizer.cu:393-412): planes outside the band untouched;
- 6-neighbour TV on caged grids: a caged grid with every vertex in the cage
  equals the dense grid; the reference's in-cage test at the record number
  (RenderUtils.cu:2669) by brute force of the textual rule on a 6^3 grid;
  an affine field gives zero at interior vertices;
- VarianceType 2 (RenderUtils.cu:2840-2845): a spike in a column its launch
  does not cover gets only the disparity term, 7 lambda h in closed form.
"""
import math

import numpy as np
import pytest

from test_oracle import _bwd_scene


def _grid_like(orc, g, planes=None, **kw):
    args = dict(bbox=g.bbox, occupancy=g.occupancy, index_offsets=g.index_offsets,
                adaptive=g.adaptive, sh_active=g.sh_active, sh_first=g.sh_first,
                in_cage=g.in_cage)
    args.update(kw)
    return orc.Grid(g.W, g.H, g.D, g.S, g.planes if planes is None else planes, g.step, **args)


def _fd(orc, g, fwd_fn, idx, fn, h=2 ** -14):
    base = g.planes.copy()
    out = []
    for (p, v) in idx:
        vals = []
        for sgn in (1, -1):
            pl = base.copy()
            pl[p, v] = base[p, v] + sgn * h
            vals.append((fn(fwd_fn(_grid_like(orc, g, pl))), float(pl[p, v]) - float(base[p, v])))
        (fp, hp), (fm, hm) = vals
        out.append((fp - fm) / (hp - hm))
    return np.array(out)


def _check_bwd_fd(orc, g, rays, i, grad, fwd, fwd_fn, huber=0.05):
    C = np.array([fwd["r"][i], fwd["g"][i], fwd["b"][i]])
    A = fwd["w"][i]
    gt = np.array([rays.d["gt_r"][i], rays.d["gt_g"][i], rays.d["gt_b"][i]]) / 255.0
    r = C - gt
    d = huber
    r_used = np.where(r > d, d - d * d / 2, np.where(r < -d, d * d / 2 - d, r))
    gm = rays.d["gt_mask"][i]
    m = (gm / 255.0 - A) if (gm and A < 1) else 0.0
    wp = min(max(A, 1e-6), 1 - 1e-6)
    m += (math.log(wp) - math.log(1 - wp)) * 1e-3
    nz = np.argwhere(np.abs(grad) > 1e-9 * np.abs(grad).max())
    pick = nz[np.random.default_rng(i).choice(len(nz), min(10, len(nz)), replace=False)]
    idx = [tuple(x) for x in pick]
    fd_c = _fd(orc, g, fwd_fn, idx, lambda o: float((np.array([o["r"][i], o["g"][i],
                                                               o["b"][i]]) * r_used).sum()))
    fd_a = _fd(orc, g, fwd_fn, idx, lambda o: o["w"][i])
    for j, (p, v) in enumerate(idx):
        k = float(np.float32(0.6667)) if p > 0 else 0.6667
        exp = k * fd_c[j] - m * fd_a[j]
        assert abs(grad[p, v] - exp) <= 2e-6 * max(1e-3, abs(exp)) + 1e-9, (p, v, grad[p, v], exp)
    return len(idx)


def test_masked_backward_finite_differences(orc):
    """orc_render_bwd_masked = 0.6667 sum_ch r_used dC_ch/dtheta - m dA/dtheta
    of the masked forward (the steps of cleared bins skipped in both)."""
    g, rays = _bwd_scene(orc)
    rng = np.random.default_rng(11)
    checked = 0
    for i in range(rays.n):
        mask = np.array([rng.integers(0, 2 ** 63, dtype=np.uint64) | np.uint64(1)], np.uint64)
        si = [i]
        fwd_fn = lambda gg, si=si, mask=mask: orc.render_fwd_masked(gg, rays, si, mask)
        fwd = fwd_fn(g)
        plain = orc.render_fwd(g, rays, si)
        _, grad = orc.render_bwd_masked(g, rays, si, fwd, bin_mask=mask, mask_coef=1.0,
                                        huber_delta=0.05)
        if fwd["w"][i] == plain["w"][i]:
            continue  # the mask removed nothing on this ray
        checked += _check_bwd_fd(orc, g, rays, i, grad, fwd, fwd_fn)
    assert checked >= 20


def test_band_colour_limit_equals_zeroed_coefficients(orc):
    g, rays = _bwd_scene(orc, S=9, seed=5)
    si = np.arange(rays.n)
    for act in (1, 4):
        lim = _grid_like(orc, g, sh_active=act)
        pl = g.planes.copy()
        for ch in range(3):
            pl[1 + ch * 9 + act:1 + (ch + 1) * 9] = 0.0
        zeroed = _grid_like(orc, g, pl)
        a, b = orc.render_fwd(lim, rays, si), orc.render_fwd(zeroed, rays, si)
        for k in ("r", "g", "b", "w"):
            np.testing.assert_allclose(a[k], b[k], rtol=1e-13, atol=1e-15)


def test_band_gradient_range(orc):
    """Band [1, 4) of S = 9: zero outside the band's planes (density always
    kept), equal to the full gradient of the band-limited grid inside; and
    a finite-difference check of a band coefficient."""
    g, rays = _bwd_scene(orc, S=9, seed=6)
    si = np.arange(rays.n)
    band = _grid_like(orc, g, sh_active=4, sh_first=1)
    full = _grid_like(orc, g, sh_active=4, sh_first=0)
    fwd = orc.render_fwd(band, rays, si)
    _, gb = orc.render_bwd(band, rays, si, fwd, 1.0, 0.05)
    _, gf = orc.render_bwd(full, rays, si, fwd, 1.0, 0.05)
    live = np.zeros(28, bool)
    live[0] = True
    for ch in range(3):
        live[1 + ch * 9 + 1:1 + ch * 9 + 4] = True
    assert np.abs(gb[~live]).max() == 0.0
    assert np.abs(gf[1::9]).max() > 0.0  # coefficient 0 has a gradient outside band mode
    np.testing.assert_array_equal(gb[live], gf[live])
    for i in range(2):
        fwd_fn = lambda gg, i=i: orc.render_fwd(gg, rays, [i])
        f1 = fwd_fn(band)
        _, g1 = orc.render_bwd(band, rays, [i], f1, 1.0, 0.05)
        _check_bwd_fd(orc, band, rays, i, g1, f1, fwd_fn)


def test_band_variance_losses_and_adam(orc):
    rng = np.random.default_rng(2)
    N, S = 6, 9
    P = 1 + 3 * S
    planes = rng.normal(0, 1, (P, N ** 3)).astype(np.float32)
    planes[0] = np.abs(planes[0])
    live = np.zeros(P, bool)
    live[0] = True
    for ch in range(3):
        live[1 + ch * 9 + 4:1 + ch * 9 + 9] = True
    full = orc.Grid(N, N, N, S, planes, 0.5)
    band = orc.Grid(N, N, N, S, planes, 0.5, sh_active=9, sh_first=4)
    for fn in (lambda gg: orc.tv_disparity(gg, 0.5, 0.25, 0),
               lambda gg: orc.tv_neighbour6(gg, 0.5, 0.25)):
        gb, gf = fn(band), fn(full)
        assert np.abs(gb[~live]).max() == 0.0
        np.testing.assert_array_equal(gb[live], gf[live])
    p0 = planes.copy()
    p, gr = planes.copy(), rng.normal(size=planes.shape).astype(np.float32)
    m, v = np.zeros_like(p), np.zeros_like(p)
    g0 = gr.copy()
    orc.adam(p, gr, m, v, 0.1, 0.3, 1, band=(4, 9))
    assert np.array_equal(p[~live], p0[~live]) and np.array_equal(gr[~live], g0[~live])
    assert np.all(m[~live] == 0) and np.all(gr[live] == 0)
    exp = p0[live] - np.where(np.arange(P)[live, None] == 0, 0.1, 0.3) * g0[live] / (
        np.abs(g0[live]) + 1e-6)
    exp[0] = np.maximum(exp[0], 0)
    np.testing.assert_allclose(p[live], exp, rtol=2e-6, atol=2e-7)


def _cage_grid(orc, planes_dense, cage, N, S):
    io, n = orc.build_index(cage, N, N, N)
    c2d = np.flatnonzero(cage)
    occ = np.ones((N - 1) ** 3, np.uint8)
    return orc.Grid(N, N, N, S, planes_dense[:, c2d], 0.5, occupancy=occ, index_offsets=io,
                    in_cage=cage), io, c2d


def test_tv_neighbour6_caged_all_in_equals_dense(orc):
    rng = np.random.default_rng(4)
    N, S = 6, 1
    planes = rng.normal(0, 1, (4, N ** 3)).astype(np.float32)
    cage = np.ones(N ** 3, np.uint8)
    gc, _, _ = _cage_grid(orc, planes, cage, N, S)
    gd = orc.Grid(N, N, N, S, planes, 0.5)
    np.testing.assert_array_equal(orc.tv_neighbour6(gc, 0.5, 0.25), orc.tv_neighbour6(gd, 0.5, 0.25))


def test_tv_neighbour6_caged_brute_force(orc):
    """totalVarianceLoss's neighboorNums == 6 branch on a caged 6^3 grid, by
    its text: for each thread column (x, y) inside the faces, depth 1..D-2:
    indexs = the 27 records around (indexOffsets), skip unless
    inCageInfo[indexs[13]] (an array over dense vertices read at the record
    number), a neighbour record 0 reads the vertex's own value."""
    rng = np.random.default_rng(8)
    N, S = 6, 1
    planes = rng.normal(0, 1, (4, N ** 3)).astype(np.float32)
    cage = (rng.random(N ** 3) < 0.7).astype(np.uint8)
    cage[0] = 1
    g, io, c2d = _cage_grid(orc, planes, cage, N, S)
    got = orc.tv_neighbour6(g, 0.5, 0.25)
    exp = np.zeros_like(got)
    comp = g.planes.astype(np.float64)
    lam = [0.5] + [0.25] * 3
    skipped = 0
    for y in range(1, N - 1):
        for x in range(1, N - 1):
            for z in range(1, N - 1):
                idx = {}
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            k = (dz + 1) * 9 + (dy + 1) * 3 + dx + 1
                            idx[k] = io[(z + dz) * N * N + (y + dy) * N + (x + dx)]
                if not cage[idx[13]]:
                    skipped += 1
                    continue
                for p in range(4):
                    u = comp[p, idx[13]]
                    vals = [comp[p, idx[k]] if idx[k] != 0 else u for k in (4, 10, 12, 14, 16, 22)]
                    sd = math.sqrt(sum((u - w) ** 2 for w in vals)) / 6.0
                    exp[p, idx[13]] += lam[p] / (sd + 1e-6) * sum(u - w for w in vals)
    assert skipped > 0  # the record-number test removes some vertices here
    np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)


def test_tv_neighbour6_affine_field_zero(orc):
    N = 7
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    planes = np.zeros((4, N ** 3), np.float32)
    planes[0] = (0.5 * x - 0.25 * y + 0.125 * z).reshape(-1)
    grad = orc.tv_neighbour6(orc.Grid(N, N, N, 1, planes, 0.5), 1.0, 1.0).reshape(4, N, N, N)
    assert np.abs(grad[0, 1:-1, 1:-1, 1:-1]).max() < 1e-9


@pytest.mark.parametrize("row,covered", [(14, True), (15, False)])
def test_tv_both_column_limit(orc, row, covered):
    """W = H = 17: the type-2 launch (one 16 x 16 block of 256 threads)
    covers the columns y W + x < 256, so row y = 15 is left out of the
    6-neighbour term.  A spike h at (5, row, 5) gets the disparity term
    7 lambda h (8 cells, each lambda (h - h / 8)) and, when covered, the
    6-neighbour term lambda / (sqrt(6 h^2) / 6 + 1e-6) 6 h."""
    N, h, lam = 17, 2.0, 0.5
    planes = np.zeros((4, N ** 3), np.float32)
    v = 5 * N * N + row * N + 5
    planes[0, v] = h
    g = orc.Grid(N, N, N, 1, planes, 0.5)
    both = orc.tv_both(g, lam, 0.0)
    t6 = lam / (math.sqrt(6 * h * h) / 6 + 1e-6) * 6 * h
    exp = 7 * lam * h + (t6 if covered else 0.0)
    assert both[0, v] == pytest.approx(exp, rel=1e-12)
    np.testing.assert_allclose(both, orc.tv_disparity(g, lam, 0.0, grad=orc.tv_neighbour6(
        g, lam, 0.0, col_limit=256)), rtol=0, atol=0)


def test_render_bwd_cond_bounds(orc):
    """The magnitude sum of render_bwd_cond: its gradient is render_bwd's;
    |grad| <= grad_abs everywhere (triangle inequality); on the planes of SH
    coefficient 0 (basis 1) a single ray's colour terms all have the sign of
    its clipped residual, so there grad_abs == |grad|."""
    g, rays = _bwd_scene(orc, S=4, seed=9)
    # Keep the samples inside the grid, where the trilinear weights are >= 0
    # (outside it the clamped base cell extrapolates with negative weights).
    d = rays.d
    for i in range(rays.n):
        o = np.array([d["ox"][i], d["oy"][i], d["oz"][i]], np.float64)
        v = np.array([d["dx"][i], d["dy"][i], d["dz"][i]], np.float64)
        with np.errstate(divide="ignore"):
            t_exit = np.where(v > 0, (g.W - 1 - o) / v, np.where(v < 0, -o / v, np.inf))
        d["far_t"][i] = np.float32(min(4.0, 0.999 * t_exit.min()))
    for i in range(rays.n):
        fwd = orc.render_fwd(g, rays, [i])
        _, grad = orc.render_bwd(g, rays, [i], fwd, 1.0, 0.05)
        _, grad2, gabs = orc.render_bwd_cond(g, rays, [i], fwd, 1.0, 0.05)
        assert np.array_equal(grad, grad2)
        assert np.all(np.abs(grad) <= gabs * (1 + 1e-12) + 1e-300)
        for ch in range(3):
            p = 1 + ch * 4
            np.testing.assert_allclose(gabs[p], np.abs(grad[p]), rtol=1e-12, atol=0)
