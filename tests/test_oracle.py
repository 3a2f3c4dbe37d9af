"""Pins of the CPU oracle (no GPU).

Each oracle function is checked against something that does not restate its
own formula: the reference's pose fixtures (Poses.zip), its testRender
inputs, closed forms, invariants, brute force from the textual rule and
central finite differences.
"""
import math
import os
import zipfile
from fractions import Fraction

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def fma(a, b, c):
    """Correctly rounded a*b + c (math.fma needs Python 3.13)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def load_poses(scene="lego", split="train"):
    """cameraInfo.txt + centerInfo.txt of the committed pose fixture."""
    path = os.path.join(GOLDEN, "poses_lego.npz")
    d = np.load(path)
    return d[f"{split}_K"], d[f"{split}_t"], d[f"{split}_w2c"], d[f"{split}_centre"]


# ------------------------------------------------------------------ camera


@pytest.mark.parametrize("split", ["train", "val", "test"])
def test_camera_matches_pose_file(orc, split):
    """The look-at rotation and the -M^-1 p4 centre reproduce the w2c
    matrices and centres stored in Poses.zip (lego), to the file's 6
    decimals."""
    K, T, W2C, C = load_poses(split=split)
    for i in range(len(T)):
        c2w, w2c, ctr = orc.camera_from_pose(T[i], [0, 0, 0], 4.031129, K[0], K[4], K[2], K[5],
                                             need_rescale=False, z_reversed=False)
        assert np.abs(w2c[:3, :3] - W2C[i][:3, :3]).max() < 2e-6
        assert np.abs(w2c[:3, 3] - W2C[i][:3, 3]).max() < 2e-6
        assert np.abs(ctr[:3] - C[i]).max() < 1e-8
        assert np.abs(c2w @ w2c - np.eye(4)).max() < 1e-9


def test_camera_rescale_and_flip(orc):
    """Rescaled centre = t * oc / dist + oc; the z flip negates camera y, z."""
    t = np.array([1.0, -2.0, 3.0])
    oc = np.array([64.0, 64.0, 64.0])
    _, w2c0, c0 = orc.camera_from_pose(t, oc, 1.0, 100, 100, 50, 50, True, False)
    _, w2c1, c1 = orc.camera_from_pose(t, oc, 1.0, 100, 100, 50, 50, True, True)
    np.testing.assert_allclose(c0[:3], t * 64 + 64, rtol=0, atol=1e-9)
    np.testing.assert_allclose(c1[:3], c0[:3], rtol=0, atol=1e-9)
    np.testing.assert_allclose(w2c1[0, :3], w2c0[0, :3], atol=1e-12)
    np.testing.assert_allclose(w2c1[1:3, :3], -w2c0[1:3, :3], atol=1e-12)
    # camera z axis points from the target to the camera
    z = (c0[:3] - oc) / np.linalg.norm(c0[:3] - oc)
    np.testing.assert_allclose(w2c0[2, :3], z, atol=1e-12)


# -------------------------------------------------------------------- rays


def test_rays_optical_axis_and_slab(orc):
    """Pixel centre at the principal point looks down the optical axis; the
    slab interval of an axis-aligned ray through a box is closed-form."""
    N = 33
    oc = np.array([16.0, 16.0, 16.0])
    # camera on +x axis looking at the centre (level with it: XOY branch)
    cam_t = np.array([3.0, 0.0, 0.0])
    c2w, w2c, ctr = orc.camera_from_pose(cam_t, oc, 1.0, 10.0, 10.0, 1.0, 1.0, True, True)
    r = orc.rays_init(c2w, ctr, 10.0, 10.0, 1.0, 1.0, (0, N - 1, 0, N - 1, 0, N - 1), 2, 2)
    # the 4 pixel centres are (+-0.5 px) off-axis; their mean direction is the axis
    d = np.stack([r["dx"], r["dy"], r["dz"]], 1).astype(np.float64)
    axis = (oc - ctr[:3]) / np.linalg.norm(oc - ctr[:3])
    mean = d.mean(0)
    np.testing.assert_allclose(mean / np.linalg.norm(mean), axis, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # single pixel exactly on the axis: 1x1 image, principal point at its centre
    r1 = orc.rays_init(c2w, ctr, 10.0, 10.0, 0.5, 0.5, (0, N - 1, 0, N - 1, 0, N - 1), 1, 1)
    d1 = np.array([r1["dx"][0], r1["dy"][0], r1["dz"][0]], np.float64)
    np.testing.assert_allclose(d1, axis, atol=1e-6)
    cx = ctr[0]
    assert abs(r1["near_t"][0] - (cx - (N - 1))) < 1e-4
    assert abs(r1["far_t"][0] - (cx - 0)) < 1e-4
    assert r1["ox"][0] == np.float32(ctr[0])


def test_rays_ground_truth_white_background(orc):
    """White background: gt = rgb*mask/255 + 255 - mask, truncated
    (Rays.cu:118-124)."""
    rgb = np.array([[200, 10, 255, 0]], np.uint8).repeat(3, 0)
    mask = np.array([255, 128, 0, 77], np.uint8)
    c2w = np.eye(4)
    r = orc.rays_init(c2w, [0, 0, -10, 1], 1, 1, 2, 0.5, (0, 4, 0, 4, 0, 4), 4, 1,
                      img=(rgb[0], rgb[1], rgb[2], mask), white_bg=True)
    exp = [int(np.float32(int(v) * int(m)) / np.float32(255) + np.float32(255) - np.float32(m))
           for v, m in zip(rgb[0], mask)]
    assert list(r["gt_r"]) == exp
    assert exp == [200, 132, 255, 178]
    assert list(r["gt_mask"]) == list(mask)


# --------------------------------------------------- occupancy and index


def brute_occupancy(dens, thr):
    D, H, W = dens.shape
    occ = np.zeros((D - 1, H - 1, W - 1), np.uint8)
    cage = np.zeros(D * H * W, np.uint8)
    for z in range(D - 1):
        for y in range(H - 1):
            for x in range(W - 1):
                if (dens[z:z + 2, y:y + 2, x:x + 2] > thr).any():
                    occ[z, y, x] = 1
                    cage[z * W * H + y * W + x] = 1
                    if z == D - 2:
                        cage[(D - 1) * W * H + y * W + x] = 1
                    if y == H - 2:
                        cage[z * W * H + (H - 1) * W + x] = 1
                    if x == W - 2:  # the reference's index for this face
                        cage[z * W * H + H * W + x + 1] = 1
    return occ.reshape(-1), cage


@pytest.mark.parametrize("shape", [(5, 6, 7), (9, 9, 9), (4, 11, 3)])
def test_occupancy_brute_force(orc, shape):
    D, H, W = shape
    rng = np.random.default_rng(sum(shape))
    dens = rng.uniform(0, 10, shape).astype(np.float32)
    occ, cage, n = orc.build_occupancy(dens, W, H, D, 8.5)
    eo, ec = brute_occupancy(dens, np.float32(8.5))
    assert n == int(eo.sum())
    np.testing.assert_array_equal(occ, eo)
    np.testing.assert_array_equal(cage, ec)


def test_index_is_prefix_count(orc):
    rng = np.random.default_rng(5)
    cage = (rng.uniform(size=7 * 8 * 9) < 0.3).astype(np.uint8)
    io, n = orc.build_index(cage, 9, 8, 7)
    assert n == int(cage.sum())
    pre = np.cumsum(cage) - cage
    np.testing.assert_array_equal(io, np.where(cage != 0, pre, 0).astype(np.uint32))
    assert io[np.flatnonzero(cage)[0]] == 0  # the first caged vertex aliases "absent"


# ------------------------------------------------------------------ render


def line_grid(orc, sig_raw, col_raw, S=1, white=False):
    """Grid of W = len+2 vertices along x (H = D = 2) whose vertices at
    y = z = 0 carry the given raw values, and one ray along +x through
    those vertices with step 1: every sample sits exactly on a vertex."""
    n = len(sig_raw)
    W, H, D = n + 2, 2, 2
    P = 1 + 3 * S
    planes = np.zeros((P, W * H * D), np.float32)
    for k in range(n):
        v = k  # (x=k, y=0, z=0)
        planes[0, v] = sig_raw[k]
        for ch in range(3):
            planes[1 + ch * S, v] = col_raw[ch][k]
    g = orc.Grid(W, H, D, S, planes, 1.0, bbox=(0, W - 1, 0, 1, 0, 1))
    rays = orc.Rays({"ox": [0.0], "oy": [0.0], "oz": [0.0], "dx": [1.0], "dy": [0.0],
                     "dz": [0.0], "near_t": [0.0], "far_t": [n - 0.5],
                     "gt_r": [0], "gt_g": [0], "gt_b": [0], "gt_mask": [0]}, white_bg=white)
    return g, rays


# testRender inputs (VolumeFieldRender.cu:113-156): activated opacities and
# colours of 10 samples, step 1, GT (0.4, 0.5, 0.6).
TR_SIGMA = np.array([0.35082109, 0.49677446, 0.10159139, 0.699579, 0.70196073, 0.28737161,
                     0.88559283, 0.04455409, 0.53510863, 0.52025912])
TR_COL = np.array([
    [0.25436352, 0.41225156, 0.62032918, 0.33643962, 0.48815315, 0.10610127, 0.85595129,
     0.8505483, 0.69670436, 0.23520701],
    [0.40833749, 0.97833265, 0.36634089, 0.91709214, 0.1882598, 0.36327668, 0.51981918,
     0.4919779, 0.04402864, 0.72275637],
    [0.63738149, 0.66299872, 0.44536173, 0.04668531, 0.76686737, 0.66925641, 0.01505942,
     0.68553697, 0.81408741, 0.58484179]])
TR_GT = np.array([0.4, 0.5, 0.6])


def _inv_activations():
    sig_raw = 13.0 + np.log(np.expm1(TR_SIGMA))
    col_raw = 10.0 * np.log(TR_COL / (1.0 - TR_COL))
    return sig_raw, col_raw


def test_render_testrender_inputs(orc):
    """Forward colour of the reference's testRender inputs.  The grid holds
    float32 raw values, so the comparison uses the activations of those
    float32 values; the sample weights w_i = T_i (1 - e^{-sigma_i}) sum by
    telescoping to 1 - e^{-sum sigma}."""
    sig_raw, col_raw = _inv_activations()
    g, rays = line_grid(orc, sig_raw, col_raw)
    out = orc.render_fwd(g, rays, [0])
    s32 = g.planes[0, :10].astype(np.float64)
    sig = np.log1p(np.exp(s32 - 13.0))
    col = 1.0 / (1.0 + np.exp(-0.1 * np.stack([g.planes[1, :10], g.planes[2, :10],
                                               g.planes[3, :10]]).astype(np.float64)))
    T = np.exp(-np.concatenate([[0.0], np.cumsum(sig)[:-1]]))
    w = T * (1 - np.exp(-sig))
    np.testing.assert_allclose(sig, TR_SIGMA, rtol=1e-5)
    np.testing.assert_allclose([out["r"][0], out["g"][0], out["b"][0]], col @ w, rtol=1e-12)
    assert abs(out["w"][0] - (1 - math.exp(-sig.sum()))) < 1e-12
    assert abs(out["density"][0] - sig.sum()) < 1e-12


def test_render_constant_field_closed_form(orc):
    """Constant raw density s and colour on every vertex, white background:
    A = 1 - exp(-n sigma step), C = c A + (1 - A), n = number of samples."""
    N, S = 12, 4
    P = 1 + 3 * S
    planes = np.zeros((P, N ** 3), np.float32)
    planes[0] = 15.0
    planes[1 + 0 * S] = 3.0      # red, k = 0
    planes[1 + 1 * S + 2] = -2.0  # green, k = 2 (basis z)
    g = orc.Grid(N, N, N, S, planes, 0.5)
    rng = np.random.default_rng(0)
    n = 16
    o = rng.uniform(1, N - 2, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    near = np.zeros(n)
    far = rng.uniform(1, 6, n)
    rays = orc.Rays({"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1],
                     "dz": d[:, 2], "near_t": near, "far_t": far, "gt_r": np.zeros(n),
                     "gt_g": np.zeros(n), "gt_b": np.zeros(n), "gt_mask": np.zeros(n)})
    out = orc.render_fwd(g, rays, np.arange(n), trace=True)
    sig = math.log1p(math.exp(15.0 - 13.0))
    for i in range(n):
        far32 = float(np.float32(far[i]))
        ns = sum(1 for k in range(100) if fma(k, 0.5, 0.0) < far32)
        assert out["trace"][i, 0] == ns
        A = 1 - math.exp(-ns * sig * 0.5)
        dz = float(np.float32(d[i, 2]))
        cr = 1 / (1 + math.exp(-0.3))
        cg = 1 / (1 + math.exp(0.2 * dz))
        cb = 0.5
        assert abs(out["w"][i] - A) < 1e-12
        assert abs(out["r"][i] - (cr * A + 1 - A)) < 1e-12
        assert abs(out["g"][i] - (cg * A + 1 - A)) < 1e-12
        assert abs(out["b"][i] - (cb * A + 1 - A)) < 1e-12


def test_render_affine_field(orc):
    """Trilinear interpolation reproduces an affine raw density exactly, so
    the sum of activated samples equals the activation of the affine
    function at the sample positions."""
    N = 10
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    lin = (10.0 + 0.75 * x - 0.375 * y + 0.25 * z).reshape(-1)
    planes = np.zeros((4, N ** 3), np.float32)
    planes[0] = lin
    g = orc.Grid(N, N, N, 1, planes, 0.5)
    o = np.array([0.3, 8.7, 0.2])
    d = np.array([0.6, -0.5, 0.62])
    d = (d / np.linalg.norm(d)).astype(np.float32)
    rays = orc.Rays({"ox": [o[0]], "oy": [o[1]], "oz": [o[2]], "dx": [d[0]], "dy": [d[1]],
                     "dz": [d[2]], "near_t": [0.0], "far_t": [9.0], "gt_r": [0], "gt_g": [0],
                     "gt_b": [0], "gt_mask": [0]})
    out = orc.render_fwd(g, rays, [0])
    o32, d64 = o.astype(np.float32).astype(np.float64), d.astype(np.float64)
    sig = []
    for k in range(40):
        t = fma(k, 0.5, 0.0)
        if not t < 9.0:
            break
        p = [fma(t, d64[a], o32[a]) for a in range(3)]
        f = 10.0 + 0.75 * p[0] - 0.375 * p[1] + 0.25 * p[2]
        sig.append(math.log1p(math.exp(f - 13.0)))
    assert abs(out["density"][0] - sum(sig)) < 1e-9 * sum(sig)
    assert abs(out["w"][0] - (1 - math.exp(-0.5 * sum(sig)))) < 1e-12


def test_render_adaptive_step_rule(orc):
    """Adaptive rule: after an evaluated sample t += |step - w|.  In a
    near-empty field w ~ 1e-6, so away from a boundary the sample count
    matches the fixed rule; at far = 9.0 exactly the slightly shorter
    adaptive steps admit one more sample."""
    N = 8
    planes = np.zeros((4, N ** 3), np.float32)
    g_fixed = orc.Grid(N, N, N, 1, planes, 0.5, adaptive=False)
    g_adapt = orc.Grid(N, N, N, 1, planes, 0.5, adaptive=True)
    rays = orc.Rays({"ox": [0.2], "oy": [0.3], "oz": [0.1], "dx": [0.6], "dy": [0.64],
                     "dz": [0.48], "near_t": [0.0], "far_t": [9.0], "gt_r": [0], "gt_g": [0],
                     "gt_b": [0], "gt_mask": [0]})
    a = orc.render_fwd(g_fixed, rays, [0], trace=True)["trace"]
    b = orc.render_fwd(g_adapt, rays, [0], trace=True)["trace"]
    assert a[0, 0] == 18 and b[0, 0] == 19
    rays.d["far_t"][0] = 8.9
    a = orc.render_fwd(g_fixed, rays, [0], trace=True)["trace"]
    b = orc.render_fwd(g_adapt, rays, [0], trace=True)["trace"]
    assert a[0, 0] == b[0, 0] == 18
    rays.d["far_t"][0] = 9.0
    # a dense field makes w large, which shortens the adaptive steps
    planes[0] = 40.0
    g_dense = orc.Grid(N, N, N, 1, planes, 0.5, adaptive=True)
    c = orc.render_fwd(g_dense, rays, [0], trace=True)["trace"]
    assert c[0, 0] > 18


# ---------------------------------------------------------------- backward


def _fd_grid_gradient(orc, g, rays, si, idx_list, fn, h):
    """Central finite difference of fn(forward output) w.r.t. plane entries."""
    base = g.planes.copy()
    res = []
    for (p, v) in idx_list:
        vals = []
        for sgn in (1, -1):
            pl = base.copy()
            pl[p, v] = base[p, v] + sgn * h
            g2 = orc.Grid(g.W, g.H, g.D, g.S, pl, g.step, bbox=g.bbox, occupancy=g.occupancy,
                          index_offsets=g.index_offsets, adaptive=g.adaptive)
            h_eff = float(pl[p, v]) - float(base[p, v])
            vals.append((fn(orc.render_fwd(g2, rays, si)), h_eff))
        (fp, hp), (fm, hm) = vals
        res.append((fp - fm) / (hp - hm))
    return np.array(res)


def _bwd_scene(orc, S=4, seed=3):
    N = 9
    rng = np.random.default_rng(seed)
    P = 1 + 3 * S
    planes = np.zeros((P, N ** 3), np.float32)
    planes[0] = rng.uniform(8, 20, N ** 3)
    planes[1:] = rng.normal(0, 2.0, (3 * S, N ** 3))
    # float64-exact values so that +-h perturbations are exact
    planes = np.round(planes * 64) / 64
    g = orc.Grid(N, N, N, S, planes, 0.5)
    n = 6
    o = rng.uniform(0.5, N - 1.5, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = orc.Rays({"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1],
                     "dz": d[:, 2], "near_t": np.zeros(n), "far_t": np.full(n, 4.0),
                     "gt_r": rng.integers(0, 255, n), "gt_g": rng.integers(0, 255, n),
                     "gt_b": rng.integers(0, 255, n), "gt_mask": rng.integers(0, 255, n)})
    return g, rays


@pytest.mark.parametrize("huber", [10.0, 0.05])
def test_backward_finite_differences(orc, huber):
    """Gradient of the backward = 0.6667 * sum_ch r_used * dC_ch/dtheta
    - m * dA/dtheta, with r_used the (clipped) residual and m the mask and
    opacity-logit coefficient held fixed (RenderUtils.cu:2164-2193), checked
    by central differences of the forward."""
    g, rays = _bwd_scene(orc)
    n = rays.n
    for i in range(n):
        si = [i]
        fwd = orc.render_fwd(g, rays, si)
        losses, grad = orc.render_bwd(g, rays, si, fwd, mask_coef=1.0, huber_delta=huber)
        C = np.array([fwd["r"][i], fwd["g"][i], fwd["b"][i]])
        A = fwd["w"][i]
        gt = np.array([rays.d["gt_r"][i], rays.d["gt_g"][i], rays.d["gt_b"][i]]) / 255.0
        r = C - gt
        assert abs(losses[0] - np.abs(r).sum() * 255 / 3) < 1e-9
        d = huber
        r_used = np.where(r > d, d - d * d / 2, np.where(r < -d, d * d / 2 - d, r))
        gm = rays.d["gt_mask"][i]
        m = (gm / 255.0 - A) * 1.0 if (gm and A < 1) else 0.0
        wp = min(max(A, 1e-6), 1 - 1e-6)
        m += (math.log(wp) - math.log(1 - wp)) * 1e-3
        nz = np.argwhere(np.abs(grad) > 1e-9 * np.abs(grad).max())
        pick = nz[np.random.default_rng(i).choice(len(nz), min(12, len(nz)), replace=False)]
        idx = [tuple(x) for x in pick]

        def obj(out):
            Cc = np.array([out["r"][i], out["g"][i], out["b"][i]])
            return float((Cc * r_used).sum()), out["w"][i]

        fd_c = _fd_grid_gradient(orc, g, rays, si, idx, lambda o: obj(o)[0], 2 ** -14)
        fd_a = _fd_grid_gradient(orc, g, rays, si, idx, lambda o: obj(o)[1], 2 ** -14)
        for j, (p, v) in enumerate(idx):
            k = float(np.float32(0.6667)) if p > 0 else 0.6667
            exp = k * fd_c[j] - m * fd_a[j]
            assert abs(grad[p, v] - exp) <= 2e-6 * max(1e-3, abs(exp)) + 1e-9, (p, v, grad[p, v], exp)


def test_backward_testrender_gradients(orc):
    """testRender's own gradient expressions (VolumeFieldRender.cu:157-180)
    wrt the activated opacities and colours, chained through the
    activations and rescaled from its 0.667 to the kernel's 0.6667, plus the
    opacity-logit term -step sigma' T_final m of the kernel."""
    sig_raw, col_raw = _inv_activations()
    g, rays = line_grid(orc, sig_raw, col_raw)
    fwd = orc.render_fwd(g, rays, [0])
    C = np.array([fwd["r"][0], fwd["g"][0], fwd["b"][0]])
    # GT (0.4, 0.5, 0.6) as the nearest bytes, since rays carry byte targets
    gt_u8 = np.round(TR_GT * 255).astype(np.uint8)
    rays.d["gt_r"][0], rays.d["gt_g"][0], rays.d["gt_b"][0] = gt_u8
    gt = gt_u8 / 255.0
    losses, grad = orc.render_bwd(g, rays, [0], fwd, mask_coef=1.0, huber_delta=100.0)
    s32 = g.planes[0, :10].astype(np.float64)
    sig = np.log1p(np.exp(s32 - 13.0))
    col = 1.0 / (1.0 + np.exp(-0.1 * g.planes[1:4, :10].astype(np.float64)))
    # testRender's loops, with GT -> gt
    acc_op, Cacc = 0.0, np.zeros(3)
    dcol = np.zeros((3, 10))
    dop = np.zeros(10)
    for i in range(10):
        Ti = math.exp(-acc_op)
        w = Ti * (1 - math.exp(-sig[i]))
        acc_op += sig[i]
        for ch in range(3):
            Cacc[ch] += w * col[ch, i]
            dcol[ch, i] += 0.667 * (C[ch] - gt[ch]) * w
        Ti = math.exp(-acc_op)
        for ch in range(3):
            dop[i] += 0.667 * (C[ch] - gt[ch]) * Ti * col[ch, i]
            dop[i] -= 0.667 * (C[ch] - gt[ch]) * (C[ch] - Cacc[ch])
    A = fwd["w"][0]
    wp = min(max(A, 1e-6), 1 - 1e-6)
    m = (math.log(wp) - math.log(1 - wp)) * 1e-3
    Tfin = math.exp(-sig.sum())
    dsig = 1 / (1 + np.exp(13.0 - s32))
    exp_sig = dop * (0.6667 / 0.667) * dsig - 1.0 * dsig * Tfin * m
    np.testing.assert_allclose(grad[0, :10], exp_sig, rtol=1e-9, atol=1e-15)
    kc = float(np.float32(0.6667))
    for ch in range(3):
        exp_c = dcol[ch] * (kc / 0.667) * 0.1 * col[ch] * (1 - col[ch])
        np.testing.assert_allclose(grad[1 + ch, :10], exp_c, rtol=1e-9, atol=1e-15)


# --------------------------------------------------------------------- TV


def test_tv_disparity_closed_forms(orc):
    N = 7
    P = 4
    planes = np.zeros((P, N ** 3), np.float32)
    planes[:] = 3.25
    g = orc.Grid(N, N, N, 1, planes, 0.5)
    assert np.abs(orc.tv_disparity(g, 1.0, 1.0)).max() == 0.0
    # single spike at an interior vertex
    planes[:] = 0.0
    c = (3, 3, 3)
    v = c[2] * N * N + c[1] * N + c[0]
    planes[0, v] = 8.0
    planes[2, v] = -4.0
    g = orc.Grid(N, N, N, 1, planes, 0.5)
    lam_d, lam_s = 0.5, 0.25
    grad = orc.tv_disparity(g, lam_d, lam_s, optimize_step=0)
    assert grad[0, v] == pytest.approx(8 * lam_d * (8.0 - 1.0))
    assert grad[2, v] == pytest.approx(8 * lam_s * (-4.0 + 0.5))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if (dx, dy, dz) == (0, 0, 0):
                    continue
                u = (c[2] + dz) * N * N + (c[1] + dy) * N + (c[0] + dx)
                shared = (2 - abs(dx)) * (2 - abs(dy)) * (2 - abs(dz))
                assert grad[0, u] == pytest.approx(-lam_d * 1.0 * shared)
    # SH terms switch off above hierarchy level 1 (RenderUtils.cu:2493)
    g2 = orc.tv_disparity(g, lam_d, lam_s, optimize_step=2)
    assert np.abs(g2[1:]).max() == 0.0 and g2[0, v] == grad[0, v]


def test_tv_disparity_affine_interior_zero(orc):
    """For an affine field, every interior vertex receives zero: each cell
    mean is the value at the cell centre, and the 8 cells around a vertex
    are symmetric about it."""
    N = 8
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    planes = np.zeros((4, N ** 3), np.float32)
    planes[0] = (0.5 * x - 0.25 * y + 0.125 * z).reshape(-1)
    g = orc.Grid(N, N, N, 1, planes, 0.5)
    grad = orc.tv_disparity(g, 1.0, 0.0).reshape(4, N, N, N)
    assert np.abs(grad[0, 1:-1, 1:-1, 1:-1]).max() < 1e-12


def test_tv_neighbour6_closed_form(orc):
    N = 7
    planes = np.zeros((4, N ** 3), np.float32)
    v = 3 * N * N + 3 * N + 3
    planes[0, v] = 2.0
    g = orc.Grid(N, N, N, 1, planes, 0.5)
    grad = orc.tv_neighbour6(g, 1.0, 1.0)
    s = math.sqrt(6 * 4.0) / 6
    assert grad[0, v] == pytest.approx(1.0 / (s + 1e-6) * 12.0)
    u = v + 1
    assert grad[0, u] == pytest.approx(1.0 / (2.0 / 6 + 1e-6) * -2.0)
    assert np.count_nonzero(grad) == 7


# ------------------------------------------------------------------- Adam


def test_adam_first_step_closed_form(orc):
    rng = np.random.default_rng(1)
    P, n = 5, 64
    params = rng.normal(size=(P, n)).astype(np.float32)
    params[0] = np.abs(params[0])
    grad = rng.normal(size=(P, n)).astype(np.float32)
    m = np.zeros_like(params)
    v = np.zeros_like(params)
    p0, g0 = params.copy(), grad.copy()
    orc.adam(params, grad, m, v, 0.1, 0.3, 1)
    # t = 1: m_hat = g, v_hat = g^2 -> theta -= lr * g / (|g| + eps)
    lr = np.array([0.1] + [0.3] * (P - 1))[:, None]
    exp = p0 - lr * g0 / (np.abs(g0) + 1e-6)
    exp[0] = np.maximum(exp[0], 0)
    np.testing.assert_allclose(params, exp, rtol=2e-6, atol=2e-7)
    assert np.all(grad == 0)
    # beta1 = 0.9f and beta2 = 0.999f are float constants (Optimizer.cuh:22-23)
    b1, b2 = float(np.float32(0.9)), float(np.float32(0.999))
    np.testing.assert_allclose(m, (1 - b1) * g0.astype(np.float64), rtol=1e-6)
    np.testing.assert_allclose(v, (1 - b2) * g0.astype(np.float64) ** 2, rtol=1e-6)


def test_adam_zero_gradient_decay(orc):
    params = np.full((2, 4), 5.0, np.float32)
    grad = np.zeros_like(params)
    m = np.full_like(params, 0.5)
    v = np.full_like(params, 0.25)
    orc.adam(params, grad, m, v, 0.1, 0.1, 3)
    np.testing.assert_allclose(m, 0.45, rtol=1e-6)
    np.testing.assert_allclose(v, 0.24975, rtol=1e-6)
    mh = 0.45 / (1 - 0.9 ** 3)
    vh = 0.24975 / (1 - 0.999 ** 3)
    np.testing.assert_allclose(params, 5.0 - 0.1 * mh / (math.sqrt(vh) + 1e-6), rtol=1e-6)


def test_mse(orc):
    assert orc.mse(np.array([3.0, 4.0])) == pytest.approx(12.5)
