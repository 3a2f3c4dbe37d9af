"""CPU pins of the learned-proposal oracle (oracle/proposal.py) and of the
oracle's masked marching: finite differences, the bfloat16 rounding rule,
closed forms on a constant field."""
import math

import numpy as np
import pytest

from oracle import proposal as prop


def test_bf16_round_to_nearest_even():
    import torch

    # Halfway cases go to the even mantissa.
    assert prop.bf16(1.0 + 2.0 ** -8) == 1.0
    assert prop.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert prop.bf16(-(1.0 + 3 * 2.0 ** -8)) == -(1.0 + 2.0 ** -6)
    rng = np.random.default_rng(3)
    x = (rng.normal(size=10000) * 10.0 ** rng.uniform(-6, 6, 10000)).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy().astype(np.float64)
    assert np.array_equal(prop.bf16(x), ref)


def _random_case(rng, n=6):
    params = rng.normal(scale=0.2, size=prop.PARAMS)
    X = rng.normal(size=(n, prop.IN))
    X[:, 126:] = 0.0
    targets = rng.integers(0, 2 ** 63, size=n, dtype=np.int64).astype(np.uint64)
    valid = np.ones(n, np.uint8)
    valid[1] = 0
    return params, X, targets, valid


def test_proposal_gradient_finite_differences():
    """The analytic backward of the FP64 network (no bfloat16 rounding)
    against central differences of the mean weighted BCE."""
    rng = np.random.default_rng(7)
    params, X, targets, valid = _random_case(rng)
    n_valid = int(valid.sum())
    _, g = prop.loss_grad(params, X, targets, valid, pos_weight=2.0, bf16_round=False)

    def mean_loss(p):
        return prop.loss_grad(p, X, targets, valid, 2.0, bf16_round=False)[0] / (n_valid * 64)

    idx = np.concatenate([rng.choice(prop.W1, 20, replace=False),
                          prop.W1 + rng.choice(prop.W2, 10, replace=False),
                          prop.W1 + prop.W2 + rng.choice(prop.W3, 10, replace=False),
                          prop.WEIGHTS + rng.choice(192, 10, replace=False)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(params)
        e[i] = h
        fd = (mean_loss(params + e) - mean_loss(params - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_proposal_padding_rows_have_zero_gradient():
    rng = np.random.default_rng(8)
    params, X, targets, valid = _random_case(rng)
    _, g = prop.loss_grad(params, X, targets, valid)
    w1g = g[:prop.W1].reshape(prop.IN, prop.WIDTH)
    assert np.all(w1g[126:] == 0.0)


def test_proposal_adam_first_step():
    """t = 1: bias correction makes the step lr * g / (|g| + eps)."""
    rng = np.random.default_rng(9)
    master = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    m = np.zeros(100, np.float32)
    v = np.zeros(100, np.float32)
    before = master.astype(np.float64).copy()
    prop.adam(master, g, m, v, 0.01, 1)
    mg = np.float32(0.1) * g
    vg = np.float32(0.001) * g * g
    expect = before - 0.01 * (mg / 0.1) / (np.sqrt(vg.astype(np.float64) / 0.001) + 1e-8)
    np.testing.assert_allclose(master, expect.astype(np.float32), rtol=1e-6)
    np.testing.assert_allclose(master, before - 0.01 * np.sign(g), atol=1e-5)


def test_proposal_features_closed_form():
    N = 65
    o = np.array([[0.0, 32.0, 64.0]])
    d = np.array([[1.0, 0.0, 0.0]])
    f = prop.features(o, d, np.array([-3.0]), (N, N, N))
    np.testing.assert_allclose(f[0, :3], [-1.0, 0.0, 1.0])
    np.testing.assert_allclose(f[0, 63:66], [1.0, 0.0, 0.0])
    # frequency i = 2, component x (= -1): sin(4 pi * -1) = 0, cos = 1
    assert abs(f[0, 3 + 2 * 6 + 0]) < 1e-12 and abs(f[0, 3 + 2 * 6 + 3] - 1.0) < 1e-12
    # direction, i = 0, component x (= 1): sin(pi) ~ 0, cos(pi) = -1
    assert abs(f[0, 63 + 3 + 3] + 1.0) < 1e-12
    assert np.all(f[:, 126:] == 0.0)


def _constant_field(orc, n=12):
    N, S = 12, 4
    P = 1 + 3 * S
    planes = np.zeros((P, N ** 3), np.float32)
    planes[0] = 15.0
    planes[1] = 3.0
    g = orc.Grid(N, N, N, S, planes, 0.5)
    rng = np.random.default_rng(0)
    o = rng.uniform(1, N - 2, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    far = rng.uniform(2, 6, n)
    rays = orc.Rays({"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1],
                     "dz": d[:, 2], "near_t": np.zeros(n), "far_t": far, "gt_r": np.zeros(n),
                     "gt_g": np.zeros(n), "gt_b": np.zeros(n), "gt_mask": np.zeros(n)})
    return g, rays, far


def _bin(k, tf):
    q = math.floor(k * 0.5 * 64.0 / tf)
    return min(q, 63)


def test_masked_march_constant_field(orc):
    """All-ones mask = plain march; zero mask = nothing; one allowed bin =
    exactly the steps of that bin; targets with a zero threshold = the bins
    of the evaluated steps."""
    g, rays, far = _constant_field(orc)
    n = len(far)
    si = np.arange(n)
    plain = orc.render_fwd(g, rays, si, trace=True)
    ones = orc.render_fwd_masked(g, rays, si, np.full(n, np.uint64(2 ** 64 - 1)), trace=True)
    for k in ("r", "g", "b", "w", "density"):
        assert np.array_equal(plain[k], ones[k])
    assert np.array_equal(plain["trace"], ones["trace"])
    zero = orc.render_fwd_masked(g, rays, si, np.zeros(n, np.uint64), trace=True)
    assert np.all(zero["trace"][:, 0] == 0) and np.all(zero["w"] == 0.0)
    sig = math.log1p(math.exp(2.0))
    for j in (0, 5, 40):
        out = orc.render_fwd_masked(g, rays, si, np.full(n, np.uint64(1) << np.uint64(j)),
                                    target_wmin=0.0, trace=True)
        for i in range(n):
            tf = float(np.float32(far[i]))
            ks = [k for k in range(200) if k * 0.5 < tf and _bin(k, tf) == j]
            assert out["trace"][i, 0] == len(ks)
            assert abs(out["w"][i] - (1 - math.exp(-len(ks) * sig * 0.5))) < 1e-12
            assert out["targets"][i] == (np.uint64(1) << np.uint64(j) if ks else 0)
    full = orc.render_fwd_masked(g, rays, si, None, target_wmin=0.0)
    for i in range(n):
        tf = float(np.float32(far[i]))
        bins = {_bin(k, tf) for k in range(200) if k * 0.5 < tf}
        assert full["targets"][i] == sum(np.uint64(1) << np.uint64(b) for b in bins)


# ------------------------------------------------------------ eval image


def _eval_camera(orc, N, img):
    f = img / 2.0 / math.tan(math.radians(20.0))
    c = img / 2.0
    c2w, _, centre = orc.camera_from_pose((2.2, -3.1, 1.7), (N / 2.0,) * 3, 1.0, f, f, c, c)
    return c2w, centre, f, c


def test_eval_rays_match_training_rays(orc):
    """The eval ray of each pixel (getBound in FP64) is the training ray of
    the same pixel (the FP32 replica of the training ray builder) up to FP32
    rounding: same direction, near and far."""
    N, img = 40, 48
    planes = np.zeros((4, N ** 3), np.float32)
    g = orc.Grid(N, N, N, 1, planes, 0.5)
    c2w, centre, f, c = _eval_camera(orc, N, img)
    ev = orc.render_image(g, c2w, centre, f, f, c, c, img, img)
    tr = orc.rays_init(c2w.reshape(-1), centre, f, f, c, c, (0, N - 1, 0, N - 1, 0, N - 1),
                       img, img)
    d_tr = np.stack([tr["dx"], tr["dy"], tr["dz"]], 1).astype(np.float64)
    d_ev = ev["rays"][:, 2:5]
    # FP32 pixel coordinates, c2w product and centre subtraction in the
    # training builder: ~1e-6 on the unit direction (a swapped row / column,
    # sign or missing normalisation would be ~1e-1).
    assert np.abs(d_ev - d_tr).max() < 1e-5
    hit = ev["rays"][:, 0] <= ev["rays"][:, 1]
    assert hit.any() and (~hit).any()
    np.testing.assert_allclose(ev["rays"][hit, 0], tr["near_t"][hit], rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(ev["rays"][hit, 1], tr["far_t"][hit], rtol=1e-5, atol=1e-3)


def test_eval_image_output_conventions(orc):
    """Zero density: white bytes, float white only on a white background,
    mask 0; missed pixels white with zero disparity.  Opaque grid: mask
    254-255, disparity = weight-averaged depth inside [near, far]."""
    N, img = 40, 48
    c2w, centre, f, c = _eval_camera(orc, N, img)
    empty = orc.Grid(N, N, N, 1, np.full((4, N ** 3), -200.0, np.float32), 0.5)
    for wb in (True, False):
        out = orc.render_image(empty, c2w, centre, f, f, c, c, img, img, white_bg=wb)
        assert np.all(out["rgb8"] == 255) and np.all(out["mask"] == 0)
        np.testing.assert_allclose(out["rgb"], 1.0 if wb else 0.0, atol=1e-12)
    planes = np.zeros((4, N ** 3), np.float32)
    planes[0] = 60.0
    planes[1] = 10.0
    opaque = orc.Grid(N, N, N, 1, planes, 0.5)
    out = orc.render_image(opaque, c2w, centre, f, f, c, c, img, img)
    hit = out["rays"][:, 0] <= out["rays"][:, 1]
    assert np.all(out["mask"][hit] >= 254)
    assert np.all(out["disparity"][~hit] == 0.0)
    dmean = out["disparity"][hit] / np.maximum(1e-12, out["mask"][hit] / 255.0)
    assert np.all(dmean >= out["rays"][hit, 0] - 1e-6)
    assert np.all(dmean <= out["rays"][hit, 0] + 1.0)
    col = 1.0 / (1.0 + math.exp(-1.0))
    np.testing.assert_allclose(out["rgb"][0, hit], col, atol=1e-6)
    assert np.all(out["rgb8"][0, hit] == int(col * 255 + 0.5))
