"""Learned proposal on the GPU: the MFMA network against its oracle, the
masked marcher and its training targets against the oracle's masked march,
and what a trained proposal gains on the 128^3 scene.

Network bars (BF16 operands, FP32 accumulation; the oracle rounds to
bfloat16 at the same points in FP64): features within one bfloat16 step;
logits and gradients within 1e-2 of their largest magnitude (an FP32 sum
that lands on the other side of a bfloat16 rounding boundary changes one
hidden value by one bfloat16 step); Adam 1e-6 relative.  Marcher bars as in
test_gpu_parity: pixels 1e-5, gradients 1e-4 relative, sample counts and
hashes bit-exact; targets bit-exact except rays where a sample's weight sits
within FP64 rounding of the threshold.
"""
import sys

import numpy as np
import pytest

import oracle
import pxl
from oracle import proposal as orp
from helpers import c1_planes, cameras, from_records, grad_close, oracle_rays, to_records

pytestmark = pytest.mark.gpu

N1, S1, IMG1 = 128, 9, 200


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scene(dev):
    import torch

    say("[prop] building 128^3 scene")
    planes = c1_planes(N1, S1)
    cams = cameras(20, IMG1, 4.0, 40.0, N1)
    bbox = (0, N1 - 1, 0, N1 - 1, 0, N1 - 1)
    rays_np, _ = oracle_rays(cams, IMG1, bbox, seed=9)
    nrays = len(rays_np["ox"])
    order = np.random.default_rng(3).permutation(nrays).astype(np.int32)
    return dict(planes=planes, rays_np=rays_np, order=order, dev=dev,
                rec=to_records(planes, dev), si=torch.from_numpy(order).to(dev))


def _bf16_bits_to_float(t):
    import torch

    return (t.to(torch.int32) << 16).view(torch.float32)


def _features_np(rays_np, ids):
    o = np.stack([rays_np[k][ids] for k in ("ox", "oy", "oz")], 1).astype(np.float32)
    d = np.stack([rays_np[k][ids] for k in ("dx", "dy", "dz")], 1).astype(np.float32)
    tn = rays_np["near_t"][ids].astype(np.float32)
    return orp.features(o, d, tn, (N1, N1, N1), f32=True)


def test_proposal_features(scene):
    import torch

    say("[prop] features, 1000 rays")
    dev = scene["dev"]
    n = 1000
    grid = pxl.Grid(N1, N1, N1, S1, scene["rec"], step=0.5)
    rays = pxl.Rays.from_arrays(scene["rays_np"], device=dev)
    ws = pxl.proposal_workspace(n, dev)
    pxl.proposal_features(grid, rays, scene["si"], ws, length=n)
    torch.cuda.synchronize()
    rows = (n + 63) // 64 * 64
    X = _bf16_bits_to_float(ws[:rows * 128 * 2].view(torch.int16)).view(rows, 128).cpu().numpy()
    ref = orp.bf16(_features_np(scene["rays_np"], scene["order"][:n]))
    assert np.all(X[n:] == 0.0)
    diff = np.abs(X[:n] - ref)
    step = np.abs(ref) * 2.0 ** -7 + 1e-30
    say(f"    features beyond one bf16 step: {(diff > step).sum()}; differing {(diff > 0).mean():.2e}")
    assert np.all(diff <= step)


def _read_X(ws, n):
    import torch

    rows = (n + 63) // 64 * 64
    return _bf16_bits_to_float(ws[:rows * 128 * 2].view(torch.int16)).view(rows, 128)[:n] \
        .cpu().numpy().astype(np.float64)


def test_proposal_forward_and_gradient(scene):
    import torch

    say("[prop] forward + gradient vs oracle, 1000 rays")
    dev = scene["dev"]
    n = 1000
    grid = pxl.Grid(N1, N1, N1, S1, scene["rec"], step=0.5)
    rays = pxl.Rays.from_arrays(scene["rays_np"], device=dev)
    p = pxl.Proposal(dev, seed=5)
    ws = pxl.proposal_workspace(n, dev)
    pxl.proposal_features(grid, rays, scene["si"], ws, length=n)
    logits = torch.zeros((n, 64), dtype=torch.float32, device=dev)
    pxl.proposal_forward(p, n, logits, ws)
    rng = np.random.default_rng(11)
    tg = rng.integers(0, 2 ** 63, size=n, dtype=np.int64).astype(np.uint64)
    valid = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    t_dev = torch.from_numpy(tg.view(np.int64)).to(dev)
    v_dev = torch.from_numpy(valid).to(dev)
    loss = torch.zeros(1, dtype=torch.float64, device=dev)
    pxl.proposal_train_step(p, n, t_dev, ws, t=1, lr=0.0, valid=v_dev, n_valid=int(valid.sum()),
                            pos_weight=2.0, loss_out=loss)
    torch.cuda.synchronize()
    X = _read_X(ws, n)
    params = p.master.cpu().numpy().astype(np.float64)
    z_ref, _ = orp.forward(params, X)
    z = logits.cpu().numpy()
    zerr = np.abs(z - z_ref).max() / np.abs(z_ref).max()
    l_ref, g_ref = orp.loss_grad(params, X, tg, valid, pos_weight=2.0)
    g = p.grad.cpu().numpy().astype(np.float64)
    gerr = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    lerr = abs(float(loss.item()) - l_ref) / abs(l_ref)
    say(f"    logits {zerr:.2e}, gradient {gerr:.2e}, loss {lerr:.2e} (relative to max)")
    assert zerr <= 1e-2 and gerr <= 1e-2 and lerr <= 1e-3
    # The weight rows of the two zero features get no gradient.
    assert np.all(g[:orp.W1].reshape(128, 64)[126:] == 0.0)


def test_proposal_adam(scene):
    import torch

    say("[prop] Adam vs oracle")
    dev = scene["dev"]
    n = 256
    grid = pxl.Grid(N1, N1, N1, S1, scene["rec"], step=0.5)
    rays = pxl.Rays.from_arrays(scene["rays_np"], device=dev)
    p = pxl.Proposal(dev, seed=6)
    ws = pxl.proposal_workspace(n, dev)
    pxl.proposal_features(grid, rays, scene["si"], ws, length=n)
    tg = torch.from_numpy(np.random.default_rng(2).integers(0, 2 ** 62, n)).to(dev)
    master = p.master.cpu().numpy().copy()
    m = np.zeros_like(master)
    v = np.zeros_like(master)
    for t in (1, 2):
        pxl.proposal_train_step(p, n, tg, ws, t=t, lr=1e-2)
        torch.cuda.synchronize()
        orp.adam(master, p.grad.cpu().numpy(), m, v, 1e-2, t)
    got = p.master.cpu().numpy()
    np.testing.assert_allclose(got, master, rtol=1e-6, atol=1e-7)
    # The BF16 copy is the round-to-nearest-even of the GPU's own master.
    w16 = p.weights.cpu().numpy().view(np.uint16)
    assert np.array_equal(w16, (orp.bf16(got[:orp.WEIGHTS]).astype(np.float32)
                                .view(np.uint32) >> 16).astype(np.uint16))


def _random_masks(rng, n):
    m = rng.integers(0, 2 ** 63, size=n, dtype=np.int64).astype(np.uint64)
    m |= rng.integers(0, 2 ** 63, size=n, dtype=np.int64).astype(np.uint64)  # ~75 % bins
    m[:8] = np.uint64(2 ** 64 - 1)
    m[8:12] = 0
    return m


def test_masked_render_matches_oracle(scene):
    import torch

    say("[prop] masked forward vs oracle, 1024 rays")
    dev = scene["dev"]
    n = 1024
    grid = pxl.Grid(N1, N1, N1, S1, scene["rec"], step=0.5)
    rays = pxl.Rays.from_arrays(scene["rays_np"], device=dev)
    masks = _random_masks(np.random.default_rng(4), n)
    tr = torch.zeros(2 * n, dtype=torch.int32, device=dev)
    pxl.render_fwd_masked(grid, rays, scene["si"], torch.from_numpy(masks.view(np.int64)).to(dev),
                          length=n, trace=tr)
    torch.cuda.synchronize()
    grid_o = oracle.Grid(N1, N1, N1, S1, scene["planes"], 0.5)
    rays_o = oracle.Rays(scene["rays_np"])
    ids = scene["order"][:n]
    ref = oracle.render_fwd_masked(grid_o, rays_o, ids, masks, trace=True)
    got_tr = tr.view(-1, 2).cpu().numpy().view(np.uint32)
    assert np.array_equal(got_tr, ref["trace"])
    for k_g, k_o in (("acc_r", "r"), ("acc_g", "g"), ("acc_b", "b"), ("acc_w", "w")):
        err = np.abs(getattr(rays, k_g).cpu().numpy()[ids] - ref[k_o][ids]).max()
        assert err <= 1e-5, (k_g, err)


@pytest.mark.parametrize("cap", [512, 64])
@pytest.mark.parametrize("binned", [False, True])
def test_masked_train_and_targets(scene, cap, binned):
    import torch

    say(f"[prop] masked train step + targets vs oracle, 1024 rays, log {cap}, "
        f"{'brick-binned' if binned else 'per-ray'} gradient")
    dev = scene["dev"]
    n = 1024
    grid = pxl.Grid(N1, N1, N1, S1, scene["rec"], step=0.5)
    rays = pxl.Rays.from_arrays(scene["rays_np"], device=dev)
    optim = pxl.Optim(grid)
    masks = _random_masks(np.random.default_rng(5), n)
    losses = torch.zeros(n, dtype=torch.float64, device=dev)
    ws = pxl.train_workspace(n, cap, dev, grid=grid if binned else None)
    tg = torch.zeros(n, dtype=torch.int64, device=dev)
    tv = torch.zeros(n, dtype=torch.uint8, device=dev)
    wmin = 1e-4
    pxl.train_rays_prop(grid, optim, rays, losses, scene["si"], ws, cap,
                        bin_mask=torch.from_numpy(masks.view(np.int64)).to(dev), bin_targets=tg,
                        target_valid=tv, target_wmin=wmin, length=n, binned=binned)
    torch.cuda.synchronize()
    grid_o = oracle.Grid(N1, N1, N1, S1, scene["planes"], 0.5)
    rays_o = oracle.Rays(scene["rays_np"])
    ids = scene["order"][:n]
    fwd = oracle.render_fwd_masked(grid_o, rays_o, ids, masks, target_wmin=wmin)
    l_ref, g_ref = oracle.render_bwd_masked(grid_o, rays_o, ids, fwd, masks)
    assert np.abs(losses.cpu().numpy() - l_ref).max() <= 255 * 1e-5
    worst = grad_close(from_records(optim.grad, grid.P), g_ref, floor=1e-3)
    say(f"    worst gradient error / tolerance = {worst:.3g}")
    assert worst <= 1.0
    valid = tv.cpu().numpy().astype(bool)
    got = tg.cpu().numpy().view(np.uint64)
    ref_t = fwd["targets"]
    same = got[valid] == ref_t[valid]
    say(f"    rays with targets {valid.sum()}, differing {(~same).sum()}")
    assert (valid.sum() == n) == (cap == 512)
    assert (~same).sum() <= max(2, n // 200)
    # Targets never name a bin the mask forbids.
    assert np.all((got[valid] & ~masks[valid]) == 0)


def test_proposal_gain(scene):
    """Trained on the grid's own targets (online, unmasked batches), the
    proposal prunes most samples of held-out rays while the masked render
    stays within 0.1 dB of the plain one against the perturbed-grid targets."""
    import torch

    say("[prop] training the proposal on the 128^3 grid")
    dev = scene["dev"]
    grid = pxl.Grid(N1, N1, N1, S1, scene["rec"], step=0.5)
    rays = pxl.Rays.from_arrays(scene["rays_np"], device=dev)
    optim = pxl.Optim(grid)
    p = pxl.Proposal(dev, seed=1)
    B, steps, cap = 4096, 300, 512
    losses = torch.zeros(B, dtype=torch.float64, device=dev)
    ws = pxl.train_workspace(B, cap, dev)
    pws = pxl.proposal_workspace(B, dev)
    tg = torch.zeros(B, dtype=torch.int64, device=dev)
    tv = torch.zeros(B, dtype=torch.uint8, device=dev)
    si = scene["si"]
    held = si[-8192:]
    lossv = torch.zeros(1, dtype=torch.float64, device=dev)
    for t in range(1, steps + 1):
        s0 = ((t - 1) * B) % (si.numel() - 8192 - B)
        pxl.train_rays_prop(grid, optim, rays, losses, si, ws, cap, bin_targets=tg,
                            target_valid=tv, target_wmin=1e-5, start=s0, length=B)
        pxl.proposal_features(grid, rays, si, pws, start=s0, length=B)
        pxl.proposal_train_step(p, B, tg, pws, t=t, lr=3e-3, valid=tv, pos_weight=4.0,
                                loss_out=lossv)
        if t % 100 == 0:
            torch.cuda.synchronize()
            say(f"    step {t}: proposal BCE {lossv.item() / (B * 64):.4f}")
    # Held-out rays: plain and masked renders.
    n = held.numel()
    pws2 = pxl.proposal_workspace(n, dev)
    pxl.proposal_features(grid, rays, held, pws2)
    logits = torch.zeros((n, 64), dtype=torch.float32, device=dev)
    pxl.proposal_forward(p, n, logits, pws2)
    masks = torch.zeros(n, dtype=torch.int64, device=dev)
    pxl.proposal_masks(logits, n, masks, threshold=0.05, dilate=1)
    tr_p = torch.zeros(2 * n, dtype=torch.int32, device=dev)
    tr_m = torch.zeros(2 * n, dtype=torch.int32, device=dev)
    pxl.render_fwd_masked(grid, rays, held, None, trace=tr_p)
    plain = torch.stack([rays.acc_r[held.long()], rays.acc_g[held.long()],
                         rays.acc_b[held.long()]], 1).clone()
    pxl.render_fwd_masked(grid, rays, held, masks, trace=tr_m)
    masked = torch.stack([rays.acc_r[held.long()], rays.acc_g[held.long()],
                          rays.acc_b[held.long()]], 1).clone()
    torch.cuda.synchronize()
    # Targets: the perturbed copy of the grid (sigma + N(0, 1) >= 0, SH + N(0, 0.05)).
    rng = np.random.default_rng(2)
    pert = scene["planes"].copy()
    pert[0] = np.maximum(pert[0] + rng.normal(size=pert[0].shape).astype(np.float32), 0.0)
    pert[1:] += rng.normal(scale=0.05, size=pert[1:].shape).astype(np.float32)
    grid_t = pxl.Grid(N1, N1, N1, S1, to_records(pert, dev), step=0.5)
    pxl.render_fwd_masked(grid_t, rays, held, None)
    target = torch.stack([rays.acc_r[held.long()], rays.acc_g[held.long()],
                          rays.acc_b[held.long()]], 1).clone()
    torch.cuda.synchronize()

    def psnr(a, b):
        mse = float(((a - b) ** 2).mean())
        return 10.0 * np.log10(1.0 / mse)

    s_plain = tr_p.view(-1, 2)[:, 0].double().sum().item()
    s_masked = tr_m.view(-1, 2)[:, 0].double().sum().item()
    d_psnr = psnr(masked, target) - psnr(plain, target)
    say(f"    samples: plain {s_plain:.0f}, proposal {s_masked:.0f} "
        f"({100 * (1 - s_masked / s_plain):.1f} % pruned); PSNR vs targets plain "
        f"{psnr(plain, target):.3f} dB, proposal {psnr(masked, target):.3f} dB "
        f"(delta {d_psnr:+.4f}); proposal vs plain {psnr(masked, plain):.1f} dB")
    assert s_masked < 0.6 * s_plain
    assert abs(d_psnr) <= 0.1
