"""Shared scene builders for the parity tests (numpy on the host, torch for
device memory)."""
import numpy as np

import oracle
import pxl
from pxl import scenes


def c1_planes(N=128, S=9, n_blobs=16, seed=1):
    """Dense grid of the small plain-method workload: density = sum of seeded
    anisotropic Gaussian blobs, SH ~ N(0, 0.3).  float32 planes [P, N^3]."""
    P = 1 + 3 * S
    planes = np.empty((P, N ** 3), np.float32)
    planes[0] = scenes.blob_density(N, n_blobs, seed).reshape(-1)
    planes[1:] = scenes.sh_planes(N, S, seed + 100)
    return planes


def cameras(n_views, img, radius, fov, N, upper_only=False):
    """Product cameras (pxl.camera_from_pose) and the oracle's for the same poses."""
    f, _, c, _ = scenes.view_intrinsics(img, fov)
    oc = [N / 2.0] * 3
    out = []
    for t in scenes.fibonacci_sphere(n_views, radius, upper_only):
        cam, _ = pxl.camera_from_pose(t, oc, 1.0, f, f, c, c)
        c2w, _, ctr = oracle.camera_from_pose(t, oc, 1.0, f, f, c, c)
        out.append((cam, c2w, ctr, f, c))
    return out


def oracle_rays(cams, img, bbox, seed=0, white_bg=True):
    """All rays of the views (oracle) with random seeded ground truth."""
    rng = np.random.default_rng(seed)
    parts = []
    for (_, c2w, ctr, f, c) in cams:
        ims = tuple(rng.integers(0, 256, img * img, dtype=np.uint8) for _ in range(4))
        r = oracle.rays_init(c2w, ctr, f, f, c, c, bbox, img, img, img=ims, white_bg=white_bg)
        r["_img"] = ims
        parts.append(r)
    keys = [k for k in parts[0] if not k.startswith("_")]
    allr = {k: np.concatenate([p[k] for p in parts]) for k in keys}
    return allr, [p["_img"] for p in parts]


def to_records(planes, dev, stride=None):
    import torch

    P, n = planes.shape
    S = (P - 1) // 3
    stride = stride or pxl.record_stride(S)
    rec = torch.zeros((n, stride), dtype=torch.float32, device=dev)
    rec[:, :P] = torch.from_numpy(np.ascontiguousarray(planes.T)).to(dev)
    return rec


def from_records(rec, P):
    return rec[:, :P].T.contiguous().cpu().numpy().astype(np.float64)


def sparse_from_dense(planes, N, tau, dev):
    """Occupancy, index and compact planes built with the oracle; returns
    (oracle Grid kwargs, device arrays)."""
    occ, cage, _ = oracle.build_occupancy(planes[0], N, N, N, tau)
    io, n = oracle.build_index(cage, N, N, N)
    c2d = np.flatnonzero(cage).astype(np.uint32)
    compact = np.ascontiguousarray(planes[:, c2d])
    return occ, cage, io, n, c2d, compact


def pack_bits(occ_bytes):
    """Cell bytes -> little-endian 32-bit words (bit c of word c // 32)."""
    n = len(occ_bytes)
    pad = (-n) % 32
    b = np.concatenate([occ_bytes.astype(np.uint8) != 0, np.zeros(pad, bool)])
    return np.packbits(b.reshape(-1, 8), axis=1, bitorder="little").reshape(-1).view(np.uint32)


def device_sparse_grid(N, S, compact, occ, io, cage, c2d, dev, step=0.5, march=0,
                       bricks=True):
    import torch

    rec = to_records(compact, dev)
    occ_d = torch.from_numpy(pack_bits(occ).view(np.int32)).to(dev)
    br = pxl.build_brick_occupancy(occ_d, N, N, N) if bricks else None
    return pxl.Grid(N, N, N, S, rec, step=step, march_mode=march, occupancy=occ_d,
                    index_offsets=torch.from_numpy(io.view(np.int32)).to(dev),
                    in_cage=torch.from_numpy(cage).to(dev),
                    compact_to_dense=torch.from_numpy(c2d.view(np.int32)).to(dev),
                    brick_occupancy=br)


def brick_bits(occ_bytes, W, H, D, dilated=False):
    """Reference brick map from cell bytes: any occupied cell per 8^3 brick;
    dilated: any occupied cell in the brick or one of its 26 neighbours."""
    cw, ch, cd = W - 1, H - 1, D - 1
    c = occ_bytes.reshape(cd, ch, cw) != 0
    nbx, nby, nbz = -(-cw // 8), -(-ch // 8), -(-cd // 8)
    pad = np.zeros((nbz * 8, nby * 8, nbx * 8), bool)
    pad[:cd, :ch, :cw] = c
    b = pad.reshape(nbz, 8, nby, 8, nbx, 8).any(axis=(1, 3, 5))
    if dilated:
        p = np.zeros((nbz + 2, nby + 2, nbx + 2), bool)
        p[1:-1, 1:-1, 1:-1] = b
        d = np.zeros_like(b)
        for z in range(3):
            for y in range(3):
                for x in range(3):
                    d |= p[z:z + nbz, y:y + nby, x:x + nbx]
        b = d
    return pack_bits(b.reshape(-1).astype(np.uint8))


def grad_close_alias(g, ref, rel=1e-4, floor=1e-3, rel0=1e-3):
    """grad_close over every record but 0, plus record 0 within rel0 (the
    floors from each plane's largest entry, record 0 included).  Two GPU
    runs of a large caged batch differ at record 0 by the order of its FP32
    atomics: it collects every out-of-cage corner of the batch (the caged
    index maps them there), a sum of ~10^5 terms that largely cancel.
    Returns the worst ratio."""
    worst = 0.0
    where = None
    for p in range(ref.shape[0]):
        scale = np.abs(ref[p]).max()
        if scale == 0:
            assert np.abs(g[p]).max() == 0, p
            continue
        tol = rel * (np.abs(ref[p]) + floor * scale)
        tol[0] = rel0 * abs(ref[p][0]) + rel * floor * scale
        r = np.abs(g[p] - ref[p]) / tol
        i = int(np.argmax(r))
        if r[i] > worst:
            worst = float(r[i])
            where = (p, i, float(g[p][i]), float(ref[p][i]), float(scale))
    if worst > 1.0:
        print(f"    worst at plane {where[0]} record {where[1]}: got {where[2]:.6g} "
              f"ref {where[3]:.6g} (plane max {where[4]:.3g})", flush=True)
    return worst


def grad_close(g, ref, rel=1e-4, floor=1e-3):
    """Gradient parity bar: per plane p, |g - ref| <= rel * (|ref| +
    floor * max|ref[p]|) for every element.  Returns the worst ratio."""
    worst = 0.0
    where = None
    for p in range(ref.shape[0]):
        scale = np.abs(ref[p]).max()
        if scale == 0:
            assert np.abs(g[p]).max() == 0, p
            continue
        tol = rel * (np.abs(ref[p]) + floor * scale)
        r = np.abs(g[p] - ref[p]) / tol
        i = int(np.argmax(r))
        if r[i] > worst:
            worst = float(r[i])
            where = (p, i, float(g[p][i]), float(ref[p][i]), float(scale))
    if worst > 1.0:
        print(f"    worst at plane {where[0]} record {where[1]}: got {where[2]:.6g} "
              f"ref {where[3]:.6g} (plane max {where[4]:.3g})", flush=True)
    return worst


def round_bf16(x):
    """float32 -> bfloat16 bits (round to nearest even), as uint16."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def widen_bf16(h):
    return (h.astype(np.uint32) << 16).view(np.float32)


def to_16bit_planes(planes, storage):
    """The values a 16-bit grid holds, widened back to float32 (exact)."""
    if storage == pxl.STORE_BF16:
        return widen_bf16(round_bf16(planes))
    return planes.astype(np.float16).astype(np.float32)


def grad_close_cond(g, ref, absum, rel=1e-4, floor=1e-3, kappa=1e-6, rel0=None):
    """grad_close, or within kappa of the entry's magnitude sum (absum, from
    oracle.render_bwd_cond: the sum of the magnitudes of every term that
    enters the entry).  An entry whose terms cancel is known to FP32
    arithmetic only to a few rounding units (2^-24 each) of that sum, which
    can exceed 1e-4 of the entry itself; kappa = 1e-6 allows ~16 of them.
    rel0 (optional): record 0 within rel0 relative instead.  Returns the
    worst ratio and the number of entries held by the magnitude term only."""
    worst = 0.0
    where = None
    n_cond = 0
    for p in range(ref.shape[0]):
        scale = np.abs(ref[p]).max()
        if scale == 0:
            assert np.abs(g[p]).max() == 0, p
            continue
        tol_rel = rel * (np.abs(ref[p]) + floor * scale)
        if rel0 is not None:
            tol_rel[0] = rel0 * abs(ref[p][0]) + rel * floor * scale
        err = np.abs(g[p] - ref[p])
        n_cond += int(((err > tol_rel) & (err <= kappa * absum[p])).sum())
        tol = np.maximum(tol_rel, kappa * absum[p])
        if rel0 is not None:
            tol[0] = tol_rel[0]
        r = err / tol
        i = int(np.argmax(r))
        if r[i] > worst:
            worst = float(r[i])
            where = (p, i, float(g[p][i]), float(ref[p][i]), float(scale), float(absum[p][i]))
    if worst > 1.0:
        print(f"    worst at plane {where[0]} record {where[1]}: got {where[2]:.6g} "
              f"ref {where[3]:.6g} (plane max {where[4]:.3g}, magnitude sum {where[5]:.3g})",
              flush=True)
    return worst, n_cond
