"""Seeded synthetic scenes for the train-step workloads.

Grids are in the reference's voxel convention: a vertex grid of N^3 points,
world cube [-1, 1]^3 <-> voxel [0, N-1] around the object centre N/2, camera
poses given in world units and rescaled with camera_distance = 1
(readFromMyBlender, SparseVolumeRadianceField/CameraUtils.cu:362-374).

All generators take a numpy Generator seed and return numpy arrays (CPU) or
torch tensors on the requested device.
"""
import math

import numpy as np


def blob_params(n_blobs, N, seed, amp=(5.0, 50.0), sigma=(4.0, 16.0), centre=(0.25, 0.75)):
    """Anisotropic axis-aligned Gaussian blobs: amplitude, centre[3], sigma[3]."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(amp[0], amp[1], n_blobs)
    c = rng.uniform(centre[0] * N, centre[1] * N, (n_blobs, 3))
    s = rng.uniform(sigma[0], sigma[1], (n_blobs, 3))
    return a, c, s


def blob_density(N, n_blobs, seed, device=None, **kw):
    """Raw density sum_b a_b exp(-1/2 sum_axis ((x - c)/s)^2) on the N^3
    vertex grid, float32, shape [N, N, N] in z-y-x order.  Separable per
    blob, so each blob costs one N^3 outer product."""
    a, c, s = blob_params(n_blobs, N, seed, **kw)
    if device is None:
        x = np.arange(N, dtype=np.float64)
        out = np.zeros((N, N, N), np.float32)
        for b in range(n_blobs):
            gx = np.exp(-0.5 * ((x - c[b, 0]) / s[b, 0]) ** 2)
            gy = np.exp(-0.5 * ((x - c[b, 1]) / s[b, 1]) ** 2)
            gz = np.exp(-0.5 * ((x - c[b, 2]) / s[b, 2]) ** 2)
            out += (a[b] * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]).astype(
                np.float32)
        return out
    import torch

    x = torch.arange(N, dtype=torch.float64, device=device)
    out = torch.zeros((N, N, N), dtype=torch.float32, device=device)
    for b in range(n_blobs):
        gx = torch.exp(-0.5 * ((x - c[b, 0]) / s[b, 0]) ** 2)
        gy = torch.exp(-0.5 * ((x - c[b, 1]) / s[b, 1]) ** 2)
        gz = torch.exp(-0.5 * ((x - c[b, 2]) / s[b, 2]) ** 2)
        out += (a[b] * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]).float()
    return out


def sh_planes(N, S, seed, std=0.3):
    """3S SH coefficient planes ~ N(0, std), float32 [3S, N^3]."""
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, std, (3 * S, N * N * N)).astype(np.float32)


def fibonacci_sphere(n, radius, upper_only=False):
    """n points on a sphere (or the upper hemisphere z > 0)."""
    pts = []
    golden = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(n):
        if upper_only:
            z = 1.0 - (i + 0.5) / n * 0.9  # z in (0.1, 1)
        else:
            z = 1.0 - 2.0 * (i + 0.5) / n
        r = math.sqrt(max(0.0, 1.0 - z * z))
        th = golden * i
        pts.append((radius * r * math.cos(th), radius * r * math.sin(th), radius * z))
    return pts


def view_intrinsics(img, fov_deg):
    """Focal length and principal point of a square image of side img."""
    f = (img / 2.0) / math.tan(math.radians(fov_deg) / 2.0)
    return f, f, img / 2.0, img / 2.0


def shell_density(N, seed, n_shells=3, thickness=1.0, n_blobs=32, radius=(0.12, 0.27),
                  device="cuda"):
    """Thin spherical shells (amplitude 40, radius U(radius) N, centres in the
    central third) plus small blobs: the uneven-ray workload.  At N = 512
    with the defaults, cells with a corner above 20 are 1.02 % of all."""
    import torch

    rng = np.random.default_rng(seed)
    dens = blob_density(N, n_blobs, seed + 1, device=device, sigma=(2.0, 6.0))
    z, y, x = torch.meshgrid(*(torch.arange(N, device=device, dtype=torch.float32),) * 3,
                             indexing="ij")
    for _ in range(n_shells):
        c = rng.uniform(0.35 * N, 0.65 * N, 3)
        rad = rng.uniform(radius[0], radius[1]) * N
        d = torch.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
        dens += 40.0 * (torch.abs(d - rad) < thickness / 2).float()
    return dens
