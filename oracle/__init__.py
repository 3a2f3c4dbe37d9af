"""CPU oracle of the reference's train-step maths (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker and never as the thing
measured or shipped.  The product path (libpxl.so) does not use it.

Arrays follow the reference's layout: P = 1 + 3S parameter planes of n_vox
values, plane-major (SparseVolumeRadianceField/SceneUtils.cu:148-171).

Parity status of each function (how it is pinned; see tests/test_oracle.py):
  camera_from_pose  - Poses.zip lego centres and w2c rotations, closed forms
  rays_init         - closed forms (optical axis, slab distances), invariants
  build_occupancy   - brute force from the textual rule on tiny grids
  build_index       - brute force / prefix-sum identity
  render_fwd        - closed forms (constant field, slab), testRender inputs
  render_bwd        - central finite differences of a differentiable restatement
  render_bwd_cond   - the gradient equals render_bwd's; the magnitude sum
                      bounds it and equals it for terms of one sign
  render_bwd_masked - central finite differences of the masked forward
  tv_disparity      - closed forms (constant / affine field, single spike)
  tv_neighbour6     - closed forms (constant field, single spike, affine
                      field), brute force of the caged rule on a tiny grid
  tv_both           - the column limit of the type-2 launch by hand; sum of
                      the two terms
  adam              - closed form at t = 1 and zero-gradient decay
  render_fwd_masked - all-ones mask == plain march, one-bin masks on a constant
                      field (closed form)
  render_image      - eval rays == training rays; output conventions on empty
                      and opaque grids (closed forms)
  upsample_density,
  upsample          - even vertices copy the source, affine fields reproduced
                      away from the clamped upper faces, the clamp itself,
                      unoccupied source cells, the threshold rule, band masks

rp_* (librefprec.so, oracle/refprec.c) is the benchmark's CPU baseline: the
train step in the reference's arithmetic types, pinned against render_fwd /
render_bwd / tv_disparity / adam above (tests/test_refprec.py).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class OrcGrid(ctypes.Structure):
    _fields_ = [
        ("W", ctypes.c_int32), ("H", ctypes.c_int32), ("D", ctypes.c_int32),
        ("S", ctypes.c_int32), ("sparse", ctypes.c_int32), ("n_vox", ctypes.c_int64),
        ("planes", ctypes.c_void_p), ("occupancy", ctypes.c_void_p),
        ("index_offsets", ctypes.c_void_p), ("bbox", ctypes.c_int32 * 6),
        ("step", ctypes.c_float), ("adaptive", ctypes.c_int32),
        ("sh_active", ctypes.c_int32), ("sh_first", ctypes.c_int32), ("in_cage", ctypes.c_void_p),
    ]


class OrcRays(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t",
                 "gt_r", "gt_g", "gt_b", "gt_mask")] + [("white_bg", ctypes.c_int32)]


_P = ctypes.c_void_p
_I32, _I64, _F32, _F64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double


def _declare(L):
    L.orc_camera_from_pose.argtypes = [_P, _P, _F64, _I32, _I32, _F64, _F64, _F64, _F64, _P, _P, _P]
    L.orc_camera_from_pose.restype = None
    L.orc_rays_init.argtypes = [_P, _P, _F64, _F64, _F64, _F64, _P, _I32, _I32, _P, _P, _P, _P,
                                _I32] + [_P] * 12
    L.orc_rays_init.restype = None
    L.orc_build_occupancy.argtypes = [_P, _I32, _I32, _I32, _F32, _P, _P]
    L.orc_build_occupancy.restype = _I64
    L.orc_build_index.argtypes = [_P, _I32, _I32, _I32, _P]
    L.orc_build_index.restype = _I64
    L.orc_render_fwd.argtypes = [ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcRays), _P, _I64, _I64,
                                 _P, _P, _P, _P, _P, _P]
    L.orc_render_fwd.restype = None
    L.orc_render_bwd.argtypes = [ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcRays), _P, _I64, _I64,
                                 _P, _P, _P, _P, _F64, _F32, _P, _P]
    L.orc_render_bwd.restype = None
    L.orc_render_bwd_cond.argtypes = [ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcRays), _P,
                                      _I64, _I64, _P, _P, _P, _P, _F64, _F32, _P, _P, _P]
    L.orc_render_bwd_cond.restype = None
    L.orc_render_fwd_masked.argtypes = [ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcRays), _P, _I64,
                                        _I64, _P, ctypes.c_double, _P, _P, _P, _P, _P, _P, _P]
    L.orc_render_fwd_masked.restype = None
    L.orc_render_bwd_masked.argtypes = [ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcRays), _P, _I64,
                                        _I64, _P, _P, _P, _P, _P, ctypes.c_double,
                                        ctypes.c_float, _P, _P]
    L.orc_render_bwd_masked.restype = None
    L.orc_render_image.argtypes = [ctypes.POINTER(OrcGrid), _P, _P, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, _I32, _I32,
                                   _I32, _P, _P, _P, _P, _P]
    L.orc_render_image.restype = None
    L.orc_tv_disparity.argtypes = [ctypes.POINTER(OrcGrid), _F64, _F64, _I32, _P]
    L.orc_tv_disparity.restype = None
    L.orc_tv_neighbour6.argtypes = [ctypes.POINTER(OrcGrid), _F64, _F64, _P]
    L.orc_tv_neighbour6.restype = None
    L.orc_tv_neighbour6_cols.argtypes = [ctypes.POINTER(OrcGrid), _F64, _F64, _I64, _P]
    L.orc_tv_neighbour6_cols.restype = None
    L.orc_adam.argtypes = [_P, _P, _P, _P, _P, _I32, _I64, _F64, _F64, _I32]
    L.orc_adam.restype = None
    L.orc_adam_band.argtypes = [_P, _P, _P, _P, _P, _I32, _I32, _I64, _F64, _F64, _I32, _I32,
                                _I32]
    L.orc_adam_band.restype = None
    L.orc_upsample_density.argtypes = [ctypes.POINTER(OrcGrid), _I32, _I32, _I32, _F32, _P]
    L.orc_upsample_density.restype = None
    L.orc_upsample.argtypes = [ctypes.POINTER(OrcGrid), _I32, _I32, _I32, _I32, _P, _P, _I64,
                               _F32, _I32, _P]
    L.orc_upsample.restype = None
    L.orc_set_threads.argtypes = [_I32]
    L.orc_set_threads.restype = None
    L.orc_mse.argtypes = [_P, _I64]
    L.orc_mse.restype = _F64
    L.orc_now.argtypes = []
    L.orc_now.restype = _F64


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def camera_from_pose(translation, object_centre, camera_distance, fx, fy, cx, cy,
                     need_rescale=True, z_reversed=True):
    """Returns (c2w[4,4], w2c[4,4], centre[4]) of one pose."""
    t = _c(translation, np.float64)
    oc = _c(object_centre, np.float64)
    c2w = np.zeros(16, np.float64)
    w2c = np.zeros(16, np.float64)
    ctr = np.zeros(4, np.float64)
    lib().orc_camera_from_pose(_p(t), _p(oc), camera_distance, int(need_rescale),
                               int(z_reversed), fx, fy, cx, cy, _p(c2w), _p(w2c), _p(ctr))
    return c2w.reshape(4, 4), w2c.reshape(4, 4), ctr


def rays_init(c2w, centre, fx, fy, cx, cy, bbox, img_w, img_h, img=None, white_bg=True):
    """Rays of one view; img = (r, g, b, mask) uint8 arrays of img_h*img_w or None."""
    n = img_w * img_h
    c2w = _c(c2w, np.float64).reshape(16)
    centre = _c(centre, np.float64)
    bb = _c(bbox, np.int32)
    out = {k: np.zeros(n, np.float32) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t")}
    out.update({k: np.zeros(n, np.uint8) for k in ("gt_r", "gt_g", "gt_b", "gt_mask")})
    ims = [None] * 4 if img is None else [_c(x, np.uint8).reshape(n) for x in img]
    lib().orc_rays_init(_p(c2w), _p(centre), fx, fy, cx, cy, _p(bb), img_w, img_h,
                        *[_p(x) for x in ims], int(white_bg),
                        *[_p(out[k]) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "near_t",
                                               "far_t", "gt_r", "gt_g", "gt_b", "gt_mask")])
    return out


def build_occupancy(density, W, H, D, threshold):
    """Returns (occupancy bytes per cell, in_cage bytes per vertex, n_occupied)."""
    d = _c(density, np.float32).reshape(-1)
    occ = np.zeros((W - 1) * (H - 1) * (D - 1), np.uint8)
    cage = np.zeros(W * H * D, np.uint8)
    n = lib().orc_build_occupancy(_p(d), W, H, D, float(threshold), _p(occ), _p(cage))
    return occ, cage, n


def build_index(in_cage, W, H, D):
    """Returns (index_offsets uint32[W*H*D], inCageNums)."""
    c = _c(in_cage, np.uint8).reshape(-1)
    io = np.zeros(W * H * D, np.uint32)
    n = lib().orc_build_index(_p(c), W, H, D, _p(io))
    return io, n


class Grid:
    """Grid for the oracle: planes float32[P, n_vox] (reference layout)."""

    def __init__(self, W, H, D, S, planes, step, bbox=None, occupancy=None,
                 index_offsets=None, adaptive=False, sh_active=0, sh_first=0, in_cage=None):
        self.W, self.H, self.D, self.S = W, H, D, S
        self.planes = _c(planes, np.float32).reshape(1 + 3 * S, -1)
        self.n_vox = self.planes.shape[1]
        self.sparse = occupancy is not None
        self.occupancy = None if occupancy is None else _c(occupancy, np.uint8)
        self.index_offsets = None if index_offsets is None else _c(index_offsets, np.uint32)
        self.step = step
        self.bbox = bbox if bbox is not None else (0, W - 1, 0, H - 1, 0, D - 1)
        self.adaptive = adaptive
        self.sh_active = int(sh_active)
        self.sh_first = int(sh_first)
        self.in_cage = None if in_cage is None else _c(in_cage, np.uint8)

    def c(self):
        g = OrcGrid()
        g.W, g.H, g.D, g.S = self.W, self.H, self.D, self.S
        g.sparse = int(self.sparse)
        g.n_vox = self.n_vox
        g.planes = _p(self.planes)
        g.occupancy = _p(self.occupancy)
        g.index_offsets = _p(self.index_offsets)
        g.bbox = (ctypes.c_int32 * 6)(*self.bbox)
        g.step = self.step
        g.adaptive = int(self.adaptive)
        g.sh_active = self.sh_active
        g.sh_first = self.sh_first
        g.in_cage = _p(self.in_cage)
        return g


class Rays:
    """Ray set for the oracle: dict of numpy arrays as produced by rays_init."""

    def __init__(self, d, white_bg=True):
        self.d = {k: _c(v, np.uint8 if k.startswith("gt_") else np.float32) for k, v in d.items()}
        self.white_bg = white_bg
        self.n = len(self.d["ox"])

    def c(self):
        r = OrcRays()
        for k in ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t", "gt_r", "gt_g", "gt_b",
                  "gt_mask"):
            setattr(r, k, _p(self.d[k]))
        r.white_bg = int(self.white_bg)
        return r


def render_fwd(grid, rays, sample_index, start=0, length=None, trace=False):
    """Returns dict r, g, b, w, density (float64 per ray) and optional trace[len, 2]."""
    si = _c(sample_index, np.int32)
    if length is None:
        length = len(si) - start
    out = {k: np.zeros(rays.n, np.float64) for k in ("r", "g", "b", "w", "density")}
    tr = np.zeros((length, 2), np.uint32) if trace else None
    g, r = grid.c(), rays.c()
    lib().orc_render_fwd(ctypes.byref(g), ctypes.byref(r), _p(si), start, length,
                         _p(out["r"]), _p(out["g"]), _p(out["b"]), _p(out["w"]),
                         _p(out["density"]), _p(tr))
    if trace:
        out["trace"] = tr
    return out


def render_bwd(grid, rays, sample_index, fwd, mask_coef=1.0, huber_delta=0.1, start=0,
               length=None, grad=None):
    """Returns (losses float64[len], grad float64[P, n_vox])."""
    si = _c(sample_index, np.int32)
    if length is None:
        length = len(si) - start
    P = 1 + 3 * grid.S
    if grad is None:
        grad = np.zeros((P, grid.n_vox), np.float64)
    losses = np.zeros(length, np.float64)
    f = {k: _c(fwd[k], np.float64) for k in ("r", "g", "b", "w")}
    g, r = grid.c(), rays.c()
    lib().orc_render_bwd(ctypes.byref(g), ctypes.byref(r), _p(si), start, length,
                         _p(f["r"]), _p(f["g"]), _p(f["b"]), _p(f["w"]), mask_coef,
                         huber_delta, _p(losses), _p(grad))
    return losses, grad


def render_bwd_cond(grid, rays, sample_index, fwd, mask_coef=1.0, huber_delta=0.1, start=0,
                    length=None):
    """render_bwd that also returns the magnitude sum of every entry's terms:
    (losses, grad, grad_abs)."""
    si = _c(sample_index, np.int32)
    if length is None:
        length = len(si) - start
    P = 1 + 3 * grid.S
    grad = np.zeros((P, grid.n_vox), np.float64)
    gabs = np.zeros((P, grid.n_vox), np.float64)
    losses = np.zeros(length, np.float64)
    f = {k: _c(fwd[k], np.float64) for k in ("r", "g", "b", "w")}
    g, r = grid.c(), rays.c()
    lib().orc_render_bwd_cond(ctypes.byref(g), ctypes.byref(r), _p(si), start, length,
                              _p(f["r"]), _p(f["g"]), _p(f["b"]), _p(f["w"]), mask_coef,
                              huber_delta, _p(losses), _p(grad), _p(gabs))
    return losses, grad, gabs


def render_fwd_masked(grid, rays, sample_index, bin_mask=None, target_wmin=None, start=0,
                      length=None, trace=False):
    """render_fwd with a per-batch-entry proposal mask (uint64[len] or None);
    with target_wmin, also out["targets"]: per batch entry the bins holding a
    sample of weight T alpha > target_wmin."""
    si = _c(sample_index, np.int32)
    if length is None:
        length = len(si) - start
    out = {k: np.zeros(rays.n, np.float64) for k in ("r", "g", "b", "w", "density")}
    tr = np.zeros((length, 2), np.uint32) if trace else None
    bm = None if bin_mask is None else _c(bin_mask, np.uint64)
    tg = np.zeros(length, np.uint64) if target_wmin is not None else None
    g, r = grid.c(), rays.c()
    lib().orc_render_fwd_masked(ctypes.byref(g), ctypes.byref(r), _p(si), start, length, _p(bm),
                                0.0 if target_wmin is None else float(target_wmin), _p(tg),
                                _p(out["r"]), _p(out["g"]), _p(out["b"]), _p(out["w"]),
                                _p(out["density"]), _p(tr))
    if trace:
        out["trace"] = tr
    if tg is not None:
        out["targets"] = tg
    return out


def render_image(grid, c2w, centre, fx, fy, cx, cy, width, height, white_bg=True):
    """Eval image (renderImageGPU): dict rgb [3, h*w] float64, rgb8 [3, h*w]
    uint8, mask [h*w] uint8, disparity [h*w], rays [h*w, 5] (t_min, t_max,
    direction)."""
    n = width * height
    out = {"rgb": np.zeros((3, n), np.float64), "rgb8": np.zeros((3, n), np.uint8),
           "mask": np.zeros(n, np.uint8), "disparity": np.zeros(n, np.float64),
           "rays": np.zeros((n, 5), np.float64)}
    c2w = _c(np.asarray(c2w, np.float64).reshape(16), np.float64)
    centre = _c(np.asarray(centre, np.float64).reshape(4), np.float64)
    g = grid.c()
    lib().orc_render_image(ctypes.byref(g), _p(c2w), _p(centre), fx, fy, cx, cy, width, height,
                           int(white_bg), _p(out["rgb"]), _p(out["rgb8"]), _p(out["mask"]),
                           _p(out["disparity"]), _p(out["rays"]))
    return out


def render_bwd_masked(grid, rays, sample_index, fwd, bin_mask=None, mask_coef=1.0,
                      huber_delta=0.1, start=0, length=None, grad=None):
    si = _c(sample_index, np.int32)
    if length is None:
        length = len(si) - start
    P = 1 + 3 * grid.S
    if grad is None:
        grad = np.zeros((P, grid.n_vox), np.float64)
    losses = np.zeros(length, np.float64)
    f = {k: _c(fwd[k], np.float64) for k in ("r", "g", "b", "w")}
    bm = None if bin_mask is None else _c(bin_mask, np.uint64)
    g, r = grid.c(), rays.c()
    lib().orc_render_bwd_masked(ctypes.byref(g), ctypes.byref(r), _p(si), start, length, _p(bm),
                                _p(f["r"]), _p(f["g"]), _p(f["b"]), _p(f["w"]), mask_coef,
                                huber_delta, _p(losses), _p(grad))
    return losses, grad


def tv_disparity(grid, lambda_density, lambda_sh, optimize_step=0, grad=None):
    P = 1 + 3 * grid.S
    if grad is None:
        grad = np.zeros((P, grid.n_vox), np.float64)
    g = grid.c()
    lib().orc_tv_disparity(ctypes.byref(g), lambda_density, lambda_sh, optimize_step, _p(grad))
    return grad


def tv_neighbour6(grid, lambda_density, lambda_sh, grad=None, col_limit=None):
    """totalVarianceLoss, 6 neighbours; col_limit: columns y W + x covered
    (default W H, type 0; tv_both passes the type-2 launch's)."""
    P = 1 + 3 * grid.S
    if grad is None:
        grad = np.zeros((P, grid.n_vox), np.float64)
    if grid.sparse and grid.in_cage is None:
        raise ValueError("the 6-neighbour TV of a caged grid needs in_cage")
    g = grid.c()
    limit = grid.W * grid.H if col_limit is None else int(col_limit)
    lib().orc_tv_neighbour6_cols(ctypes.byref(g), lambda_density, lambda_sh, limit, _p(grad))
    return grad


def tv_both(grid, lambda_density, lambda_sh, optimize_step=0, grad=None):
    """VarianceType 2 (RenderUtils.cu:2840-2845): totalVarianceLoss over the
    columns of a ceil((W-1)/16) x ceil((H-1)/16) launch of 16 x 16 blocks,
    then addVarianceLossGPU."""
    limit = ((grid.W - 1 + 15) // 16) * ((grid.H - 1 + 15) // 16) * 256
    grad = tv_neighbour6(grid, lambda_density, lambda_sh, grad=grad, col_limit=limit)
    return tv_disparity(grid, lambda_density, lambda_sh, optimize_step, grad=grad)


def adam(params, grad, m, v, lr, sh_lr, t, adaptive_lr=None, band=None):
    """In-place Adam on float32 planes [P, n] (params, grad, m, v); band =
    (sh_lo, sh_hi) restricts the SH planes updated (the density always is)."""
    for a in (params, grad, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    P, n = params.shape
    alr = None if adaptive_lr is None else _c(adaptive_lr, np.float32)
    S, (lo, hi) = (0, (0, 0)) if band is None else ((P - 1) // 3, band)
    lib().orc_adam_band(_p(params), _p(grad), _p(m), _p(v), _p(alr), P, S, n, lr, sh_lr, t, lo,
                        hi)


def mse(losses):
    a = _c(losses, np.float64)
    return lib().orc_mse(_p(a), len(a))


def now():
    return lib().orc_now()


def upsample_density(grid, W2, H2, D2, threshold):
    """Thresholded density of every vertex of the W2 x H2 x D2 grid
    (upsampleScene into the dense temporary, SceneUtils.cu:540-546)."""
    out = np.zeros(W2 * H2 * D2, np.float32)
    g = grid.c()
    lib().orc_upsample_density(ctypes.byref(g), W2, H2, D2, threshold, _p(out))
    return out


def upsample(grid, W2, H2, D2, S2, threshold, low_band=False, in_cage=None, index_offsets=None,
             n=None):
    """upsampleScene (SceneUtils.cu:419-485): planes float32[1 + 3 S2, n] of
    the destination grid, dense (n = W2 H2 D2) or caged (in_cage, index)."""
    if in_cage is None:
        n = W2 * H2 * D2
    out = np.zeros((1 + 3 * S2, n), np.float32)
    cage = None if in_cage is None else _c(in_cage, np.uint8)
    io = None if index_offsets is None else _c(index_offsets, np.uint32)
    g = grid.c()
    lib().orc_upsample(ctypes.byref(g), W2, H2, D2, S2, _p(cage), _p(io), n, threshold,
                       int(low_band), _p(out))
    return out


def set_threads(n):
    """Threads for the oracle's loops (1 = reproducible sums, the default)."""
    lib().orc_set_threads(int(n))


# ------------------------------------------------------------ refprec --
# The train step in the reference's arithmetic types (refprec.c): the
# benchmark's CPU baseline, pinned against the FP64 functions above by
# tests/test_refprec.py.

_RP_PATH = os.path.join(_HERE, "librefprec.so")
_rp = None


def rp_lib():
    global _rp
    if _rp is None:
        if not os.path.exists(_RP_PATH):
            build()
        L = ctypes.CDLL(_RP_PATH)
        L.rp_render_train.argtypes = [ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcRays), _P, _I64,
                                      _F64, _F32, _P, _P, _P, _I32]
        L.rp_render_train.restype = None
        L.rp_tv_disparity.argtypes = [ctypes.POINTER(OrcGrid), _F64, _F64, _I32, _P, _I32]
        L.rp_tv_disparity.restype = None
        L.rp_adam.argtypes = [_P, _P, _P, _P, _I32, _I64, _F64, _F64, _I32, _I32]
        L.rp_adam.restype = None
        L.rp_now.argtypes = []
        L.rp_now.restype = _F64
        _rp = L
    return _rp


def rp_render_train(grid, rays, sample_index, grad, mask_coef=1.0, huber_delta=0.1, threads=1):
    """Forward + backward of the batch in the reference's types: returns
    (fwd float32[len, 5] = colour, opacity, density sum; losses float64[len]);
    gradients added into grad (float32[P, n_vox])."""
    si = _c(sample_index, np.int32)
    assert grad.dtype == np.float32 and grad.flags.c_contiguous
    fwd = np.zeros((len(si), 5), np.float32)
    losses = np.zeros(len(si), np.float64)
    g, r = grid.c(), rays.c()
    rp_lib().rp_render_train(ctypes.byref(g), ctypes.byref(r), _p(si), len(si), mask_coef,
                             huber_delta, _p(fwd), _p(losses), _p(grad), int(threads))
    return fwd, losses


def rp_tv_disparity(grid, lambda_density, lambda_sh, optimize_step, grad, threads=1):
    """addVarianceLossGPU in the reference's types into grad (float32[P, n_vox])."""
    assert grad.dtype == np.float32 and grad.flags.c_contiguous
    g = grid.c()
    rp_lib().rp_tv_disparity(ctypes.byref(g), lambda_density, lambda_sh, optimize_step, _p(grad),
                             int(threads))


def rp_adam(params, grad, m, v, lr, sh_lr, t, threads=1):
    """updateRGBSceneGPU_ADAM over float32[P, n] arrays (in place)."""
    for a in (params, grad, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    P, n = params.shape
    rp_lib().rp_adam(_p(params), _p(grad), _p(m), _p(v), P, n, lr, sh_lr, t, int(threads))


def rp_now():
    return rp_lib().rp_now()
