"""Train-step throughput of the plain method on one MI355X.

Workload (the 256^3 SH-2 configuration the metric is quoted on): 256^3
vertex grid, density from 64 seeded Gaussian blobs thresholded to ~10%
occupied cells and stored in the caged (sparse) layout; 100 views of
800x800 on the upper hemisphere of radius 4, fov 40 deg; targets rendered by
the plain method from the seeded ground-truth grid; the trained grid starts
at density 0, SH 0; 65536 rays per batch, step 0.5 voxel.

One step = one batch of Render::objectSceneLearner
(SparseVolumeRadianceField/Render.cu:352-503): forward + backward as the
two-kernel logged step (pxl_train_rays_sched; the trainer times every batch
order and launch schedule on the first, untimed batch and keeps the
fastest), FP64 batch loss on the GPU, the disparity TV gradient every third
batch (sparse grids, VarianceLossInterval 3), Adam, LR decay.

Prints one JSON line.  `value` is rays/s over the timed steps with all
inputs resident in HBM.
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "clear-plenoxels_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=("plain256", "uneven512", "dense128"),
                    default="plain256",
                    help="plain256: the 256^3 FP32 configuration the metric is quoted on; "
                         "uneven512: 512^3 shells + blobs, BF16 records, 262144 "
                         "rays/batch from a radius-6 sphere, autotuned schedule; "
                         "dense128: the small plain-method scene (dense 128^3, 16 blobs, "
                         "20 views 200x200 on a sphere, 4096 rays/batch), the scene of the "
                         "learned-proposal configuration")
    ap.add_argument("--grid", type=int, default=None)
    ap.add_argument("--views", type=int, default=None)
    ap.add_argument("--img", type=int, default=None)
    ap.add_argument("--batch", type=int, default=None)
    ap.add_argument("--occupancy", type=float, default=None)
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: the trainer times every batch order x schedule on the first "
                         "(untimed) batch and keeps the fastest; 0: given order, 4 waves per "
                         "block")
    ap.add_argument("--optim-stride", type=int, default=0,
                    help="floats per moment and master record; 0: P rounded "
                         "up to 4 (28 for SH-2), which streams no padding")
    ap.add_argument("--grad-stride", type=int, default=0,
                    help="floats per gradient record; 0: the FP32 record stride (32 up to 32 "
                         "parameters), so that the gradient walk flushes whole 128-byte lines")
    ap.add_argument("--bf16-master", type=int, default=1,
                    help="with --storage bf16: Adam updates an FP32 master copy (1) or the "
                         "widened BF16 values themselves (0)")
    ap.add_argument("--storage", choices=("f32", "bf16"), default="f32",
                    help="plain256: grid records in FP32 or BF16 (FP32 master copy for Adam)")
    ap.add_argument("--proposal", action="store_true",
                    help="plain256: train with the learned proposal (pretrain untimed, then "
                         "time pruned steps) and report held-out PSNR of pruned vs plain renders")
    ap.add_argument("--pretrain", type=int, default=None,
                    help="untimed training steps before warmup (default 2000 with --proposal, "
                         "else 0)")
    ap.add_argument("--prop-interval", type=int, default=8,
                    help="every n-th timed batch runs unpruned and trains the proposal")
    ap.add_argument("--prop-threshold", type=float, default=0.05,
                    help="bin probability below which the proposal prunes the bin")
    ap.add_argument("--prop-dilate", type=int, default=1,
                    help="bins kept on each side of a bin the proposal keeps")
    ap.add_argument("--cpu-rays", type=int, default=0,
                    help="rays of the batch the CPU baseline marches (0: the whole batch)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-order", choices=("shuffled", "sorted"), default="sorted",
                    help="order of the rays inside each batch (the batches themselves are "
                         "the same consecutive chunks of one seeded shuffle)")
    ap.add_argument("--log-samples", type=int, default=512,
                    help="samples per ray the two-kernel fused step logs (0: one fused kernel)")
    ap.add_argument("--march", choices=("fixed", "adaptive"), default="fixed",
                    help="marching rule of the trained grid: fixed step (the benchmark "
                         "configurations) or the reference's t += |step - w|")
    ap.add_argument("--fast-path", type=int, default=None,
                    help="1: after the plain measurement, measure the learned proposal with "
                         "BF16 records on the same rays (default on for the plain256 workload)")
    ap.add_argument("--fast-pretrain", type=int, default=1000,
                    help="fast path: untimed steps with the proposal learning on every batch")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a kernel of "
                         "known traffic for calibrating profiler byte counters)")
    a = ap.parse_args()
    uneven = a.workload == "uneven512"
    dense = a.workload == "dense128"
    a.grid = a.grid or (512 if uneven else (128 if dense else 256))
    a.batch = a.batch or (262144 if uneven else (4096 if dense else 65536))
    a.views = a.views or (20 if dense else 100)
    a.img = a.img or (200 if dense else 800)
    a.occupancy = a.occupancy if a.occupancy is not None else (0.01 if uneven else 0.10)
    if uneven and a.log_samples == 512:
        a.log_samples = 1024
    if a.log_samples == 0:
        # The single fused kernel has one fixed schedule: nothing to tune.
        a.autotune = 0
    if a.pretrain is None:
        a.pretrain = 2000 if a.proposal else 0
    if a.fast_path is None:
        a.fast_path = int(a.workload == "plain256" and not a.proposal and a.storage == "f32"
                          and a.march == "fixed")
    return a


def occupancy_threshold(dens, N, frac):
    """Density threshold that leaves `frac` of the (N-1)^3 cells occupied
    (a cell is occupied when a corner exceeds it)."""
    import torch

    ncell = (N - 1) ** 3
    cellmax = torch.nn.functional.max_pool3d(dens.view(1, 1, N, N, N), 2, stride=1).reshape(-1)
    k = int(round((1.0 - frac) * ncell))
    tau = float(torch.kthvalue(cellmax.float().cpu(), k).values)
    return tau


def caged_grid(dens, N, tau, dev):
    """Occupancy bits, cage, caged index and brick bits of a dense density."""
    import torch

    import pxl

    ncell = (N - 1) ** 3
    occ = torch.zeros((ncell + 31) // 32, dtype=torch.int32, device=dev)
    cage = torch.zeros(N ** 3, dtype=torch.uint8, device=dev)
    nocc = torch.zeros(1, dtype=torch.int64, device=dev)
    pxl.build_occupancy(dens, N, N, N, tau, occ, cage, nocc)
    io = torch.zeros(N ** 3, dtype=torch.int32, device=dev)
    c2d = torch.zeros(N ** 3, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    scratch = torch.zeros(pxl.sparse_index_scratch_bytes(N ** 3) // 4 + 1, dtype=torch.int32,
                          device=dev)
    pxl.build_sparse_index(cage, N, N, N, io, c2d, cnt, scratch)
    torch.cuda.synchronize()
    n_vox = int(cnt.item())
    c2d = c2d[:n_vox].clone()
    bricks = pxl.build_brick_occupancy(occ, N, N, N)
    kw = dict(step=0.5, occupancy=occ, index_offsets=io, in_cage=cage, compact_to_dense=c2d,
              brick_occupancy=bricks)
    return kw, n_vox, int(nocc.item()) / ncell


def set_targets(rays, grid, n_rays, dev):
    """Render `grid` without background and store the un-premultiplied
    colours and the opacity mask as the rays' 8-bit targets, so the
    white-background compositing of the train step reproduces the render."""
    import torch

    import pxl

    allidx = torch.arange(n_rays, dtype=torch.int32, device=dev)
    chunk = 1 << 22
    for s0 in range(0, n_rays, chunk):
        pxl.render_fwd(grid, rays, allidx, start=s0, length=min(chunk, n_rays - s0))
    torch.cuda.synchronize()
    A = rays.acc_w.clamp(0, 1)
    inv = 1.0 / A.clamp_min(1e-6)
    to_u8 = lambda x: (x.clamp(0, 1) * 255.0 + 0.5).floor().to(torch.uint8)
    rays.gt_r.copy_(to_u8(rays.acc_r * inv))
    rays.gt_g.copy_(to_u8(rays.acc_g * inv))
    rays.gt_b.copy_(to_u8(rays.acc_b * inv))
    rays.gt_mask.copy_(to_u8(A))
    rays.white_bg = True


class UnevenWorkload:
    """512^3 thin shells + blobs (seed 41), cells occupied where a corner
    exceeds half the shell amplitude (a few % of cells), caged,
    SH-2 stored in BF16 with an FP32 master copy; rays with origins uniform
    on a radius-6 sphere (grid half-width 1) aimed at uniform points of
    [-0.5, 0.5]^3, so per-ray sample counts run from 0 to ~2000; four
    262144-ray batches with targets rendered from the seeded ground truth."""

    def __init__(self, args, dev):
        import torch

        import pxl
        from pxl import scenes

        self.args = args
        N, S = args.grid, 9
        P = 1 + 3 * S
        self.N, self.S, self.P = N, S, P
        t0 = time.time()
        dens = scenes.shell_density(N, 41, device=dev).reshape(-1)
        # Half the shell amplitude: every shell cell and the blob cores are
        # occupied (a quantile of the cell maxima would fall above the
        # shells, whose density is one constant).
        tau = 20.0
        grid_kw, n_vox, self.occ_frac = caged_grid(dens, N, tau, dev)
        self.n_vox = n_vox
        c2d = grid_kw["compact_to_dense"]
        gen = torch.Generator(device=dev)
        gen.manual_seed(42)
        rec32 = torch.zeros((n_vox, pxl.record_stride(S)), dtype=torch.float32, device=dev)
        rec32[:, 0] = dens[c2d.long()]
        rec32[:, 1:P].normal_(0.0, 0.3, generator=gen)
        del dens
        gt16 = pxl.records_to_16bit(rec32, pxl.STORE_BF16)
        self.gt_grid = pxl.Grid(N, N, N, S, gt16, storage=pxl.STORE_BF16, **grid_kw)
        log(f"grid {N}^3: occupied cells {self.occ_frac:.4f}, caged vertices {n_vox} "
            f"({time.time() - t0:.1f}s)")

        B = args.batch
        self.n_rays = 4 * B
        gen.manual_seed(43)
        half = N / 2.0
        u = torch.randn((self.n_rays, 3), generator=gen, device=dev, dtype=torch.float64)
        origin = 6.0 * u / u.norm(dim=1, keepdim=True)
        target = torch.rand((self.n_rays, 3), generator=gen, device=dev, dtype=torch.float64) - 0.5
        o = origin * half + half
        d = (target - origin)
        d = (d / d.norm(dim=1, keepdim=True)).float()
        lo, hi = 0.0, float(N - 1)
        dd = d.double()
        inv = 1.0 / torch.where(dd == 0, torch.full_like(dd, 1e-300), dd)
        t1, t2 = (lo - o) * inv, (hi - o) * inv
        tmin = torch.minimum(t1, t2).amax(dim=1).clamp_min(0.0)
        tmax = torch.maximum(t1, t2).amin(dim=1)
        rays = pxl.Rays(self.n_rays, device=dev, white_bg=False)
        for i, k in enumerate(("ox", "oy", "oz")):
            getattr(rays, k).copy_(o[:, i].float())
        for i, k in enumerate(("dx", "dy", "dz")):
            getattr(rays, k).copy_(d[:, i])
        rays.near_t.copy_(tmin.float())
        rays.far_t.copy_(tmax.float())
        set_targets(rays, self.gt_grid, self.n_rays, dev)
        self.rays = rays
        self.order = torch.arange(self.n_rays, dtype=torch.int32, device=dev)
        tr = torch.zeros(2 * B, dtype=torch.int32, device=dev)
        pxl.render_fwd(self.gt_grid, rays, self.order, length=B, trace=tr)
        cnt = tr.view(-1, 2)[:, 0].double().cpu()
        self.samples = {"mean": round(cnt.mean().item(), 2), "max": int(cnt.max().item()),
                        "zero_frac": round((cnt == 0).double().mean().item(), 4)}
        log(f"rays {self.n_rays}, samples/ray {self.samples} ({time.time() - t0:.1f}s)")

        zero32 = torch.zeros_like(rec32)
        self.rec = pxl.records_to_16bit(zero32, pxl.STORE_BF16)
        self.grid = pxl.Grid(N, N, N, S, self.rec, storage=pxl.STORE_BF16, **grid_kw)
        ostride = args.optim_stride or (P + 3) // 4 * 4
        self.optim = pxl.Optim(self.grid, stride=ostride,
                               grad_stride=args.grad_stride or pxl.record_stride(S),
                               master_from=zero32[:, :ostride] if args.bf16_master else None)
        del rec32
        self.trainer = pxl.Trainer(self.grid, self.optim, self.rays, self.order, batch_size=B,
                                   learning_rate=0.3, sh_learning_rate=1.0,
                                   variance_coefficient=1e-5, sh_coefficient=1e-6,
                                   variance_type=pxl.TV_DISPARITY, neighbour_nums=6,
                                   variance_interval=3, mask_coefficient=1.0, huber_delta=0.1,
                                   fused=True, log_samples=args.log_samples,
                                   autotune=bool(args.autotune))

    def step(self):
        self.trainer.step()


class Workload:
    """GPU-resident scene, rays and optimiser state."""

    def __init__(self, args, dev):
        import numpy as np
        import torch

        import pxl
        from pxl import scenes

        self.args = args
        N, S = args.grid, 9
        P = 1 + 3 * S
        stride = pxl.record_stride(S)
        self.N, self.S, self.P = N, S, P
        self.dense = args.workload == "dense128"
        t0 = time.time()
        if self.dense:
            gt_rec, grid_kw = self._dense_scene(dev, stride)
        else:
            gt_rec, grid_kw = self._caged_scene(dev, stride)
        self.gt_grid = pxl.Grid(N, N, N, S, gt_rec, **grid_kw)
        if self.dense:
            log(f"grid {N}^3 dense, {self.n_vox} vertices ({time.time() - t0:.1f}s)")
        else:
            log(f"grid {N}^3: tau={self.tau:.3f}, occupied cells {self.occ_frac:.4f}, "
                f"caged vertices {self.n_vox} ({time.time() - t0:.1f}s)")
        self._rays_and_trainer(dev, gt_rec, grid_kw, t0)

    def _dense_scene(self, dev, stride):
        """The small plain-method scene: 16 blobs (seed 1), SH ~ N(0, 0.3)."""
        import torch

        from pxl import scenes

        N, P = self.N, self.P
        self.occ_frac, self.tau = 1.0, None
        self.n_vox = N ** 3
        gen = torch.Generator(device=dev)
        gen.manual_seed(101)
        rec = torch.zeros((N ** 3, stride), dtype=torch.float32, device=dev)
        rec[:, 0] = scenes.blob_density(N, 16, 1, device=dev).reshape(-1)
        rec[:, 1:P].normal_(0.0, 0.3, generator=gen)
        return rec, dict(step=0.5)

    def _caged_scene(self, dev, stride):
        import torch

        import pxl
        from pxl import scenes

        N, P, args = self.N, self.P, self.args
        # Ground-truth grid, dense first.
        dens = scenes.blob_density(N, 64, 11, device=dev).reshape(-1)
        ncell = (N - 1) ** 3
        cellmax = torch.nn.functional.max_pool3d(dens.view(1, 1, N, N, N), 2, stride=1).reshape(-1)
        k = int(round((1.0 - args.occupancy) * ncell))
        tau = float(torch.kthvalue(cellmax.float().cpu(), k).values)
        del cellmax
        occ = torch.zeros((ncell + 31) // 32, dtype=torch.int32, device=dev)
        cage = torch.zeros(N ** 3, dtype=torch.uint8, device=dev)
        nocc = torch.zeros(1, dtype=torch.int64, device=dev)
        pxl.build_occupancy(dens, N, N, N, tau, occ, cage, nocc)
        io = torch.zeros(N ** 3, dtype=torch.int32, device=dev)
        c2d = torch.zeros(N ** 3, dtype=torch.int32, device=dev)
        cnt = torch.zeros(1, dtype=torch.int32, device=dev)
        scratch = torch.zeros(pxl.sparse_index_scratch_bytes(N ** 3) // 4 + 1, dtype=torch.int32,
                              device=dev)
        pxl.build_sparse_index(cage, N, N, N, io, c2d, cnt, scratch)
        torch.cuda.synchronize()
        n_vox = int(cnt.item())
        self.occ_frac = int(nocc.item()) / ncell
        self.tau = tau
        self.n_vox = n_vox
        c2d = c2d[:n_vox].clone()
        gen = torch.Generator(device=dev)
        gen.manual_seed(12)
        dense_rec = torch.zeros((N ** 3, stride), dtype=torch.float32, device=dev)
        dense_rec[:, 0] = dens
        dense_rec[:, 1:P].normal_(0.0, 0.3, generator=gen)
        gt_rec = torch.zeros((n_vox, stride), dtype=torch.float32, device=dev)
        pxl.grid_compact(dense_rec, c2d, n_vox, gt_rec)
        del dense_rec, dens
        bricks = pxl.build_brick_occupancy(occ, N, N, N)
        grid_kw = dict(step=0.5, occupancy=occ, index_offsets=io, in_cage=cage,
                       compact_to_dense=c2d, brick_occupancy=bricks)
        return gt_rec, grid_kw

    def _rays_and_trainer(self, dev, gt_rec, grid_kw, t0):
        import torch

        import pxl
        from pxl import scenes

        N, S, args = self.N, self.S, self.args
        # Rays of all views (upper hemisphere; the whole sphere for the dense
        # scene).
        img, V = args.img, args.views
        f, _, c, _ = scenes.view_intrinsics(img, 40.0)
        self.cams = [pxl.camera_from_pose(t, [N / 2.0] * 3, 1.0, f, f, c, c)[0]
                     for t in scenes.fibonacci_sphere(V, 4.0, upper_only=not self.dense)]
        npix = img * img
        self.n_rays = V * npix
        rays = pxl.Rays(self.n_rays, device=dev, white_bg=False)
        bbox = (0, N - 1, 0, N - 1, 0, N - 1)
        self.bbox = bbox
        for v, cam in enumerate(self.cams):
            pxl.rays_init(cam, bbox, img, img, rays, start_index=v * npix)
        # Targets: render the ground truth without background, then set the
        # per-view images (un-premultiplied colour, alpha mask) so that the
        # white-background compositing of the ray builder reproduces the
        # white-background render.
        allidx = torch.arange(self.n_rays, dtype=torch.int32, device=dev)
        chunk = 1 << 22
        for i, s0 in enumerate(range(0, self.n_rays, chunk)):
            pxl.render_fwd(self.gt_grid, rays, allidx, start=s0,
                           length=min(chunk, self.n_rays - s0))
            if i % 4 == 3:
                torch.cuda.synchronize()
                log(f"targets: {s0 + chunk} / {self.n_rays} rays ({time.time() - t0:.1f}s)")
        torch.cuda.synchronize()
        A = rays.acc_w.clamp(0, 1)
        inv = 1.0 / A.clamp_min(1e-6)
        to_u8 = lambda x: (x.clamp(0, 1) * 255.0 + 0.5).floor().to(torch.uint8)
        im_r, im_g, im_b = to_u8(rays.acc_r * inv), to_u8(rays.acc_g * inv), to_u8(rays.acc_b * inv)
        im_m = to_u8(A)
        rays.white_bg = True
        for v, cam in enumerate(self.cams):
            sl = slice(v * npix, (v + 1) * npix)
            pxl.rays_init(cam, bbox, img, img, rays, start_index=v * npix,
                          img=(im_r[sl], im_g[sl], im_b[sl], im_m[sl]))
        del im_r, im_g, im_b, im_m, A, inv, allidx
        self.rays = rays
        gen = torch.Generator(device=dev)
        gen.manual_seed(3)
        self.order = torch.randperm(self.n_rays, generator=gen, device=dev).to(torch.int32)
        if args.batch_order == "sorted":
            # Same batches (consecutive chunks of the shuffle), rays inside
            # each batch in ray-index order: a view's rays become neighbours.
            B = args.batch
            full = (self.n_rays // B) * B
            head = self.order[:full].view(-1, B).sort(dim=1).values.reshape(-1)
            self.order = torch.cat([head, self.order[full:].sort().values])
        torch.cuda.synchronize()
        log(f"rays {self.n_rays} built with targets ({time.time() - t0:.1f}s)")

        # Trained grid: same caged structure, InitialDensity 0, InitialColor 0.
        self.grid_kw = grid_kw
        if args.march == "adaptive":
            grid_kw = dict(grid_kw, march_mode=pxl.MARCH_REF_ADAPTIVE)
        self.rec = torch.zeros_like(gt_rec)
        # Moment and master records of P floats (the Adam pass streams no
        # padding); gradient records of 32 floats, one 128-byte line each,
        # which the gradient walk's atomic flushes fill whole.
        ostride = args.optim_stride or (self.P + 3) // 4 * 4
        if args.storage == "bf16":
            zero32 = self.rec
            self.rec = pxl.records_to_16bit(zero32, pxl.STORE_BF16)
            self.grid = pxl.Grid(N, N, N, S, self.rec, storage=pxl.STORE_BF16, **grid_kw)
            self.optim = pxl.Optim(self.grid, stride=ostride,
                                   grad_stride=args.grad_stride or pxl.record_stride(S),
                                   master_from=zero32[:, :ostride] if args.bf16_master else None)
        else:
            self.grid = pxl.Grid(N, N, N, S, self.rec, **grid_kw)
            self.optim = pxl.Optim(self.grid, stride=ostride,
                                   grad_stride=args.grad_stride or pxl.record_stride(S))
        # config_linux.txt defaults: LearningRate 0.3, SHLearningRate 1,
        # VarianceCoefficient 1e-5, ColorDisparityCoefficient 1e-6,
        # VarianceType 1, VarianceLossInterval 3, MaskRayCoefficient 1,
        # HuberDelta 0.1.  The batch loop itself runs in the C++ host layer.
        self.trainer = pxl.Trainer(self.grid, self.optim, self.rays, self.order,
                                   batch_size=args.batch, learning_rate=0.3,
                                   sh_learning_rate=1.0, variance_coefficient=1e-5,
                                   sh_coefficient=1e-6, variance_type=pxl.TV_DISPARITY,
                                   neighbour_nums=6, variance_interval=3, mask_coefficient=1.0,
                                   huber_delta=0.1, fused=True,
                                   log_samples=args.log_samples,
                                   autotune=bool(args.autotune))

    def step(self):
        self.trainer.step()


def touched_vertices(w, dev, nb=4):
    import torch

    import pxl

    bm = torch.zeros((w.n_vox + 31) // 32, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    tot = 0
    for b in range(nb):
        pxl.count_touched(w.grid, w.rays, w.order, bm, cnt, start=b * w.args.batch,
                          length=w.args.batch)
        torch.cuda.synchronize()
        tot += int(cnt.item())
    return tot / nb


def samples_per_batch(w, dev, nb=2):
    """Mean samples the plain marcher evaluates per batch (per-ray counts of
    a traced forward over the first nb batches; untimed)."""
    import torch

    import pxl

    B = w.args.batch
    tr = torch.zeros(2 * B, dtype=torch.int32, device=dev)
    tot = 0
    for b in range(nb):
        pxl.render_fwd(w.grid, w.rays, w.order, start=b * B, length=B, trace=tr)
        torch.cuda.synchronize()
        tot += int(tr.view(-1, 2)[:, 0].long().sum().item())
    return tot / nb


def cpu_cores():
    """Host threads for the CPU baseline: every core this process may run on,
    at most the OMP_NUM_THREADS the host sets (the host share of one GPU;
    16 when unset)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    try:
        cap = int(os.environ.get("OMP_NUM_THREADS", "16"))
    except ValueError:
        cap = 16
    return max(1, min(n, cap if cap > 0 else 16))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, dev, nrays):
    """The train step in the reference's arithmetic types (oracle/refprec.c:
    FP32 parameters, float interpolation and activations, FP64 t, T, weights
    and loss, float gradient atomics) on the host cores: forward + backward
    of nrays rays of the batch (0: the whole batch), the variance loss over
    the grid (every batch dense, every 3rd caged) and Adam over every
    record; each part once untimed, then the median of 5 timed runs."""
    import numpy as np

    import oracle

    cores = cpu_cores()
    P = w.P
    B = w.args.batch
    n = B if nrays <= 0 else min(nrays, B)
    planes = w.rec[:, :P].T.contiguous().cpu().numpy()
    if w.dense:
        occ = io = None
    else:
        occ_bits = w.grid.occupancy.cpu().numpy().view(np.uint32)
        ncell = (w.N - 1) ** 3
        occ = np.unpackbits(occ_bits.view(np.uint8), bitorder="little")[:ncell].astype(np.uint8)
        io = w.grid.index_offsets.cpu().numpy().view(np.uint32)
    keys = ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t", "gt_r", "gt_g", "gt_b",
            "gt_mask")
    idx = w.order[:n].long()
    sub = {k: getattr(w.rays, k)[idx].cpu().numpy() for k in keys}
    rays_o = oracle.Rays(sub, white_bg=bool(w.rays.white_bg))
    grid_o = oracle.Grid(w.N, w.N, w.N, w.S, planes, 0.5, occupancy=occ, index_offsets=io,
                         adaptive=w.args.march == "adaptive")
    si = np.arange(n, dtype=np.int32)
    grad = np.zeros_like(planes)
    params = planes.copy()
    m = np.zeros_like(planes)
    v = np.zeros_like(planes)

    def median5(fn):
        fn()
        times = []
        for _ in range(5):
            t0 = oracle.rp_now()
            fn()
            times.append(oracle.rp_now() - t0)
        return sorted(times)[2]

    t_rays = median5(lambda: oracle.rp_render_train(grid_o, rays_o, si, grad, 1.0, 0.1,
                                                    threads=cores)) / n
    t_tv = median5(lambda: oracle.rp_tv_disparity(grid_o, 1e-5, 1e-6, 0, grad, threads=cores))
    t_adam = median5(lambda: oracle.rp_adam(params, grad, m, v, 0.3, 1.0, 1, threads=cores))
    t_step = B * t_rays + (t_tv if w.dense else t_tv / 3.0) + t_adam
    nproc = os.cpu_count()
    model = cpu_model()
    sample = (f"the train step in the reference's arithmetic types (oracle/refprec.c), "
              f"{cores} OpenMP threads on {model} (nproc {nproc}); forward + backward of "
              f"{n} rays of the batch ({t_rays * 1e6:.2f} us/ray), variance loss over the "
              f"{'dense' if w.dense else 'caged'} grid ({t_tv:.3f} s, "
              f"{'every' if w.dense else 'every 3rd'} step), Adam over {w.n_vox} records "
              f"({t_adam:.3f} s); each part: 1 untimed run, median of 5; step = {B} rays")
    log(f"CPU baseline: {B / t_step:.1f} rays/s; {sample}")
    return {"value": round(B / t_step, 1), "unit": "rays/s", "cores": cores, "kind": "port",
            "sample": sample, "host": {"nproc": nproc, "cpu_model": model}}


TRAFFIC_FILE = os.path.join(REPO, "profiles", "train_step_256_sparse_f32.md")


def profiled_traffic(kernel, args):
    """HBM bytes per launch of `kernel` on the default workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (lines of the form
    `traffic_bytes_per_launch <kernel> <bytes>`); None for any other
    workload or kernel, or when the summary holds no figure for it."""
    default = (args.grid, args.views, args.img, args.batch, args.occupancy,
               args.batch_order, args.log_samples) == (256, 100, 800, 65536, 0.10, "sorted", 512)
    if not default or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.strip().strip("`").split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return int(parts[2])
    return None


def proposal_quality(w, prop, dev, args, grid=None, fp32_grid=None):
    """Held-out batch (the last of the shuffle, never trained on): sample
    counts and PSNR against the targets of the plain and the pruned render
    of the current grid (w.grid, or `grid`); with fp32_grid (the FP32 master
    records of a 16-bit grid) also the plain FP32 render's PSNR."""
    import torch

    import pxl

    grid = w.grid if grid is None else grid
    B = args.batch
    start = (w.n_rays // B - 1) * B
    idx = w.order[start:start + B]
    pws = pxl.proposal_workspace(B, dev)
    pxl.proposal_features(grid, w.rays, w.order, pws, start=start, length=B)
    logits = torch.zeros((B, 64), dtype=torch.float32, device=dev)
    pxl.proposal_forward(prop, B, logits, pws)
    masks = torch.zeros(B, dtype=torch.int64, device=dev)
    pxl.proposal_masks(logits, B, masks, threshold=args.prop_threshold, dilate=args.prop_dilate)
    gt = torch.stack([w.rays.gt_r[idx.long()], w.rays.gt_g[idx.long()],
                      w.rays.gt_b[idx.long()]], 1).double() / 255.0
    # Rays that miss the grid box are rendered black and carry no loss
    # (the reference's convention); PSNR is over the rays that enter it.
    hit = w.rays.near_t[idx.long()] <= w.rays.far_t[idx.long()]
    res = {}
    runs = [("plain", grid, None), ("proposal", grid, masks)]
    if fp32_grid is not None:
        runs.append(("plain_fp32", fp32_grid, None))
    for name, g, m in runs:
        tr = torch.zeros(2 * B, dtype=torch.int32, device=dev)
        pxl.render_fwd_masked(g, w.rays, w.order, m, start=start, length=B, trace=tr)
        c = torch.stack([w.rays.acc_r[idx.long()], w.rays.acc_g[idx.long()],
                         w.rays.acc_b[idx.long()]], 1).double()
        torch.cuda.synchronize()
        mse = float(((c - gt)[hit] ** 2).mean())
        res[name] = {"samples": int(tr.view(-1, 2)[:, 0].sum().item()),
                     "psnr": round(10.0 * math.log10(1.0 / mse), 4)}
    res["pruned_frac"] = round(1.0 - res["proposal"]["samples"] / max(1, res["plain"]["samples"]), 4)
    res["psnr_delta"] = round(res["proposal"]["psnr"] - res["plain"]["psnr"], 4)
    if fp32_grid is not None:
        res["psnr_delta_vs_fp32"] = round(res["proposal"]["psnr"] - res["plain_fp32"]["psnr"], 4)
    res["interval"] = args.prop_interval
    res["threshold"] = args.prop_threshold
    res["dilate"] = args.prop_dilate
    log(f"proposal on held-out batch: {res}")
    return res


def fast_path(w, dev, args, plain_ms):
    """The north-star fast path on the plain workload's rays and schedule:
    BF16 records with an FP32 master copy, the learned proposal learning on
    every batch for fast_pretrain untimed steps, then every prop_interval-th
    batch unpruned (learning) and the others pruned; warmup + `steps` timed
    steps.  Quality on the held-out batch: PSNR of the pruned BF16 render
    against the unpruned BF16 render and against the FP32 render of the
    master records (the plain FP32 path's values)."""
    import torch

    import pxl

    N, S = w.N, w.S
    zero32 = torch.zeros_like(w.rec)
    rec16 = pxl.records_to_16bit(zero32, pxl.STORE_BF16)
    grid = pxl.Grid(N, N, N, S, rec16, storage=pxl.STORE_BF16, **w.grid_kw)
    ostride = args.optim_stride or (w.P + 3) // 4 * 4
    optim = pxl.Optim(grid, stride=ostride, grad_stride=args.grad_stride or pxl.record_stride(S),
                      master_from=zero32[:, :ostride])
    del zero32
    tr = pxl.Trainer(grid, optim, w.rays, w.order, batch_size=args.batch, learning_rate=0.3,
                     sh_learning_rate=1.0, variance_coefficient=1e-5, sh_coefficient=1e-6,
                     variance_type=pxl.TV_DISPARITY, neighbour_nums=6, variance_interval=3,
                     mask_coefficient=1.0, huber_delta=0.1, fused=True,
                     log_samples=args.log_samples, autotune=bool(args.autotune))
    prop = pxl.Proposal(dev, seed=1)
    tr.set_proposal(prop, interval=1, prune=False)
    t0 = time.time()
    for _ in range(args.fast_pretrain):
        tr.step()
    torch.cuda.synchronize()
    log(f"fast path: {args.fast_pretrain} untimed steps with the proposal learning "
        f"({time.time() - t0:.1f}s), batch PSNR {tr.loss()[1]:.2f}")
    tr.set_proposal(prop, interval=args.prop_interval, prune=True, threshold=args.prop_threshold,
                    dilate=args.prop_dilate)
    for _ in range(args.warmup):
        tr.step()
    K = args.steps
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(K):
        tr.step()
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / K
    tr.sync()
    fp32 = pxl.Grid(N, N, N, S, optim.master, **w.grid_kw)
    q = proposal_quality(w, prop, dev, args, grid=grid, fp32_grid=fp32)
    out = {"config": "learned proposal + BF16 records (FP32 master), "
                     f"{args.fast_pretrain} untimed steps learning first, then every "
                     f"{args.prop_interval}th batch unpruned",
           "ms_per_step": round(ms, 4), "value": round(args.batch / (ms / 1e3), 1),
           "unit": "rays/s", "ratio_vs_plain": round(plain_ms / ms, 4),
           "held_out": q, "target_ratio": 3.0, "target_psnr_loss_db": 0.1}
    log(f"fast path: {ms:.3f} ms/step ({plain_ms / ms:.3f}x the plain path), held out {q}")
    tr.close()
    return out


def copy_bandwidth(dev, mib=1024, reps=5):
    """Device-to-device copy bandwidth (bytes read + written per second, GB/s)
    of a mib-MiB float buffer: the stream rate this box reaches, reported
    beside the published HBM peak."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * n * 4 * reps / (t0.elapsed_time(t1) / 1e3) / 1e9
    del src, dst
    return gbs


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU design is described, not built.", file=sys.stderr)
        sys.exit(2)
    import torch
    import pxl

    dev = torch.device("cuda:0")
    if args.copy_mib > 0:
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_mib} MiB read + {args.copy_mib} MiB written")
        del src, dst
    uneven = args.workload == "uneven512"
    w = UnevenWorkload(args, dev) if uneven else Workload(args, dev)
    prop = None
    if args.proposal:
        # Pretrain with the proposal learning on every (unpruned) batch, then
        # prune all but every prop_interval-th batch in the timed steps.
        prop = pxl.Proposal(dev, seed=1)
        w.trainer.set_proposal(prop, interval=1, prune=False)
        t0 = time.time()
        for i in range(args.pretrain):
            w.step()
            if (i + 1) % 500 == 0:
                torch.cuda.synchronize()
                mse, psnr = w.trainer.loss()
                log(f"pretrain {i + 1}: batch PSNR {psnr:.2f}, proposal BCE "
                    f"{w.trainer.proposal_loss() / (args.batch * 64):.4f} ({time.time() - t0:.1f}s)")
        w.trainer.set_proposal(prop, interval=args.prop_interval, prune=True,
                               threshold=args.prop_threshold, dilate=args.prop_dilate)
    elif args.pretrain:
        for _ in range(args.pretrain):
            w.step()
        torch.cuda.synchronize()
        log(f"pretrained {args.pretrain} steps, batch PSNR {w.trainer.loss()[1]:.2f}")
    for _ in range(args.warmup):
        w.step()
    torch.cuda.synchronize()

    K = args.steps
    # HIP events around each batch's render launch, recorded by the trainer
    # on the stream it launches on.  The roofline uses the mean over every
    # timed batch (on regulariser batches of a sparse grid the variance loss
    # runs beside the forward on the trainer's second stream, inside that
    # window); the mean over the other batches is reported beside it.
    w.trainer.time_render(K)
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    wall0 = time.time()
    for _ in range(K):
        w.step()
    end.record()
    torch.cuda.synchronize()
    wall = time.time() - wall0
    ms = start.elapsed_time(end)
    kern_ms, nk, alone_ms, n_alone = w.trainer.render_ms_split()
    assert 0 < nk <= K
    # The fused TV + Adam pass leaves the current records in either of two
    # buffers; the measurements below read the grid's.
    w.trainer.sync()
    ms_step = ms / K
    rays_per_s = args.batch * K / (ms / 1e3)
    mse, psnr = w.trainer.loss()
    log(f"{K} steps: {ms_step:.3f} ms/step, fused render {kern_ms:.3f} ms, wall {wall:.2f}s, "
        f"batch MSE {mse:.3f} PSNR {psnr:.2f}")

    # Algorithmic bytes of the render stage (the fused forward + backward,
    # one kernel or the logged pair): 44 B per ray (geometry, target,
    # outputs, loss) + 448 B per distinct touched record (its parameters read
    # by the forward and the backward, its gradient read and written).  The
    # step figure adds Adam's 896 B per record (parameters, gradient, two
    # moments: read + write).
    # 16-bit records: 2 P s_p + 8 P = 336 B per touched record and
    # P (2 s_p + 8 + 16) = 784 B per record for Adam.
    s_p = 2 if (uneven or args.storage == "bf16") else 4
    vt = touched_vertices(w, dev)
    alg = args.batch * 44 + vt * (2 * w.P * s_p + 8 * w.P)
    achieved = alg / (kern_ms / 1e3) / 1e9
    step_alg = alg + w.n_vox * w.P * (2 * s_p + 8 + 16)
    kname = pxl.render_kernel_name(w.grid, "logged" if args.log_samples > 0 else "fused")
    copy_gbs = copy_bandwidth(dev)
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "peak_measured_copy": round(copy_gbs, 1),
            "frac_of_measured_copy": round(achieved / copy_gbs, 5),
            "traffic": profiled_traffic(kname, args),
            "kernel": kname, "kernel_ms": round(kern_ms, 4), "kernel_batches_timed": int(nk),
            "kernel_ms_without_tv_beside": round(alone_ms, 4) if n_alone else None,
            "kernel_batches_without_tv_beside": int(n_alone),
            "touched_records": int(vt), "alg_bytes": int(alg),
            "step_alg_bytes": int(step_alg),
            "step_frac": round(step_alg / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 5)}
    # The L2-level figure, not HBM: per sample, the 8 corner records
    # gathered (8 P s_p bytes) and the reference's 8 P float atomics at 8 B.
    n_samp = samples_per_batch(w, dev)
    per_sample = 8 * w.P * s_p + 8 * w.P * 8
    roof["l2_level"] = {"samples_per_batch": int(n_samp), "bytes_per_sample": per_sample,
                        "GBps": round(n_samp * per_sample / (kern_ms / 1e3) / 1e9, 1)}

    cpu = None
    if (not args.no_cpu_baseline and not uneven and not args.proposal
            and args.storage == "f32" and not args.pretrain):
        log("CPU baseline (reference-precision restatement) ...")
        cpu = cpu_baseline(w, dev, args.cpu_rays)
    out = {
        "metric": f"train rays/sec (fwd+bwd+TV+Adam) {args.grid}^3 SH2"
                  + (" dense" if args.workload == "dense128" else ""),
        "value": round(rays_per_s, 1),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"batch_order": args.batch_order,
                   "workload": f"plain {args.storage.upper()}, {args.grid}^3 SH-2 "
                               + ("dense grid" if args.workload == "dense128" else
                                  f"caged grid at {w.occ_frac:.4f} cell occupancy")
                               + f", {args.views} views {args.img}x{args.img}, "
                                 f"{args.batch} rays/batch, step 0.5",
                   "global_batch": args.batch, "records": w.n_vox,
                   "train_psnr_last_batch": round(psnr, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if args.march == "adaptive":
        out["metric"] += ", reference adaptive marching rule"
        out["config"]["march"] = "adaptive: t += |step - w| (RenderUtils.cu:1704)"
    if args.fast_path:
        out["fast_path"] = fast_path(w, dev, args, ms_step)
    if prop is not None:
        out["metric"] += (" with learned proposal"
                          + (", BF16 grid" if args.storage == "bf16" else ""))
        out["config"]["proposal"] = proposal_quality(w, prop, dev, args)
    out["config"]["pretrain_steps"] = args.pretrain
    out["config"]["storage"] = args.storage
    if args.storage == "bf16" or args.workload == "uneven512":
        out["config"]["fp32_master"] = bool(args.bf16_master)
    out["config"]["record_floats"] = {"moments": w.optim.stride,
                                      "gradient": w.optim.grad_stride}
    order, wpb, pers, gmode, gwpb = w.trainer.schedule()
    out["config"]["schedule"] = {"batch_order": ["given", "morton", "length+morton"][order],
                                 "waves_per_block": wpb, "grad_waves_per_block": gwpb,
                                 "persistent": pers, "gradient": pxl.GRAD_NAMES[gmode],
                                 "autotuned": bool(args.autotune)}
    if uneven:
        out["metric"] = "train rays/sec (fwd+bwd+TV+Adam) 512^3 uneven SH2 BF16"
        out["config"]["workload"] = (
            f"uneven BF16-stored, {args.grid}^3 SH-2 caged shells + blobs at {w.occ_frac:.4f} "
            f"cell occupancy, {args.batch} rays/batch from a radius-6 sphere, step 0.5, "
            f"samples/ray {w.samples}")
        out["config"].pop("batch_order")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
