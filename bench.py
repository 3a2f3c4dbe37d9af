"""Train-step throughput of the plain method on one MI355X.

Workload (the 256^3 SH-2 configuration the metric is quoted on): 256^3
vertex grid, density from 64 seeded Gaussian blobs thresholded to ~10%
occupied cells and stored in the caged (sparse) layout; 100 views of
800x800 on the upper hemisphere of radius 4, fov 40 deg; targets rendered by
the plain method from the seeded ground-truth grid; the trained grid starts
at density 0, SH 0; 65536 rays per batch, step 0.5 voxel.

One step = one batch of Render::objectSceneLearner
(SparseVolumeRadianceField/Render.cu:352-503): fused forward + backward
(pxl_train_rays), FP64 batch loss on the GPU, the disparity TV gradient every
third batch (sparse grids, VarianceLossInterval 3), Adam, LR decay.

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
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--views", type=int, default=100)
    ap.add_argument("--img", type=int, default=800)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--occupancy", type=float, default=0.10)
    ap.add_argument("--cpu-rays", type=int, default=1024,
                    help="rays of the bounded CPU-baseline sample (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-order", choices=("shuffled", "sorted"), default="sorted",
                    help="order of the rays inside each batch (the batches themselves are "
                         "the same consecutive chunks of one seeded shuffle)")
    ap.add_argument("--log-samples", type=int, default=512,
                    help="samples per ray the two-kernel fused step logs (0: one fused kernel)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a kernel of "
                         "known traffic for calibrating profiler byte counters)")
    return ap.parse_args()


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
        t0 = time.time()
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
        self.gt_grid = pxl.Grid(N, N, N, S, gt_rec, **grid_kw)
        log(f"grid {N}^3: tau={tau:.3f}, occupied cells {self.occ_frac:.4f}, "
            f"caged vertices {n_vox} ({time.time() - t0:.1f}s)")

        # Rays of all views.
        img, V = args.img, args.views
        f, _, c, _ = scenes.view_intrinsics(img, 40.0)
        self.cams = [pxl.camera_from_pose(t, [N / 2.0] * 3, 1.0, f, f, c, c)[0]
                     for t in scenes.fibonacci_sphere(V, 4.0, upper_only=True)]
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
        self.rec = torch.zeros_like(gt_rec)
        self.grid = pxl.Grid(N, N, N, S, self.rec, **grid_kw)
        self.optim = pxl.Optim(self.grid)
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
                                   log_samples=args.log_samples)

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


def cpu_baseline(w, dev, nrays):
    """The oracle's train step on the host: a bounded sample of nrays rays
    (forward + backward) plus the per-step TV and Adam over the full caged
    grid, scaled to rays/s of a 65536-ray step."""
    import numpy as np
    import torch

    import oracle

    P = w.P
    planes = w.rec[:, :P].T.contiguous().cpu().numpy()
    occ_bits = w.grid.occupancy.cpu().numpy().view(np.uint32)
    ncell = (w.N - 1) ** 3
    occ = np.unpackbits(occ_bits.view(np.uint8), bitorder="little")[:ncell].astype(np.uint8)
    io = w.grid.index_offsets.cpu().numpy().view(np.uint32)
    keys = ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t", "gt_r", "gt_g", "gt_b",
            "gt_mask")
    idx = w.order[:nrays].long()
    sub = {k: getattr(w.rays, k)[idx].cpu().numpy() for k in keys}
    rays_o = oracle.Rays(sub, white_bg=True)
    grid_o = oracle.Grid(w.N, w.N, w.N, w.S, planes, 0.5, occupancy=occ, index_offsets=io)
    si = np.arange(nrays, dtype=np.int32)
    t0 = oracle.now()
    fwd = oracle.render_fwd(grid_o, rays_o, si)
    oracle.render_bwd(grid_o, rays_o, si, fwd, 1.0, 0.1)
    t_rays = (oracle.now() - t0) / nrays
    t0 = oracle.now()
    oracle.tv_disparity(grid_o, 1e-5, 1e-6, 0)
    t_tv = oracle.now() - t0
    p32 = planes.copy()
    g = np.zeros_like(p32)
    m = np.zeros_like(p32)
    v = np.zeros_like(p32)
    t0 = oracle.now()
    oracle.adam(p32, g, m, v, 0.3, 1.0, 1)
    t_adam = oracle.now() - t0
    B = w.args.batch
    t_step = B * t_rays + t_tv / 3.0 + t_adam
    sample = (f"{nrays} rays of the workload (FP64 forward + backward, "
              f"{t_rays * 1e3:.3f} ms/ray) + TV over the caged grid ({t_tv:.2f} s, every 3rd "
              f"step) + Adam over {w.n_vox} records ({t_adam:.2f} s); step = {B} rays")
    return {"value": B / t_step, "unit": "rays/s", "cores": 1, "kind": "port",
            "sample": sample}


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
    w = Workload(args, dev)
    for _ in range(args.warmup):
        w.step()
    torch.cuda.synchronize()

    K = args.steps
    # HIP events around each batch's fused render launch, recorded by the
    # trainer on the stream it launches on.
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
    kern_ms, nk = w.trainer.render_ms()
    assert nk == K
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
    vt = touched_vertices(w, dev)
    alg = args.batch * 44 + vt * 448
    achieved = alg / (kern_ms / 1e3) / 1e9
    step_alg = alg + w.n_vox * 896
    kname = pxl.render_kernel_name(w.grid, "logged" if args.log_samples > 0 else "fused")
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": profiled_traffic(kname, args),
            "kernel": kname, "kernel_ms": round(kern_ms, 4),
            "touched_records": int(vt), "alg_bytes": int(alg),
            "step_alg_bytes": int(step_alg),
            "step_frac": round(step_alg / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 5)}

    cpu = None
    if not args.no_cpu_baseline and args.cpu_rays > 0:
        log("CPU baseline (oracle) ...")
        cpu = cpu_baseline(w, dev, args.cpu_rays)
    out = {
        "metric": "train rays/sec (fwd+bwd+TV+Adam) 256^3 SH2",
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
                   "workload": f"plain FP32, {args.grid}^3 SH-2 caged grid at "
                               f"{w.occ_frac:.4f} cell occupancy, {args.views} views "
                               f"{args.img}x{args.img}, {args.batch} rays/batch, step 0.5",
                   "global_batch": args.batch, "caged_records": w.n_vox,
                   "train_psnr_last_batch": round(psnr, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
