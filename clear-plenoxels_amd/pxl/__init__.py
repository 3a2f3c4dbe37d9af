"""Python binding of libpxl.so, the MI355X-native Plenoxel train-step library.

Thin ctypes layer over the C ABI declared in include/pxl.h.  Device memory
comes from torch (ROCm) tensors; every call is queued on torch's current HIP
stream.  There is no fallback: if libpxl.so is missing or a call fails, an
exception is raised.

The entry points mirror the reference's operator API
(SparseVolumeRadianceField/RenderUtils.cuh:7-85, Optimizer.cuh:35-37,
SceneUtils.cuh:86-90, Rays.cuh:48-53); see each function's docstring.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_DIR, "libpxl.so")

PXL_OK, PXL_EINVAL, PXL_ENOMEM, PXL_EHIP, PXL_ENAN = 0, 1, 2, 3, 4
MARCH_FIXED, MARCH_REF_ADAPTIVE = 0, 1
TV_NEIGHBOUR, TV_DISPARITY, TV_BOTH = 0, 1, 2
OPT_MARCHER, OPT_TV_ADDR64, OPT_CLAIM_GRID = 1, 2, 3
MARCHER_GROUP, MARCHER_WAVE = 0, 1
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2

# Every symbol include/pxl.h declares.
EXPORTS = (
    "pxl_last_error", "pxl_version", "pxl_abi_sizes", "pxl_camera_from_pose", "pxl_rays_init",
    "pxl_render_fwd", "pxl_render_fwd_trace", "pxl_render_bwd_rgb", "pxl_train_rays",
    "pxl_tv_scratch_bytes", "pxl_add_variance_loss", "pxl_adam_step", "pxl_tv_adam_step",
    "pxl_build_occupancy", "pxl_sparse_index_scratch_bytes", "pxl_build_sparse_index",
    "pxl_grid_compact", "pxl_grid_import_planes", "pxl_grid_export_planes",
    "pxl_loss_reduce_f64", "pxl_count_touched", "pxl_trainer_create", "pxl_trainer_step",
    "pxl_trainer_time_render", "pxl_trainer_render_ms", "pxl_trainer_loss",
    "pxl_trainer_state", "pxl_trainer_destroy", "pxl_brick_words",
    "pxl_build_brick_occupancy", "pxl_records_to_16bit", "pxl_tv_prepare",
    "pxl_render_kernel_name", "pxl_train_rays_workspace_bytes",
    "pxl_train_rays_workspace_bytes_grid", "pxl_train_rays_ws",
    "pxl_train_rays_sched", "pxl_batch_order_scratch_bytes", "pxl_batch_order",
    "pxl_trainer_schedule", "pxl_proposal_workspace_bytes", "pxl_proposal_init",
    "pxl_proposal_features", "pxl_proposal_forward", "pxl_proposal_train_step",
    "pxl_proposal_masks", "pxl_render_fwd_masked", "pxl_train_rays_prop",
    "pxl_trainer_set_proposal", "pxl_trainer_proposal_loss", "pxl_render_image",
    "pxl_upsample_density", "pxl_upsample", "pxl_rays_init_box", "pxl_mask_bounds",
    "pxl_rays_permute", "pxl_shuffle_rank", "pxl_read_cameras", "pxl_read_image_text",
    "pxl_grid_save_text", "pxl_grid_text_records", "pxl_grid_load_text",
    "pxl_set_option", "pxl_get_option", "pxl_trainer_render_ms_split", "pxl_optim_init_band",
    "pxl_trainer_sync", "pxl_trainer_grid",
    "pxl_trainer_set_band", "pxl_trainer_learn_level", "pxl_hierarchy_plan",
    "pxl_train_rays_noalloc",
)
NOALLOC_SCRATCH_BYTES = 64 << 20

PROP_IN, PROP_WIDTH, PROP_BINS = 128, 64, 64
PROP_WEIGHTS = PROP_IN * PROP_WIDTH + 2 * PROP_WIDTH * PROP_WIDTH
PROP_PARAMS = PROP_WEIGHTS + 3 * PROP_WIDTH


class PxlError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def build():
    """Compile libpxl.so in-tree (hipcc, gfx950)."""
    import subprocess

    jobs = str(min(16, os.cpu_count() or 4))
    subprocess.run(["make", "-s", "-C", PKG_DIR, "-j" + jobs], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libpxl.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        sizes = (ctypes.c_int64 * 9)()
        L.pxl_abi_sizes(sizes, 9)
        mine = [ctypes.sizeof(c) for c in (Camera, BBox, RaysC, GridC, OptimC, TrainConfigC,
                                           LevelConfigC, LevelC, ScheduleC)]
        if list(sizes) != mine:
            raise ImportError(f"libpxl.so struct sizes {list(sizes)} != binding {mine}")
        _lib = L
    return _lib


_P = ctypes.c_void_p
_I32, _I64, _F32, _F64, _SZ = (ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double,
                               ctypes.c_size_t)


class Camera(ctypes.Structure):
    _fields_ = [("c2w", ctypes.c_double * 16), ("centre", ctypes.c_double * 4),
                ("fx", ctypes.c_double), ("fy", ctypes.c_double),
                ("cx", ctypes.c_double), ("cy", ctypes.c_double)]


class BBox(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("x_min", "x_max", "y_min", "y_max", "z_min", "z_max")]


_RAY_FLOAT = ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t")
_RAY_U8 = ("gt_r", "gt_g", "gt_b", "gt_mask")
_RAY_OUT = ("acc_r", "acc_g", "acc_b", "acc_w", "acc_density")


class RaysC(ctypes.Structure):
    _fields_ = ([(n, _P) for n in _RAY_FLOAT + _RAY_U8 + _RAY_OUT] +
                [("n", ctypes.c_int64), ("white_bg", ctypes.c_int32), ("reserved", ctypes.c_int32)])


class GridC(ctypes.Structure):
    _fields_ = [("W", _I32), ("H", _I32), ("D", _I32), ("sh_dim", _I32), ("stride", _I32),
                ("sparse", _I32), ("n_vox", _I64), ("params", _P), ("occupancy", _P),
                ("index_offsets", _P), ("in_cage", _P), ("compact_to_dense", _P),
                ("brick_occupancy", _P), ("bbox", BBox), ("step", _F32),
                ("march_mode", _I32), ("storage", _I32), ("sh_active", _I32)]


class OptimC(ctypes.Structure):
    _fields_ = [("grad", _P), ("m", _P), ("v", _P), ("adaptive_lr", _P), ("master", _P),
                ("stride", _I32), ("sh_first", _I32), ("grad_stride", _I32), ("_pad", _I32)]


class TrainConfigC(ctypes.Structure):
    _fields_ = [("batch_size", _I64), ("learning_rate", _F64), ("sh_learning_rate", _F64),
                ("variance_coefficient", _F64), ("sh_coefficient", _F64),
                ("variance_type", _I32), ("neighbour_nums", _I32),
                ("variance_interval", _I32), ("optimize_step", _I32),
                ("mask_coefficient", _F64), ("huber_delta", _F32), ("fused", _I32),
                ("log_samples", _I32), ("autotune", _I32), ("batch_order", _I32),
                ("waves_per_block", _I32), ("persistent", _I32), ("gradient", _I32),
                ("tv_inline", _I32), ("tv_separate", _I32)]


class LevelConfigC(ctypes.Structure):
    _fields_ = [(n, _I32) for n in ("epochs", "sh_band", "hierarchy_band_strategy",
                                    "sh_for_target_only", "current_hierarchy",
                                    "minimum_hierarchy", "initial_from_file", "upsampled",
                                    "available_sh_band", "check_nan")]


class LevelC(ctypes.Structure):
    _fields_ = [("level", _I32), ("epochs", _I32), ("batch_size", _I64), ("upsample", _I32),
                ("reserved", _I32)]


class ProposalC(ctypes.Structure):
    _fields_ = [("weights", _P), ("master", _P), ("m", _P), ("v", _P), ("grad", _P)]


class ScheduleC(ctypes.Structure):
    _fields_ = [("waves_per_block", _I32), ("persistent", _I32), ("gradient", _I32),
                ("grad_waves_per_block", _I32)]


# Gradient schedules of the logged step (pxl_schedule.gradient).
GRAD_PER_RAY, GRAD_CHUNKED, GRAD_BINNED = 0, 1, 2
GRAD_NAMES = {GRAD_PER_RAY: "per_ray", GRAD_CHUNKED: "chunked", GRAD_BINNED: "binned"}


def _grad_mode(chunked=False, binned=False):
    if chunked and binned:
        raise ValueError("chunked and binned are two different gradient schedules")
    return GRAD_BINNED if binned else (GRAD_CHUNKED if chunked else GRAD_PER_RAY)


def _declare(L):
    G, R, O = ctypes.POINTER(GridC), ctypes.POINTER(RaysC), ctypes.POINTER(OptimC)
    sig = {
        "pxl_last_error": ([], ctypes.c_char_p),
        "pxl_version": ([], _I32),
        "pxl_abi_sizes": ([_P, _I32], _I32),
        "pxl_camera_from_pose": ([_P, _P, _F64, _I32, _I32, _F64, _F64, _F64, _F64,
                                  ctypes.POINTER(Camera), _P], _I32),
        "pxl_rays_init": ([ctypes.POINTER(Camera), ctypes.POINTER(BBox), _I32, _I32, _P, _P, _P,
                           _P, _I64, R, _P], _I32),
        "pxl_render_fwd": ([G, R, _P, _I64, _I64, _P], _I32),
        "pxl_render_fwd_trace": ([G, R, _P, _I64, _I64, _P, _P], _I32),
        "pxl_render_bwd_rgb": ([G, O, R, _P, _P, _I64, _I64, _F64, _F32, _P], _I32),
        "pxl_train_rays": ([G, O, R, _P, _P, _I64, _I64, _F64, _F32, _P], _I32),
        "pxl_tv_scratch_bytes": ([G], _SZ),
        "pxl_add_variance_loss": ([G, O, _F64, _F64, _I32, _I32, _I32, _P, _SZ, _P], _I32),
        "pxl_adam_step": ([G, O, _F64, _F64, _I32, _P], _I32),
        "pxl_tv_adam_step": ([G, O, _F64, _F64, _I32, _P, _SZ, _F64, _F64, _I32, _P, _P], _I32),
        "pxl_build_occupancy": ([_P, _I32, _I32, _I32, _F32, _P, _P, _P, _P], _I32),
        "pxl_sparse_index_scratch_bytes": ([_I64], _SZ),
        "pxl_build_sparse_index": ([_P, _I32, _I32, _I32, _P, _P, _P, _P, _SZ, _P], _I32),
        "pxl_grid_compact": ([_P, _I32, _P, _I64, _P, _P], _I32),
        "pxl_grid_import_planes": ([_P, _I32, _I64, _P, _I32, _P], _I32),
        "pxl_grid_export_planes": ([_P, _I32, _I64, _P, _I32, _P], _I32),
        "pxl_loss_reduce_f64": ([_P, _I64, _P, _P], _I32),
        "pxl_count_touched": ([G, R, _P, _I64, _I64, _P, _P, _P], _I32),
        "pxl_trainer_create": ([ctypes.POINTER(TrainConfigC), G, O, R, _P, _I64, _P, _P, _P,
                                _SZ, _P], _P),
        "pxl_tv_prepare": ([G, _P, _SZ, _P], _I32),
        "pxl_trainer_step": ([_P], _I32),
        "pxl_trainer_sync": ([_P], _I32),
        "pxl_trainer_grid": ([_P, G], _I32),
        "pxl_trainer_time_render": ([_P, _I32], _I32),
        "pxl_trainer_render_ms": ([_P, ctypes.POINTER(_F64), ctypes.POINTER(_I64)], _I32),
        "pxl_trainer_loss": ([_P, ctypes.POINTER(_F64), ctypes.POINTER(_F64)], _I32),
        "pxl_trainer_state": ([_P, ctypes.POINTER(_I64), ctypes.POINTER(_F64),
                               ctypes.POINTER(_F64)], _I32),
        "pxl_trainer_destroy": ([_P], None),
        "pxl_brick_words": ([_I32, _I32, _I32], _I64),
        "pxl_build_brick_occupancy": ([_P, _I32, _I32, _I32, _P, _P], _I32),
        "pxl_records_to_16bit": ([_P, _I32, _I64, _P, _I32, _I32, _P], _I32),
        "pxl_render_kernel_name": ([G, _I32], ctypes.c_char_p),
        "pxl_train_rays_workspace_bytes": ([_I64, _I32], _SZ),
        "pxl_train_rays_workspace_bytes_grid": ([G, _I64, _I32], _SZ),
        "pxl_train_rays_ws": ([G, O, R, _P, _P, _I64, _I64, _F64, _F32, _P, _SZ, _I32, _P],
                              _I32),
        "pxl_train_rays_sched": ([G, O, R, _P, _P, _I64, _I64, _F64, _F32, _P, _SZ, _I32,
                                  ctypes.POINTER(ScheduleC), _P], _I32),
        "pxl_batch_order_scratch_bytes": ([_I64], _SZ),
        "pxl_batch_order": ([G, R, _P, _I64, _I64, _I32, _P, _P, _SZ, _P], _I32),
        "pxl_trainer_schedule": ([_P, ctypes.POINTER(_I32), ctypes.POINTER(ScheduleC)], _I32),
        "pxl_proposal_workspace_bytes": ([_I64], _SZ),
        "pxl_proposal_init": ([ctypes.POINTER(ProposalC), ctypes.c_uint64, _P], _I32),
        "pxl_proposal_features": ([G, R, _P, _I64, _I64, _P, _SZ, _P], _I32),
        "pxl_proposal_forward": ([ctypes.POINTER(ProposalC), _I64, _P, _P, _SZ, _P], _I32),
        "pxl_proposal_train_step": ([ctypes.POINTER(ProposalC), _I64, _P, _P, _I64, _F32,
                                     _F64, _I32, _P, _SZ, _P, _P], _I32),
        "pxl_proposal_masks": ([_P, _I64, _F32, _I32, _P, _P], _I32),
        "pxl_render_fwd_masked": ([G, R, _P, _I64, _I64, _P, _P, _P], _I32),
        "pxl_train_rays_prop": ([G, O, R, _P, _P, _I64, _I64, _F64, _F32, _P, _SZ, _I32,
                                 ctypes.POINTER(ScheduleC), _P, _P, _P, _F64, _P], _I32),
        "pxl_trainer_set_proposal": ([_P, ctypes.POINTER(ProposalC), _I32, _I32, _F32, _I32,
                                      _F64, _F32, _F64], _I32),
        "pxl_trainer_proposal_loss": ([_P, ctypes.POINTER(_F64)], _I32),
        "pxl_set_option": ([_I32, _I64], _I32),
        "pxl_train_rays_noalloc": ([G, O, R, _P, _P, _I64, _I64, _F64, _F32, _I32,
                                    ctypes.POINTER(ScheduleC), _P], _I32),
        "pxl_get_option": ([_I32, ctypes.POINTER(_I64)], _I32),
        "pxl_trainer_render_ms_split": ([_P, ctypes.POINTER(_F64), ctypes.POINTER(_I64),
                                         ctypes.POINTER(_F64), ctypes.POINTER(_I64)], _I32),
        "pxl_optim_init_band": ([G, O, _I32, _I32, _I32, _P], _I32),
        "pxl_trainer_set_band": ([_P, _I32, _I32, _I32], _I32),
        "pxl_trainer_learn_level": ([_P, ctypes.POINTER(LevelConfigC)], _I32),
        "pxl_hierarchy_plan": ([_I32, _I32, _I32, _I64, _I32, _I32, _I32, ctypes.POINTER(LevelC),
                                _I32, ctypes.POINTER(_I32)], _I32),
        "pxl_render_image": ([G, ctypes.POINTER(Camera), _I32, _I32, _I32, _P, _P, _P, _P, _P],
                             _I32),
        "pxl_upsample_density": ([G, _I32, _I32, _I32, _F32, _P, _P], _I32),
        "pxl_upsample": ([G, G, _F32, _I32, _P], _I32),
        "pxl_rays_init_box": ([ctypes.POINTER(Camera), ctypes.POINTER(BBox), _I32, _I32, _P, _P,
                               _P, _P, _P, _I64, R, _P], _I32),
        "pxl_mask_bounds": ([_P, _I32, _I32, _P, _P], _I32),
        "pxl_rays_permute": ([R, R, _P, _I64, _P], _I32),
        "pxl_shuffle_rank": ([_P, _I32, _I32, _I32, _I32], _I32),
        "pxl_read_cameras": ([ctypes.c_char_p, _P, _I32, _I32, _I32, _F64, _I32, _I32,
                              ctypes.POINTER(Camera), _I32, ctypes.POINTER(_I32)], _I32),
        "pxl_read_image_text": ([ctypes.c_char_p, _I32, _I32, _I32, _P, _P, _P, _P, _P], _I32),
        "pxl_grid_save_text": ([G, ctypes.c_char_p, _I32, _I32, _I32, _P], _I32),
        "pxl_grid_text_records": ([ctypes.c_char_p, _I32, _I32, _I32, _I32, _I32,
                                   ctypes.POINTER(_I64)], _I32),
        "pxl_grid_load_text": ([ctypes.c_char_p, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32,
                                _I64, _P, _P, _P, _P, _P, _P], _I32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def _check(fn, code):
    if code != PXL_OK:
        raise PxlError(fn, code, lib().pxl_last_error().decode())


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


# ---------------------------------------------------------------- data ---


class Rays:
    """Ray set in device memory (SoA, Rays.cuh:7-56)."""

    def __init__(self, n, device="cuda", white_bg=True):
        import torch

        self.n = int(n)
        self.white_bg = bool(white_bg)
        for k in _RAY_FLOAT + _RAY_OUT:
            setattr(self, k, torch.zeros(self.n, dtype=torch.float32, device=device))
        for k in _RAY_U8:
            setattr(self, k, torch.zeros(self.n, dtype=torch.uint8, device=device))

    @classmethod
    def from_arrays(cls, arrays, device="cuda", white_bg=True):
        import torch

        n = len(arrays["ox"])
        r = cls(n, device=device, white_bg=white_bg)
        for k, v in arrays.items():
            if hasattr(r, k):
                getattr(r, k).copy_(torch.as_tensor(v).to(getattr(r, k).dtype))
        return r

    def c(self):
        s = RaysC()
        for k in _RAY_FLOAT + _RAY_U8 + _RAY_OUT:
            setattr(s, k, getattr(self, k).data_ptr())
        s.n = self.n
        s.white_bg = int(self.white_bg)
        return s


def record_stride(S, storage=STORE_F32):
    """Elements per record: 32 (128 B FP32 / 64 B 16-bit) up to 32
    parameters, else rounded up to 4 (FP32) or 8 (16-bit) elements."""
    P = 1 + 3 * S
    if P <= 32:
        return 32
    a = 4 if storage == STORE_F32 else 8
    return ((P + a - 1) // a) * a


class Grid:
    """Voxel grid in device memory: records [n_vox, stride] float32 plus the
    caged-layout index arrays when sparse (SceneUtils.cuh:4-98)."""

    def __init__(self, W, H, D, S, params, step=0.5, bbox=None, march_mode=MARCH_FIXED,
                 occupancy=None, index_offsets=None, in_cage=None, compact_to_dense=None,
                 brick_occupancy=None, storage=STORE_F32, sh_active=0):
        self.W, self.H, self.D, self.S = int(W), int(H), int(D), int(S)
        self.params = params
        self.stride = int(params.shape[1])
        self.n_vox = int(params.shape[0])
        self.step = float(step)
        self.bbox = tuple(bbox) if bbox is not None else (0, W - 1, 0, H - 1, 0, D - 1)
        self.march_mode = int(march_mode)
        self.occupancy = occupancy
        self.index_offsets = index_offsets
        self.in_cage = in_cage
        self.compact_to_dense = compact_to_dense
        self.brick_occupancy = brick_occupancy
        self.storage = int(storage)
        self.sh_active = int(sh_active)
        if self.storage != STORE_F32:
            assert params.element_size() == 2, "16-bit storage needs a 2-byte tensor"
        self.sparse = occupancy is not None

    @property
    def P(self):
        return 1 + 3 * self.S

    def c(self):
        g = GridC()
        g.W, g.H, g.D, g.sh_dim, g.stride = self.W, self.H, self.D, self.S, self.stride
        g.sparse = int(self.sparse)
        g.n_vox = self.n_vox
        g.params = self.params.data_ptr()
        g.occupancy = None if self.occupancy is None else self.occupancy.data_ptr()
        g.index_offsets = None if self.index_offsets is None else self.index_offsets.data_ptr()
        g.in_cage = None if self.in_cage is None else self.in_cage.data_ptr()
        g.compact_to_dense = (None if self.compact_to_dense is None
                              else self.compact_to_dense.data_ptr())
        g.brick_occupancy = (None if self.brick_occupancy is None
                             else self.brick_occupancy.data_ptr())
        g.bbox = BBox(*self.bbox)
        g.step = self.step
        g.march_mode = self.march_mode
        g.storage = self.storage
        g.sh_active = self.sh_active
        return g


class Optim:
    """Gradient and Adam moments (FP32 records of `stride` floats: by
    default the grid's stride, or 32 for a 16-bit grid; any multiple of 4
    from P up, e.g. 28 for S = 9, which streams no padding).  For a 16-bit
    grid, `master` (FP32 parameters that Adam updates) is created from the
    given FP32 records when master_from is passed.  The gradient may use its
    own record stride (`grad_stride`, a multiple of 4 from P up; default
    `stride`), e.g. 32 so that every gradient record is one 128-byte line
    while the moments stream 28 floats."""

    def __init__(self, grid, adaptive_lr=None, master_from=None, sh_first=0, stride=None,
                 grad_stride=None):
        import torch

        if stride is None:
            stride = grid.stride if grid.storage == STORE_F32 else record_stride(grid.S)
        if grad_stride is None:
            grad_stride = stride
        shape = (grid.n_vox, stride)
        dev = grid.params.device
        self.stride = stride
        self.grad_stride = grad_stride
        self.grad = torch.zeros((grid.n_vox, grad_stride), dtype=torch.float32, device=dev)
        self.m = torch.zeros(shape, dtype=torch.float32, device=dev)
        self.v = torch.zeros(shape, dtype=torch.float32, device=dev)
        self.adaptive_lr = adaptive_lr
        self.sh_first = int(sh_first)
        self.master = None
        if master_from is not None:
            self.master = torch.zeros(shape, dtype=torch.float32, device=dev)
            self.master[:, :master_from.shape[1]] = master_from

    def c(self):
        o = OptimC()
        o.grad = self.grad.data_ptr()
        o.m = self.m.data_ptr()
        o.v = self.v.data_ptr()
        o.adaptive_lr = None if self.adaptive_lr is None else self.adaptive_lr.data_ptr()
        o.master = None if self.master is None else self.master.data_ptr()
        o.stride = self.stride
        o.sh_first = self.sh_first
        o.grad_stride = self.grad_stride
        return o


# ---------------------------------------------------------- entry points --


def camera_from_pose(translation, object_centre, camera_distance, fx, fy, cx, cy,
                     need_rescale=True, z_reversed=True):
    """Camera of one pose (Camera::readFromMyBlender per view,
    CameraUtils.cu:339-377).  Returns (Camera, w2c list of 16)."""
    t = (ctypes.c_double * 3)(*translation)
    oc = (ctypes.c_double * 3)(*object_centre)
    cam = Camera()
    w2c = (ctypes.c_double * 16)()
    _check("pxl_camera_from_pose",
           lib().pxl_camera_from_pose(t, oc, camera_distance, int(need_rescale), int(z_reversed),
                                      fx, fy, cx, cy, ctypes.byref(cam), w2c))
    return cam, list(w2c)


def rays_init(cam, bbox, img_w, img_h, rays, start_index=0, img=None, stream=None):
    """Rays of one view (_rays::initialSelf / initialRays, Rays.cu:77-240).
    img = (r, g, b, mask) uint8 device tensors of img_h*img_w or None."""
    ims = [None] * 4 if img is None else [_ptr(x) for x in img]
    rc = rays.c()
    _check("pxl_rays_init",
           lib().pxl_rays_init(ctypes.byref(cam), ctypes.byref(BBox(*bbox)), img_w, img_h,
                               *ims, start_index, ctypes.byref(rc), _stream(stream)))


def rays_init_box(cam, bbox, img_w, img_h, pixel_box, rays, start_index=0, img=None,
                  stream=None):
    """Mask-area rays of one view (MaskAreaOnly, Rays.cu:86-92): the pixels of
    pixel_box = (row_min, row_max, col_min, col_max), inclusive, row-major."""
    ims = [None] * 4 if img is None else [_ptr(x) for x in img]
    box = (ctypes.c_int32 * 4)(*[int(v) for v in pixel_box])
    rc = rays.c()
    _check("pxl_rays_init_box",
           lib().pxl_rays_init_box(ctypes.byref(cam), ctypes.byref(BBox(*bbox)), img_w, img_h,
                                   box, *ims, start_index, ctypes.byref(rc), _stream(stream)))


def mask_bounds(mask, img_w, img_h, stream=None):
    """readMask's bounds of a device mask: int32 device tensor (row min, row
    max, col min, col max)."""
    import torch

    out = torch.empty(4, dtype=torch.int32, device=mask.device)
    _check("pxl_mask_bounds",
           lib().pxl_mask_bounds(_ptr(mask), img_w, img_h, _ptr(out), _stream(stream)))
    return out


def rays_permute(src, dst, perm, n=None, stream=None):
    """_rays::shuffleTo (Rays.cu:271-287): dst ray j = src ray perm[j]."""
    n = perm.numel() if n is None else n
    s, d = src.c(), dst.c()
    _check("pxl_rays_permute",
           lib().pxl_rays_permute(ctypes.byref(s), ctypes.byref(d), _ptr(perm), n,
                                  _stream(stream)))


def shuffle_rank(dest, max_index, regenerate=True, min_index=0, shuffle=True):
    """ShuffleRank (Utils.cu:488-506) into a host int32 numpy array, drawing
    from the C library's rand()."""
    import numpy as np

    assert dest.dtype == np.int32 and dest.flags["C_CONTIGUOUS"] and len(dest) >= max_index
    _check("pxl_shuffle_rank",
           lib().pxl_shuffle_rank(dest.ctypes.data_as(ctypes.c_void_p), max_index,
                                  int(regenerate), min_index, int(shuffle)))
    return dest


def render_fwd(grid, rays, sample_index, start=0, length=None, trace=None, stream=None):
    """RenderUtils::renderRays (RenderUtils.cu:1961-1971)."""
    length = sample_index.numel() - start if length is None else length
    g, r = grid.c(), rays.c()
    if trace is None:
        _check("pxl_render_fwd", lib().pxl_render_fwd(ctypes.byref(g), ctypes.byref(r),
                                                      _ptr(sample_index), start, length,
                                                      _stream(stream)))
    else:
        _check("pxl_render_fwd_trace",
               lib().pxl_render_fwd_trace(ctypes.byref(g), ctypes.byref(r), _ptr(sample_index),
                                          start, length, _ptr(trace), _stream(stream)))


def render_bwd_rgb(grid, optim, rays, losses, sample_index, start=0, length=None,
                   mask_coef=1.0, huber_delta=0.1, stream=None):
    """RenderUtils::renderRays_withGradients_RGB (RenderUtils.cu:2258-2278)."""
    length = sample_index.numel() - start if length is None else length
    g, o, r = grid.c(), optim.c(), rays.c()
    _check("pxl_render_bwd_rgb",
           lib().pxl_render_bwd_rgb(ctypes.byref(g), ctypes.byref(o), ctypes.byref(r),
                                    _ptr(losses), _ptr(sample_index), start, length, mask_coef,
                                    huber_delta, _stream(stream)))


def train_rays(grid, optim, rays, losses, sample_index, start=0, length=None, mask_coef=1.0,
               huber_delta=0.1, stream=None):
    """Fused renderRays + renderRays_withGradients_RGB in one launch."""
    length = sample_index.numel() - start if length is None else length
    g, o, r = grid.c(), optim.c(), rays.c()
    _check("pxl_train_rays",
           lib().pxl_train_rays(ctypes.byref(g), ctypes.byref(o), ctypes.byref(r), _ptr(losses),
                                _ptr(sample_index), start, length, mask_coef, huber_delta,
                                _stream(stream)))


def train_workspace(length, max_samples, device="cuda", grid=None):
    """Device workspace for train_rays_ws over `length` rays logging up to
    `max_samples` samples per ray; with `grid`, also for the brick-binned
    gradient schedule on that grid."""
    import torch

    if grid is None:
        nbytes = lib().pxl_train_rays_workspace_bytes(length, max_samples)
    else:
        g = grid.c()
        nbytes = lib().pxl_train_rays_workspace_bytes_grid(ctypes.byref(g), length, max_samples)
    if nbytes == 0:
        raise PxlError("pxl_train_rays_workspace_bytes", PXL_EINVAL,
                       f"bad length {length} / max_samples {max_samples}")
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def train_rays_ws(grid, optim, rays, losses, sample_index, workspace, max_samples, start=0,
                  length=None, mask_coef=1.0, huber_delta=0.1, stream=None):
    """train_rays as two kernels with a sample log in `workspace`
    (train_workspace(length, max_samples))."""
    length = sample_index.numel() - start if length is None else length
    g, o, r = grid.c(), optim.c(), rays.c()
    _check("pxl_train_rays_ws",
           lib().pxl_train_rays_ws(ctypes.byref(g), ctypes.byref(o), ctypes.byref(r),
                                   _ptr(losses), _ptr(sample_index), start, length, mask_coef,
                                   huber_delta, _ptr(workspace), workspace.numel(), max_samples,
                                   _stream(stream)))


ORDER_NONE, ORDER_MORTON, ORDER_LENGTH_MORTON = 0, 1, 2


def train_rays_sched(grid, optim, rays, losses, sample_index, workspace, max_samples,
                     waves_per_block=4, persistent=False, chunked=False, grad_waves_per_block=0,
                     start=0, length=None, mask_coef=1.0, huber_delta=0.1, stream=None,
                     binned=False):
    """train_rays_ws with an explicit launch schedule (chunked: one gradient
    wave per 64-sample chunk of each ray's log; binned: the brick-binned
    gradient, workspace from train_workspace(..., grid=grid);
    grad_waves_per_block: the per-ray gradient kernels' waves per workgroup,
    0 = waves_per_block)."""
    length = sample_index.numel() - start if length is None else length
    g, o, r = grid.c(), optim.c(), rays.c()
    sc = ScheduleC(waves_per_block, int(persistent), _grad_mode(chunked, binned),
                   int(grad_waves_per_block))
    _check("pxl_train_rays_sched",
           lib().pxl_train_rays_sched(ctypes.byref(g), ctypes.byref(o), ctypes.byref(r),
                                      _ptr(losses), _ptr(sample_index), start, length, mask_coef,
                                      huber_delta, _ptr(workspace), workspace.numel(),
                                      max_samples, ctypes.byref(sc), _stream(stream)))


def batch_order(grid, rays, sample_index, out, mode=ORDER_MORTON, start=0, length=None,
                scratch=None, stream=None):
    """Ray ids sample_index[start:start+length] in `mode` order, into out[:length]."""
    import torch

    length = sample_index.numel() - start if length is None else length
    nbytes = lib().pxl_batch_order_scratch_bytes(length)
    if scratch is None:
        scratch = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=sample_index.device)
    g, r = grid.c(), rays.c()
    _check("pxl_batch_order",
           lib().pxl_batch_order(ctypes.byref(g), ctypes.byref(r), _ptr(sample_index), start,
                                 length, mode, _ptr(out), _ptr(scratch), scratch.numel(),
                                 _stream(stream)))
    return out


class Proposal:
    """Learned proposal network (BF16 weights, FP32 master and Adam moments)
    in device memory, initialised from `seed`."""

    def __init__(self, device="cuda", seed=0, stream=None):
        import torch

        self.weights = torch.zeros(PROP_WEIGHTS, dtype=torch.int16, device=device)
        self.master = torch.zeros(PROP_PARAMS, dtype=torch.float32, device=device)
        self.m = torch.zeros_like(self.master)
        self.v = torch.zeros_like(self.master)
        self.grad = torch.zeros_like(self.master)
        c = self.c()
        _check("pxl_proposal_init", lib().pxl_proposal_init(ctypes.byref(c), seed,
                                                            _stream(stream)))

    def c(self):
        return ProposalC(self.weights.data_ptr(), self.master.data_ptr(), self.m.data_ptr(),
                         self.v.data_ptr(), self.grad.data_ptr())


def proposal_workspace(length, device="cuda"):
    import torch

    return torch.empty(lib().pxl_proposal_workspace_bytes(length), dtype=torch.uint8,
                       device=device)


def proposal_features(grid, rays, sample_index, workspace, start=0, length=None, stream=None):
    length = sample_index.numel() - start if length is None else length
    g, r = grid.c(), rays.c()
    _check("pxl_proposal_features",
           lib().pxl_proposal_features(ctypes.byref(g), ctypes.byref(r), _ptr(sample_index),
                                       start, length, _ptr(workspace), workspace.numel(),
                                       _stream(stream)))


def proposal_forward(prop, length, logits, workspace, stream=None):
    c = prop.c()
    _check("pxl_proposal_forward",
           lib().pxl_proposal_forward(ctypes.byref(c), length, _ptr(logits), _ptr(workspace),
                                      workspace.numel(), _stream(stream)))


def proposal_train_step(prop, length, targets, workspace, t, lr=1e-3, valid=None, n_valid=None,
                        pos_weight=1.0, loss_out=None, stream=None):
    c = prop.c()
    n_valid = length if n_valid is None else n_valid
    _check("pxl_proposal_train_step",
           lib().pxl_proposal_train_step(ctypes.byref(c), length, _ptr(targets), _ptr(valid),
                                         n_valid, pos_weight, lr, t, _ptr(workspace),
                                         workspace.numel(), _ptr(loss_out), _stream(stream)))


def proposal_masks(logits, length, masks, threshold=0.05, dilate=1, stream=None):
    _check("pxl_proposal_masks",
           lib().pxl_proposal_masks(_ptr(logits), length, threshold, dilate, _ptr(masks),
                                    _stream(stream)))


def render_image(grid, cam, width, height, white_bg=True, stream=None, device="cuda"):
    """Eval render of one image (RenderUtils::renderImage): dict of device
    planes rgb [3, h*w] float32, rgb8 [3, h*w] uint8, mask, disparity."""
    import torch

    n = width * height
    out = {"rgb": torch.zeros((3, n), dtype=torch.float32, device=device),
           "rgb8": torch.zeros((3, n), dtype=torch.uint8, device=device),
           "mask": torch.zeros(n, dtype=torch.uint8, device=device),
           "disparity": torch.zeros(n, dtype=torch.float32, device=device)}
    g = grid.c()
    _check("pxl_render_image",
           lib().pxl_render_image(ctypes.byref(g), ctypes.byref(cam), width, height,
                                  int(white_bg), _ptr(out["rgb"]), _ptr(out["rgb8"]),
                                  _ptr(out["mask"]), _ptr(out["disparity"]), _stream(stream)))
    return out


def render_fwd_masked(grid, rays, sample_index, bin_mask=None, start=0, length=None,
                      trace=None, stream=None):
    length = sample_index.numel() - start if length is None else length
    g, r = grid.c(), rays.c()
    _check("pxl_render_fwd_masked",
           lib().pxl_render_fwd_masked(ctypes.byref(g), ctypes.byref(r), _ptr(sample_index),
                                       start, length, _ptr(bin_mask), _ptr(trace),
                                       _stream(stream)))


def train_rays_prop(grid, optim, rays, losses, sample_index, workspace, max_samples,
                    bin_mask=None, bin_targets=None, target_valid=None, target_wmin=1e-4,
                    waves_per_block=4, persistent=False, chunked=False, start=0, length=None,
                    mask_coef=1.0, huber_delta=0.1, stream=None, binned=False):
    """The two-kernel step with an optional proposal mask and target output
    (batches with targets run the chunked schedule's gradient one wave per
    ray)."""
    length = sample_index.numel() - start if length is None else length
    g, o, r = grid.c(), optim.c(), rays.c()
    sc = ScheduleC(waves_per_block, int(persistent), _grad_mode(chunked, binned), 0)
    _check("pxl_train_rays_prop",
           lib().pxl_train_rays_prop(ctypes.byref(g), ctypes.byref(o), ctypes.byref(r),
                                     _ptr(losses), _ptr(sample_index), start, length, mask_coef,
                                     huber_delta, _ptr(workspace), workspace.numel(), max_samples,
                                     ctypes.byref(sc), _ptr(bin_mask), _ptr(bin_targets),
                                     _ptr(target_valid), target_wmin, _stream(stream)))


def tv_scratch(grid, stream=None):
    """The grid's prepared variance-loss scratch (sparse grids; created and
    filled by pxl_tv_prepare on first use, kept on the Grid object)."""
    import torch

    if not grid.sparse:
        return None
    if getattr(grid, "_tv_scratch", None) is None:
        g = grid.c()
        nbytes = lib().pxl_tv_scratch_bytes(ctypes.byref(g))
        buf = torch.zeros(nbytes // 8 + 64, dtype=torch.float64, device=grid.params.device)
        _check("pxl_tv_prepare", lib().pxl_tv_prepare(ctypes.byref(g), _ptr(buf), nbytes,
                                                      _stream(stream)))
        grid._tv_scratch = (buf, nbytes)
    return grid._tv_scratch


def add_variance_loss(grid, optim, lambda_density, lambda_sh, tv_type=TV_DISPARITY,
                      neighbours=6, optimize_step=0, stream=None):
    """RenderUtils::addVarianceLoss (RenderUtils.cu:2815-2851)."""
    g, o = grid.c(), optim.c()
    sc = tv_scratch(grid, stream) if tv_type in (TV_DISPARITY, TV_BOTH) else None
    buf, nbytes = (None, 0) if sc is None else sc
    _check("pxl_add_variance_loss",
           lib().pxl_add_variance_loss(ctypes.byref(g), ctypes.byref(o), lambda_density,
                                       lambda_sh, tv_type, neighbours, optimize_step,
                                       _ptr(buf), nbytes, _stream(stream)))


def set_option(key, value):
    """Library option (OPT_MARCHER: MARCHER_GROUP / MARCHER_WAVE;
    OPT_TV_ADDR64: 0 / 1; OPT_CLAIM_GRID: 0 resident, else the largest grid
    of the launches whose waves claim batch entries)."""
    _check("pxl_set_option", lib().pxl_set_option(int(key), int(value)))


def get_option(key):
    v = _I64()
    _check("pxl_get_option", lib().pxl_get_option(int(key), ctypes.byref(v)))
    return v.value


def optim_init_band(grid, optim, band, band_mode=True, density_initialized=True, stream=None):
    """Optimiser state at the start of a band (optimizer::initialBand /
    initialAll, Optimizer.cu:11-64, 139-187)."""
    g, o = grid.c(), optim.c()
    _check("pxl_optim_init_band",
           lib().pxl_optim_init_band(ctypes.byref(g), ctypes.byref(o), band, int(band_mode),
                                     int(density_initialized), _stream(stream)))


def hierarchy_plan(hierarchy, minimum_hierarchy, epochs, batch_size, shuffle_rays=True,
                   fix_batch_size=False, to_sparse_hierarchy=0):
    """Levels of the coarse-to-fine loop (VolumeFieldRender.cu:456-577), by
    the C++ host layer: list of dicts level, epochs, batch_size, upsample
    (None after the last level, else {"to_sparse": bool})."""
    n = _I32()
    args = (int(hierarchy), int(minimum_hierarchy), int(epochs), int(batch_size),
            int(shuffle_rays), int(fix_batch_size), int(to_sparse_hierarchy))
    _check("pxl_hierarchy_plan", lib().pxl_hierarchy_plan(*args, None, 0, ctypes.byref(n)))
    arr = (LevelC * max(1, n.value))()
    _check("pxl_hierarchy_plan", lib().pxl_hierarchy_plan(*args, arr, n.value, ctypes.byref(n)))
    out = []
    for L in arr[:n.value]:
        up = None if L.upsample == 0 else {"to_sparse": L.upsample == 2}
        out.append({"level": L.level, "epochs": L.epochs, "batch_size": L.batch_size,
                    "upsample": up})
    return out


def tv_adam_step(grid, optim, lambda_density, lambda_sh, lr, sh_lr, t, params_out,
                 optimize_step=0, stream=None):
    """addVarianceLossGPU (type 1) and updateRGBSceneGPU_ADAM in one pass
    (RenderUtils.cu:2469-2547, Optimizer.cu:359-415): the stepped records go
    to params_out (a tensor shaped and typed as grid.params); grid.params is
    left as it was."""
    assert params_out.shape == grid.params.shape and params_out.dtype == grid.params.dtype
    g, o = grid.c(), optim.c()
    buf, nbytes = tv_scratch(grid, stream) if grid.sparse else (None, 0)
    _check("pxl_tv_adam_step",
           lib().pxl_tv_adam_step(ctypes.byref(g), ctypes.byref(o), lambda_density, lambda_sh,
                                  optimize_step, _ptr(buf), nbytes, lr, sh_lr, t,
                                  _ptr(params_out), _stream(stream)))


def adam_step(grid, optim, lr, sh_lr, t, stream=None):
    """Optimizer::updateRGBScene (Optimizer.cu:451-479)."""
    g, o = grid.c(), optim.c()
    _check("pxl_adam_step", lib().pxl_adam_step(ctypes.byref(g), ctypes.byref(o), lr, sh_lr, t,
                                                _stream(stream)))


def build_occupancy(density, W, H, D, threshold, occupancy, in_cage, n_occupied=None,
                    stream=None):
    """objectScene::pruneSelf (SceneUtils.cu:883-913)."""
    _check("pxl_build_occupancy",
           lib().pxl_build_occupancy(_ptr(density), W, H, D, threshold, _ptr(occupancy),
                                     _ptr(in_cage), _ptr(n_occupied), _stream(stream)))


def build_brick_occupancy(occupancy, W, H, D, stream=None):
    """Brick map (one bit per 8^3 cells) of a cell bitfield; returns it."""
    import torch

    words = lib().pxl_brick_words(W, H, D)
    out = torch.zeros(max(1, words), dtype=torch.int32, device=occupancy.device)
    _check("pxl_build_brick_occupancy",
           lib().pxl_build_brick_occupancy(_ptr(occupancy), W, H, D, _ptr(out), _stream(stream)))
    return out


def sparse_index_scratch_bytes(n_vertices):
    return lib().pxl_sparse_index_scratch_bytes(n_vertices)


def build_sparse_index(in_cage, W, H, D, index_offsets, compact_to_dense, in_cage_nums,
                       scratch, stream=None):
    """objectScene::getCagedNums (SceneUtils.cu:1081-1120)."""
    nbytes = 0 if scratch is None else scratch.numel() * scratch.element_size()
    _check("pxl_build_sparse_index",
           lib().pxl_build_sparse_index(_ptr(in_cage), W, H, D, _ptr(index_offsets),
                                        _ptr(compact_to_dense), _ptr(in_cage_nums),
                                        _ptr(scratch), nbytes, _stream(stream)))


def grid_compact(dense_records, compact_to_dense, n_vox, out, stream=None):
    _check("pxl_grid_compact",
           lib().pxl_grid_compact(_ptr(dense_records), dense_records.shape[1],
                                  _ptr(compact_to_dense), n_vox, _ptr(out), _stream(stream)))


def grid_import_planes(planes, records, stream=None):
    """planes: list of P device float32 tensors [n_vox] (reference plane table)."""
    arr = (ctypes.c_void_p * len(planes))(*[p.data_ptr() for p in planes])
    _check("pxl_grid_import_planes",
           lib().pxl_grid_import_planes(arr, len(planes), records.shape[0], _ptr(records),
                                        records.shape[1], _stream(stream)))


def grid_export_planes(records, planes, stream=None):
    arr = (ctypes.c_void_p * len(planes))(*[p.data_ptr() for p in planes])
    _check("pxl_grid_export_planes",
           lib().pxl_grid_export_planes(_ptr(records), records.shape[1], records.shape[0], arr,
                                        len(planes), _stream(stream)))


def records_to_16bit(rec32, storage, stride16=None, stream=None):
    """FP32 records -> new 16-bit records (round to nearest even)."""
    import torch

    stride16 = rec32.shape[1] if stride16 is None else stride16
    out = torch.empty((rec32.shape[0], stride16), dtype=torch.int16, device=rec32.device)
    _check("pxl_records_to_16bit",
           lib().pxl_records_to_16bit(_ptr(rec32), rec32.shape[1], rec32.shape[0], _ptr(out),
                                      stride16, storage, _stream(stream)))
    return out


def loss_reduce(losses, out, length=None, stream=None):
    """FP64 MSE and PSNR of a batch (getMSEError, Utils.cu:311-320)."""
    length = losses.numel() if length is None else length
    _check("pxl_loss_reduce_f64",
           lib().pxl_loss_reduce_f64(_ptr(losses), length, _ptr(out), _stream(stream)))


def render_kernel_name(grid, mode="fused"):
    """Name of the kernel the render entry points launch for `grid`
    (mode "fwd", "bwd", "fused" or "logged" for train_rays_ws)."""
    m = {"fwd": 0, "bwd": 1, "fused": 2, "logged": 3}[mode]
    g = grid.c()
    name = lib().pxl_render_kernel_name(ctypes.byref(g), m)
    if name is None:
        raise PxlError("pxl_render_kernel_name", PXL_EINVAL, lib().pxl_last_error().decode())
    return name.decode()


def count_touched(grid, rays, sample_index, bitmap, count, start=0, length=None, stream=None):
    length = sample_index.numel() - start if length is None else length
    g, r = grid.c(), rays.c()
    _check("pxl_count_touched",
           lib().pxl_count_touched(ctypes.byref(g), ctypes.byref(r), _ptr(sample_index), start,
                                   length, _ptr(bitmap), _ptr(count), _stream(stream)))


class Trainer:
    """Batch loop of Render::objectSceneLearner (Render.cu:352-503), run by
    the C++ host layer; all buffers stay owned by the Python objects passed
    in, which must outlive the trainer."""

    def __init__(self, grid, optim, rays, order, n_rays=None, batch_size=65536,
                 learning_rate=0.3, sh_learning_rate=1.0, variance_coefficient=1e-5,
                 sh_coefficient=1e-6, variance_type=TV_DISPARITY, neighbour_nums=6,
                 variance_interval=3, optimize_step=0, mask_coefficient=1.0, huber_delta=0.1,
                 fused=True, log_samples=0, autotune=False, batch_order=0,
                 waves_per_block=0, persistent=False, chunked=False, tv_inline=False,
                 tv_separate=False, stream=None, binned=False):
        import torch

        self._keep = (grid, optim, rays, order)
        dev = grid.params.device
        self.losses = torch.zeros(batch_size, dtype=torch.float64, device=dev)
        self.loss_out = torch.zeros(2, dtype=torch.float64, device=dev)
        cfg = TrainConfigC(batch_size, learning_rate, sh_learning_rate, variance_coefficient,
                           sh_coefficient, variance_type, neighbour_nums, variance_interval,
                           optimize_step, mask_coefficient, huber_delta, int(fused),
                           int(log_samples), int(autotune), int(batch_order),
                           int(waves_per_block), int(persistent), _grad_mode(chunked, binned),
                           int(tv_inline), int(tv_separate))
        self._g, self._o, self._r = grid.c(), optim.c(), rays.c()
        n_rays = order.numel() if n_rays is None else n_rays
        sc = tv_scratch(grid, stream) if variance_type in (TV_DISPARITY, TV_BOTH) else None
        buf, nbytes = (None, 0) if sc is None else sc
        h = lib().pxl_trainer_create(ctypes.byref(cfg), ctypes.byref(self._g),
                                     ctypes.byref(self._o), ctypes.byref(self._r), _ptr(order),
                                     n_rays, _ptr(self.losses), _ptr(self.loss_out), _ptr(buf),
                                     nbytes, _stream(stream))
        if not h:
            raise PxlError("pxl_trainer_create", PXL_EINVAL, lib().pxl_last_error().decode())
        self._h = ctypes.c_void_p(h)

    def step(self):
        _check("pxl_trainer_step", lib().pxl_trainer_step(self._h))

    def sync(self):
        """Bring the current records into grid.params (the fused TV + Adam
        pass alternates them with a second, trainer-owned buffer); call it
        before reading grid.params between steps."""
        _check("pxl_trainer_sync", lib().pxl_trainer_sync(self._h))

    def grid_c(self):
        """The trainer's grid descriptor (GridC), current records included."""
        g = GridC()
        _check("pxl_trainer_grid", lib().pxl_trainer_grid(self._h, ctypes.byref(g)))
        return g

    def time_render(self, max_steps):
        _check("pxl_trainer_time_render", lib().pxl_trainer_time_render(self._h, max_steps))

    def render_ms(self):
        avg, n = ctypes.c_double(), ctypes.c_int64()
        _check("pxl_trainer_render_ms",
               lib().pxl_trainer_render_ms(self._h, ctypes.byref(avg), ctypes.byref(n)))
        return avg.value, n.value

    def render_ms_split(self):
        """(mean over every timed batch, count, mean over the batches whose
        render ran without the variance loss beside it, count)."""
        avg, n, al, na = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
        _check("pxl_trainer_render_ms_split",
               lib().pxl_trainer_render_ms_split(self._h, ctypes.byref(avg), ctypes.byref(n),
                                                 ctypes.byref(al), ctypes.byref(na)))
        return avg.value, n.value, al.value, na.value

    def set_band(self, band, band_mode=True, density_initialized=True):
        """Switch to SH band `band` (Render.cu:282-315)."""
        _check("pxl_trainer_set_band",
               lib().pxl_trainer_set_band(self._h, band, int(band_mode), int(density_initialized)))

    def learn_level(self, epochs, sh_band, hierarchy_band_strategy=False,
                    sh_for_target_only=True, current_hierarchy=0, minimum_hierarchy=0,
                    initial_from_file=False, upsampled=False, available_sh_band=0,
                    check_nan=False):
        """One level of Render::objectSceneLearner (bands x epochs x batches),
        run by the C++ host layer."""
        lc = LevelConfigC(int(epochs), int(sh_band), int(hierarchy_band_strategy),
                          int(sh_for_target_only), int(current_hierarchy),
                          int(minimum_hierarchy), int(initial_from_file), int(upsampled),
                          int(available_sh_band), int(check_nan))
        _check("pxl_trainer_learn_level", lib().pxl_trainer_learn_level(self._h, ctypes.byref(lc)))

    def loss(self):
        mse, psnr = ctypes.c_double(), ctypes.c_double()
        _check("pxl_trainer_loss", lib().pxl_trainer_loss(self._h, ctypes.byref(mse),
                                                          ctypes.byref(psnr)))
        return mse.value, psnr.value

    def state(self):
        t, lr, sl = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        _check("pxl_trainer_state", lib().pxl_trainer_state(self._h, ctypes.byref(t),
                                                            ctypes.byref(lr), ctypes.byref(sl)))
        return t.value, lr.value, sl.value

    def set_proposal(self, prop, interval=8, prune=True, threshold=0.05, dilate=1, lr=3e-3,
                     pos_weight=4.0, target_wmin=1e-5):
        """Run the learned proposal `prop` (a pxl.Proposal kept alive by the
        caller) inside the batch loop."""
        self._prop = prop
        c = prop.c()
        _check("pxl_trainer_set_proposal",
               lib().pxl_trainer_set_proposal(self._h, ctypes.byref(c), interval, int(prune),
                                              threshold, dilate, lr, pos_weight, target_wmin))

    def proposal_loss(self):
        v = ctypes.c_double()
        _check("pxl_trainer_proposal_loss",
               lib().pxl_trainer_proposal_loss(self._h, ctypes.byref(v)))
        return v.value

    def schedule(self):
        """(batch order, waves per block, persistent, gradient schedule
        GRAD_*, gradient waves per block) the trainer runs."""
        order, sc = _I32(), ScheduleC()
        _check("pxl_trainer_schedule",
               lib().pxl_trainer_schedule(self._h, ctypes.byref(order), ctypes.byref(sc)))
        return (order.value, sc.waves_per_block, bool(sc.persistent), sc.gradient,
                sc.grad_waves_per_block or sc.waves_per_block)

    def close(self):
        if getattr(self, "_h", None):
            lib().pxl_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def upsample_density(src, W2, H2, D2, threshold, stream=None):
    """Thresholded density of every vertex of the W2 x H2 x D2 grid sampled
    from `src` (upsampleScene into the dense temporary, SceneUtils.cu:540-546);
    a float32 device tensor [W2 * H2 * D2]."""
    import torch

    out = torch.empty(W2 * H2 * D2, dtype=torch.float32, device=src.params.device)
    g = src.c()
    _check("pxl_upsample_density",
           lib().pxl_upsample_density(ctypes.byref(g), W2, H2, D2, threshold, _ptr(out),
                                      _stream(stream)))
    return out


def upsample(src, dst, threshold, low_band=False, stream=None):
    """upsampleScene (SceneUtils.cu:419-485) into the FP32 records of `dst`
    (a Grid, dense or caged)."""
    gs, gd = src.c(), dst.c()
    _check("pxl_upsample",
           lib().pxl_upsample(ctypes.byref(gs), ctypes.byref(gd), threshold, int(low_band),
                              _stream(stream)))


def upsample_grid(src, threshold, to_sparse=True, low_band=False, sh_dim=None, stride=None,
                  step=None, stream=None):
    """objectScene::initSelfByUpsample (SceneUtils.cu:487-657): a new Grid of
    twice the source's dimensions.  Caged (to_sparse): the thresholded density
    is pruned with the same threshold (pxl_build_occupancy), indexed
    (pxl_build_sparse_index), and the caged records filled by pxl_upsample;
    the brick map is built as for any caged grid."""
    import torch

    dev = src.params.device
    W2, H2, D2 = 2 * src.W, 2 * src.H, 2 * src.D
    S2 = (1 if low_band else src.S) if sh_dim is None else int(sh_dim)
    P2 = 1 + 3 * S2
    stride = stride or ((P2 + 3) // 4 * 4 if P2 > 32 else 32)
    step = src.step if step is None else step
    if not to_sparse:
        rec = torch.empty((W2 * H2 * D2, stride), dtype=torch.float32, device=dev)
        dst = Grid(W2, H2, D2, S2, rec, step=step, march_mode=src.march_mode)
        upsample(src, dst, threshold, low_band, stream)
        return dst
    dens = upsample_density(src, W2, H2, D2, threshold, stream)
    ncell = (W2 - 1) * (H2 - 1) * (D2 - 1)
    occ = torch.zeros((ncell + 31) // 32, dtype=torch.int32, device=dev)
    cage = torch.zeros(W2 * H2 * D2, dtype=torch.uint8, device=dev)
    build_occupancy(dens, W2, H2, D2, threshold, occ, cage, stream=stream)
    del dens
    io = torch.zeros(W2 * H2 * D2, dtype=torch.int32, device=dev)
    c2d = torch.zeros(W2 * H2 * D2, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    scratch = torch.zeros(sparse_index_scratch_bytes(W2 * H2 * D2) // 4 + 1, dtype=torch.int32,
                          device=dev)
    build_sparse_index(cage, W2, H2, D2, io, c2d, cnt, scratch, stream=stream)
    n = int(cnt.item())
    rec = torch.empty((n, stride), dtype=torch.float32, device=dev)
    br = build_brick_occupancy(occ, W2, H2, D2, stream=stream)
    dst = Grid(W2, H2, D2, S2, rec, step=step, march_mode=src.march_mode, occupancy=occ,
               index_offsets=io, in_cage=cage, compact_to_dense=c2d[:max(n, 1)].clone(),
               brick_occupancy=br)
    upsample(src, dst, threshold, low_band, stream)
    return dst


def read_cameras(path, object_centre, target_w, target_h, camera_distance, camera_nums=0,
                 need_rescale=True, z_reversed=True):
    """Camera::readFromMyBlender (CameraUtils.cu:271-386): list of Camera."""
    oc = (ctypes.c_double * 3)(*object_centre)
    n = _I32()
    p = str(path).encode()
    _check("pxl_read_cameras",
           lib().pxl_read_cameras(p, oc, target_w, target_h, camera_nums, camera_distance,
                                  int(need_rescale), int(z_reversed), None, 0, ctypes.byref(n)))
    cams = (Camera * max(1, n.value))()
    _check("pxl_read_cameras",
           lib().pxl_read_cameras(p, oc, target_w, target_h, camera_nums, camera_distance,
                                  int(need_rescale), int(z_reversed), cams, n.value,
                                  ctypes.byref(n)))
    return [cams[i] for i in range(n.value)]


def read_image_text(prefix, width, height, gray_only=False):
    """Image::initSelfFromFile + readMask: dict of host uint8 planes r, g, b,
    mask ([height * width]) and the mask bounds."""
    import numpy as np

    n = width * height
    planes = {k: np.zeros(n, np.uint8) for k in ("r", "g", "b", "mask")}
    bounds = (ctypes.c_int32 * 4)()
    ptr = {k: v.ctypes.data_as(ctypes.c_void_p) for k, v in planes.items()}
    _check("pxl_read_image_text",
           lib().pxl_read_image_text(str(prefix).encode(), width, height, int(gray_only),
                                     ptr["r"], ptr["g"], ptr["b"], ptr["mask"], bounds))
    planes["bounds"] = list(bounds)
    return planes


def grid_save_text(grid, directory, band, density_only=False, hierarchy_mode=False,
                   stream=None):
    """objectScene::saveSelf (SceneUtils.cu:1246-1299)."""
    g = grid.c()
    _check("pxl_grid_save_text",
           lib().pxl_grid_save_text(ctypes.byref(g), str(directory).encode(), band,
                                    int(density_only), int(hierarchy_mode), _stream(stream)))


def grid_load_text(directory, W, H, D, band, sparse, sh_dim=None, have_occupancy=True,
                   step=0.5, device="cuda", stream=None):
    """objectScene::initFromFile (SceneUtils.cu:703-790): a new FP32 Grid."""
    import torch

    S = (band + 1) ** 2 if sh_dim is None else int(sh_dim)
    stride = record_stride(S)
    d = str(directory).encode()
    n = _I64()
    _check("pxl_grid_text_records",
           lib().pxl_grid_text_records(d, band, int(sparse), W, H, D, ctypes.byref(n)))
    nrec = n.value
    ncell = (W - 1) * (H - 1) * (D - 1)
    rec = torch.zeros((nrec, stride), dtype=torch.float32, device=device)
    occ = (torch.zeros((ncell + 31) // 32, dtype=torch.int32, device=device)
           if have_occupancy else None)
    cage = io = c2d = None
    if sparse:
        cage = torch.zeros(W * H * D, dtype=torch.uint8, device=device)
        io = torch.zeros(W * H * D, dtype=torch.int32, device=device)
        c2d = torch.zeros(max(1, nrec), dtype=torch.int32, device=device)
    _check("pxl_grid_load_text",
           lib().pxl_grid_load_text(d, band, int(sparse), int(have_occupancy), W, H, D, S, stride,
                                    nrec, _ptr(rec), _ptr(occ), _ptr(cage), _ptr(io), _ptr(c2d),
                                    _stream(stream)))
    if not sparse:
        return Grid(W, H, D, S, rec, step=step)
    br = build_brick_occupancy(occ, W, H, D, stream=stream)
    return Grid(W, H, D, S, rec, step=step, occupancy=occ, index_offsets=io, in_cage=cage,
                compact_to_dense=c2d, brick_occupancy=br)
