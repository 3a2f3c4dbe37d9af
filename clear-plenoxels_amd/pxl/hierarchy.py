"""Coarse-to-fine training loop over grid levels (trainSynthesisBlender,
SparseVolumeRadianceField/VolumeFieldRender.cu:456-577).

Level i trains a grid of base / 2^i vertices per axis; after every level but
the last the grid is upsampled to twice its size (initSelfByUpsample,
SceneUtils.cu:487-657, here pxl.upsample_grid).  Per level the reference
resets both learning rates to their starting values (:466-467), adds one
epoch when rays are shuffled and uses 10 otherwise (:476-489), and doubles
the batch size per level below 3 (:491-495).  The batch loop of each level
is the C++ trainer (pxl.Trainer, Render::objectSceneLearner).
"""
import math

import pxl


def plan(hierarchy, minimum_hierarchy, epochs, batch_size, shuffle_rays=True,
         fix_batch_size=False, to_sparse_hierarchy=0):
    """Per-level settings of the loop, coarsest level first: a list of dicts
    with `level`, `epochs`, `batch_size` and `upsample` (None after the last
    level, else {"to_sparse": bool}).

    - epochs: the configured count plus 3 - hierarchy (:459), then +1 per
      level with shuffled rays, 10 without (:476-489);
    - batch size: batch_size << (3 - level) unless fixed (:491-495); the
      reference shifts by a negative count above level 3, which C leaves
      undefined, so that case is rejected here;
    - upsample after level i > minimum: to a caged grid when
      i <= minimum_hierarchy - to_sparse_hierarchy (:549-557), dense
      otherwise.
    """
    if minimum_hierarchy > hierarchy:
        raise ValueError("minimum_hierarchy must not exceed hierarchy")
    if not fix_batch_size and hierarchy > 3:
        raise ValueError("batch doubling (batch_size << (3 - level)) needs levels <= 3")
    out = []
    ep = epochs + 3 - hierarchy
    for i in range(hierarchy, minimum_hierarchy - 1, -1):
        ep = ep + 1 if shuffle_rays else 10
        bs = batch_size if fix_batch_size else batch_size << (3 - i)
        up = None
        if i > minimum_hierarchy:
            up = {"to_sparse": i <= minimum_hierarchy - to_sparse_hierarchy}
        out.append({"level": i, "epochs": ep, "batch_size": bs, "upsample": up})
    return out


def train(grid, make_rays, levels, prune_threshold, on_level=None, **trainer_kw):
    """Run the levels of `plan` starting from `grid` (the coarsest level's,
    base / 2^hierarchy per axis); returns the final grid.

    make_rays(level, grid) -> (rays, order): the level's training rays (a
    pxl.Rays against this grid's bounding box) and the int32 device tensor of
    ray indices the trainer visits each epoch.  Each level runs
    epochs * ceil(len(order) / batch_size) batches of a fresh pxl.Trainer
    (fresh optimiser state, learning rates from trainer_kw, optimize_step =
    level); on_level(level_dict, grid, trainer) is called after the level's
    last batch, before the upsampling.  Every level trains the same SH band
    (the reference's SHsForTargetResoOnly / per-band strategy is not built).
    """
    for k in ("batch_size", "optimize_step"):
        if k in trainer_kw:
            raise ValueError(f"{k} is set per level by the plan")
    for lv in levels:
        rays, order = make_rays(lv["level"], grid)
        n = int(order.numel())
        if n == 0:
            raise ValueError(f"level {lv['level']}: no rays")
        optim = pxl.Optim(grid)
        tr = pxl.Trainer(grid, optim, rays, order, batch_size=lv["batch_size"],
                         optimize_step=lv["level"], **trainer_kw)
        try:
            for _ in range(lv["epochs"] * math.ceil(n / lv["batch_size"])):
                tr.step()
            if on_level is not None:
                on_level(lv, grid, tr)
        finally:
            tr.close()
        if lv["upsample"] is not None:
            grid = pxl.upsample_grid(grid, prune_threshold, to_sparse=lv["upsample"]["to_sparse"])
    return grid
