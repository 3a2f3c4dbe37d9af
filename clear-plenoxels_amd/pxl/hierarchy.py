"""Coarse-to-fine training loop over grid levels (trainSynthesisBlender,
SparseVolumeRadianceField/VolumeFieldRender.cu:456-577).

Level i trains a grid of base / 2^i vertices per axis; after every level but
the last the grid is upsampled to twice its size (initSelfByUpsample,
SceneUtils.cu:487-657, here pxl.upsample_grid).  The plan of levels
(pxl_hierarchy_plan) and each level's band, epoch and batch loops
(pxl_trainer_learn_level, Render::objectSceneLearner, Render.cu:264-503) run
in the C++ host layer; this module allocates each level's buffers (device
memory comes from torch) and chains the calls.
"""
import math

import pxl


def plan(hierarchy, minimum_hierarchy, epochs, batch_size, shuffle_rays=True,
         fix_batch_size=False, to_sparse_hierarchy=0):
    """Per-level settings of the loop, coarsest level first: a list of dicts
    with `level`, `epochs`, `batch_size` and `upsample` (None after the last
    level, else {"to_sparse": bool}), from pxl_hierarchy_plan.

    - epochs: the configured count plus 3 - hierarchy (:459), then +1 per
      level with shuffled rays, 10 without (:476-489);
    - batch size: batch_size << (3 - level) unless fixed (:491-495); the
      reference shifts by a negative count above level 3, which C leaves
      undefined, so that case is rejected (ValueError);
    - upsample after level i > minimum: to a caged grid when
      i <= minimum_hierarchy - to_sparse_hierarchy (:549-557), dense
      otherwise.
    """
    try:
        return pxl.hierarchy_plan(hierarchy, minimum_hierarchy, epochs, batch_size,
                                  shuffle_rays, fix_batch_size, to_sparse_hierarchy)
    except pxl.PxlError as e:
        raise ValueError(str(e)) from e


def train(grid, make_rays, levels, prune_threshold, on_level=None, sh_band=None,
          hierarchy_band_strategy=False, sh_for_target_only=True, **trainer_kw):
    """Run the levels of `plan` starting from `grid` (the coarsest level's,
    base / 2^hierarchy per axis); returns the final grid.

    make_rays(level, grid) -> (rays, order): the level's training rays (a
    pxl.Rays against this grid's bounding box) and the int32 device tensor of
    ray indices the trainer visits each epoch.  Each level is one
    pxl_trainer_learn_level call on a fresh pxl.Trainer (learning rates from
    trainer_kw, optimize_step = level): the SH bands of the level (with
    sh_for_target_only every level above the last trains bands 0 .. sh_band
    in turn, each from a cleared optimiser state; with
    hierarchy_band_strategy every level does), each for the level's epochs.
    on_level(level_dict, grid, trainer) is called after the level's last
    batch, before the upsampling.
    """
    for k in ("batch_size", "optimize_step"):
        if k in trainer_kw:
            raise ValueError(f"{k} is set per level by the plan")
    band = sh_band if sh_band is not None else int(round(math.sqrt(grid.S))) - 1
    if (band + 1) ** 2 != grid.S:
        raise ValueError(f"sh_band {band} does not match the grid's {grid.S} coefficients")
    minimum = levels[-1]["level"]
    upsampled = False
    for lv in levels:
        rays, order = make_rays(lv["level"], grid)
        n = int(order.numel())
        if n == 0:
            raise ValueError(f"level {lv['level']}: no rays")
        optim = pxl.Optim(grid)
        tr = pxl.Trainer(grid, optim, rays, order, batch_size=lv["batch_size"],
                         optimize_step=lv["level"], **trainer_kw)
        try:
            tr.learn_level(lv["epochs"], band, hierarchy_band_strategy=hierarchy_band_strategy,
                           sh_for_target_only=sh_for_target_only,
                           current_hierarchy=lv["level"], minimum_hierarchy=minimum,
                           upsampled=upsampled)
            # The trainer restricts the colour to the band it trained last;
            # the grid object keeps using every coefficient it holds.
            if on_level is not None:
                on_level(lv, grid, tr)
        finally:
            tr.close()
        if lv["upsample"] is not None:
            grid = pxl.upsample_grid(grid, prune_threshold, to_sparse=lv["upsample"]["to_sparse"])
            upsampled = True
    return grid
