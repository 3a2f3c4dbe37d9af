"""Level plan of the coarse-to-fine loop (VolumeFieldRender.cu:456-577),
worked by hand from the reference's rules for three configurations:

- hierarchy 2 .. 0, 5 epochs, shuffled rays, batch 4096: epochs start at
  5 + 3 - 2 = 6 (:459) and grow by one per level (:478-489) -> 7, 8, 9;
  batches 4096 << 1, << 2, << 3 (:491-495); upsampling after levels 2 and 1,
  caged when level <= 0 - (-1) = 1 (:549-557) -> dense after 2, caged after 1;
- unshuffled rays: every level runs 10 epochs; a fixed batch stays;
- a single level: no upsampling.
"""
import pytest

from pxl import hierarchy


def test_three_levels_shuffled():
    p = hierarchy.plan(2, 0, 5, 4096, shuffle_rays=True, to_sparse_hierarchy=-1)
    assert [lv["level"] for lv in p] == [2, 1, 0]
    assert [lv["epochs"] for lv in p] == [7, 8, 9]
    assert [lv["batch_size"] for lv in p] == [8192, 16384, 32768]
    assert p[0]["upsample"] == {"to_sparse": False}
    assert p[1]["upsample"] == {"to_sparse": True}
    assert p[2]["upsample"] is None


def test_unshuffled_fixed_batch():
    p = hierarchy.plan(3, 1, 2, 1000, shuffle_rays=False, fix_batch_size=True)
    assert [lv["level"] for lv in p] == [3, 2, 1]
    assert [lv["epochs"] for lv in p] == [10, 10, 10]
    assert [lv["batch_size"] for lv in p] == [1000] * 3
    # to_sparse_hierarchy 0: caged only when level <= 1, never true after a level > 1
    assert [lv["upsample"] for lv in p] == [{"to_sparse": False}, {"to_sparse": False}, None]


def test_single_level():
    p = hierarchy.plan(0, 0, 4, 512)
    assert p == [{"level": 0, "epochs": 8, "batch_size": 4096, "upsample": None}]


def test_rejects_undefined_shift_and_inverted_range():
    with pytest.raises(ValueError):
        hierarchy.plan(4, 0, 1, 256)
    hierarchy.plan(4, 0, 1, 256, fix_batch_size=True)
    with pytest.raises(ValueError):
        hierarchy.plan(0, 1, 1, 256)
