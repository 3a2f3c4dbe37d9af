"""ShuffleRank (SparseVolumeRadianceField/Utils.cu:488-506) through
pxl_shuffle_rank, a host-only entry point (no GPU needed).

Pins: the known answer for glibc's published rand() sequence after srand(1)
(1804289383, 846930886, 1681692777, 1714636915, 1957747793, 424238335,
719885386, ...); the invariant of the top-down swap with a position drawn
from [0, i) (Sattolo's algorithm): one single cycle through every index; the
untouched index below min_index 1; the rejection of min_index values whose
draws leave the range."""
import ctypes

import numpy as np
import pytest

import pxl


def _libc():
    libc = ctypes.CDLL(None)
    if not hasattr(libc, "gnu_get_libc_version"):
        pytest.skip("the known answer is glibc's rand() sequence")
    return libc


def _cycle_length(p, start=0):
    n, i = 1, p[start]
    while i != start:
        i = p[i]
        n += 1
    return n


def test_shuffle_rank_known_answer():
    libc = _libc()
    libc.srand(1)
    out = pxl.shuffle_rank(np.zeros(8, np.int32), 8)
    # j = r % i for i = 7..1 with r the first seven values above:
    # 1, 4, 2, 3, 2, 1, 0 -> swaps (7,1) (6,4) (5,2) (4,3) (3,2) (2,1) (1,0)
    assert out.tolist() == [6, 0, 7, 5, 3, 2, 4, 1]


def test_shuffle_rank_single_cycle_and_prefix():
    libc = _libc()
    libc.srand(12345)
    out = pxl.shuffle_rank(np.zeros(1000, np.int32), 1000)
    assert sorted(out.tolist()) == list(range(1000))
    assert _cycle_length(out) == 1000
    # min_index 1: index 0 keeps its value, [1, n) is permuted (the draw
    # may pick i itself, so not always one cycle)
    dest = np.arange(1000, dtype=np.int32)[::-1].copy()
    libc.srand(7)
    pxl.shuffle_rank(dest, 1000, regenerate=True, min_index=1)
    assert dest[0] == 999
    assert sorted(dest[1:].tolist()) == list(range(1, 1000))
    # shuffle off: identity (regenerate) or untouched
    same = pxl.shuffle_rank(np.full(10, 3, np.int32), 10, regenerate=False, shuffle=False)
    assert np.all(same == 3)


def test_shuffle_rank_rejects_bad_range():
    with pytest.raises(pxl.PxlError):
        pxl.shuffle_rank(np.zeros(4, np.int32), 2, min_index=3)
    # min_index 2 would draw past the range (the reference's formula)
    with pytest.raises(pxl.PxlError):
        pxl.shuffle_rank(np.zeros(8, np.int32), 8, min_index=2)
    out = pxl.shuffle_rank(np.zeros(8, np.int32), 8, min_index=2, shuffle=False)
    assert out.tolist() == [0, 0, 2, 3, 4, 5, 6, 7]
