"""CPU emulation of the partitioned counting's slice arithmetic (ci_kernels.hip, part_slice_of).

A partitioned table's z slices hold sc = floor(16384 / c(X)c(Y)) * c(X)c(Y) cells (whole z
strata), at most 4096 of them, so keys stay below 2^26.  The kernel finds a key's slice
with an FP32 reciprocal, q = trunc(float(key) * (1.0f / sc)), and corrects q by one from the
sign of the remainder.  That is exact only if the FP32 quotient is never off by more than
one, which holds because sc > 8192 for every c(X)c(Y) <= 16384.  This test checks that
bound for every c(X)c(Y), and replays the FP32 arithmetic bit for bit (numpy float32, the
same IEEE single-precision rounding) on every key next to a slice boundary and on random
keys, for a spread of slice sizes including the smallest and largest.
"""
from __future__ import annotations

import numpy as np

SLICE_CELLS = 16384        # kSliceCells
MAX_SLICES = 4096          # kPartMaxSlices
KEY_LIMIT = MAX_SLICES * SLICE_CELLS


def slice_cells(cxy):
    return (SLICE_CELLS // cxy) * cxy


def part_slice_of(key, sc):
    """The kernel's part_slice_of over uint32 arrays: (slice, key within it)."""
    inv = np.float32(1.0) / np.float32(sc)
    q = (key.astype(np.float32) * inv).astype(np.uint32)  # FP32 product, truncated
    r = (key - q * np.uint32(sc)).astype(np.int64)
    r = np.where(r >= 2 ** 31, r - 2 ** 32, r)          # the int32 view of the remainder
    low, high = r < 0, r >= sc
    q = q - low.astype(np.uint32) + high.astype(np.uint32)
    r = r + low * sc - high * sc
    return q, r


def test_slices_hold_more_than_half_the_slice_cells():
    scs = np.array([slice_cells(c) for c in range(1, SLICE_CELLS + 1)])
    assert scs.min() > SLICE_CELLS // 2 and scs.max() <= SLICE_CELLS
    # every slice is whole strata
    assert all(slice_cells(c) % c == 0 for c in range(1, SLICE_CELLS + 1))


def test_fp32_slice_quotient_is_exact():
    rng = np.random.default_rng(11)
    scs = sorted({slice_cells(c) for c in range(1, SLICE_CELLS + 1)})
    picks = set(scs[:5] + scs[-5:]) | set(rng.choice(scs, size=40, replace=False).tolist())
    for sc in sorted(picks):
        k = np.arange(0, KEY_LIMIT // sc + 1, dtype=np.int64) * sc
        keys = np.concatenate([k - 1, k, k + 1, rng.integers(0, KEY_LIMIT, 20_000)])
        keys = keys[(keys >= 0) & (keys < KEY_LIMIT)].astype(np.uint32)
        q, r = part_slice_of(keys, sc)
        assert np.array_equal(q.astype(np.int64), keys.astype(np.int64) // sc), sc
        assert np.array_equal(r, keys.astype(np.int64) % sc), sc


def test_a_wider_fp32_error_would_be_caught():
    # the correction step covers an error of one; a quotient taken from a coarser
    # reciprocal (bfloat16-like, 8 significant bits) is off by more and must fail
    sc = slice_cells(8193)
    keys = np.arange(0, KEY_LIMIT, 4099, dtype=np.uint32)
    inv = np.float32(np.float32(1.0) / np.float32(sc)).view(np.uint32) & np.uint32(0xFFFF0000)
    q = (keys.astype(np.float32) * inv.view(np.float32)).astype(np.uint32)
    r = keys.astype(np.int64) - q.astype(np.int64) * sc
    one_step = q.astype(np.int64) - (r < 0) + (r >= sc)
    assert not np.array_equal(one_step, keys.astype(np.int64) // sc)
