"""CPU emulation of the per-lane sinks' key arithmetic (ci_kernels.hip, count_rows_lane).

The per-lane sinks form the keys of 4 rows at once, one byte per row of a 32-bit word:
the Horner steps over (Z_1..Z_d, X, Y) run as packed 16-bit multiply-adds (v_pk_mad_u16)
on the unpacked column words, the wave pair's bias enters every byte (ORed into the Y
bytes for 64 and 128 cells, added to the key word for 96), and one v_perm_b32 moves a
row's key byte next to the lane's byte offset.  This test replays that arithmetic bit
for bit over every table shape of at most 128 cells and every cell of each, with the
rows past the end of a chunk holding arbitrary bytes, and checks that every valid row
lands at pair * kCells * 256 + key * 256 + lane * 4 with key the Counts3D index
(CellTable.cpp:268-290).  The ORed form of the 96-cell bias is replayed as well: it
must fail, or the check would not pin the fix.
"""
from __future__ import annotations

import numpy as np
import pytest

MASK16 = 0xFFFF


def pk_mad(a, b, c):
    """v_pk_mad_u16: a * b + c in each 16-bit half, modulo 2^16 (uint64 arrays)."""
    lo = ((a & MASK16) * (b & MASK16) + (c & MASK16)) & MASK16
    hi = (((a >> 16) & MASK16) * ((b >> 16) & MASK16) + ((c >> 16) & MASK16)) & MASK16
    return lo | (hi << 16)


def perm_offset(K, lane4, j):
    """__builtin_amdgcn_perm(K, lane4, 0x0C0C0000 | ((4 + j) << 8)): byte 0 from lane4,
    byte 1 from byte j of K, bytes 2-3 zero."""
    return (((K >> (8 * j)) & 0xFF) << 8) | (lane4 & 0xFF)


def words(vals):
    """Rows (multiple of 4) -> packed words, row 4q + j in byte j of word q."""
    v = vals.reshape(-1, 4).astype(np.uint64)
    return v[:, 0] | (v[:, 1] << 8) | (v[:, 2] << 16) | (v[:, 3] << 24)


def shapes(max_cells, max_d):
    """Every (zcards, cx, cy) with product <= max_cells: Z cardinalities >= 2, X and Y >= 1."""
    out = []

    def rec(zc, dz):
        for cx in range(1, max_cells // dz + 1):
            for cy in range(1, max_cells // (dz * cx) + 1):
                out.append((zc, cx, cy))
        if len(zc) < max_d:
            for c in range(2, max_cells // dz + 1):
                rec(zc + (c,), dz * c)

    rec((), 1)
    return out


def replay(zc, cx, cy, k_cells, pair, or_bias, rng):
    """Offsets of every cell of one table shape as the kernel computes them."""
    dz = int(np.prod(zc)) if zc else 1
    cells = dz * cx * cy
    idx = np.arange(cells, dtype=np.int64)
    digits = []
    rem = idx
    for c in (cy, cx) + tuple(reversed(zc)):
        digits.append(rem % c)
        rem = rem // c
    ydig, xdig = digits[0], digits[1]
    zdig = list(reversed(digits[2:]))
    n_rows = ((cells + 3) // 4) * 4
    pad = n_rows - cells

    def col(v):  # valid rows, then arbitrary bytes past the end (uint8 columns may hold any)
        return np.concatenate([v, rng.integers(0, 256, pad)]).astype(np.uint64)

    bias = k_cells * 0x01010101 if pair else 0
    ybias = bias if or_bias else 0
    kbias = 0 if or_bias else bias
    wz = [words(col(z)) for z in zdig]
    wx = words(col(xdig))
    wy = words(col(ydig)) | np.uint64(ybias)
    if zc:
        K = wz[0]
        for j in range(1, len(zc)):
            K = pk_mad(K, np.uint64(zc[j] * 0x10001), wz[j])
        K = pk_mad(K, np.uint64(cx * 0x10001), wx)
    else:
        K = wx
    K = (pk_mad(K, np.uint64(cy * 0x10001), wy) + np.uint64(kbias)) & np.uint64(0xFFFFFFFF)
    lane = int(rng.integers(0, 64))
    offs = np.stack([perm_offset(K, lane * 4, j) for j in range(4)], axis=1).reshape(-1)[:cells]
    want = (idx + pair * k_cells) * 256 + lane * 4
    return offs.astype(np.int64), want


@pytest.mark.parametrize("k_cells", [64, 96, 128])
def test_lane_key_bytes_every_shape(k_cells):
    rng = np.random.default_rng(k_cells)
    or_bias = (k_cells & (k_cells - 1)) == 0  # the kernel's lane_bias<kCells>::kOr
    n = 0
    for zc, cx, cy in shapes(k_cells, 5):
        for pair in (0, 1):
            got, want = replay(zc, cx, cy, k_cells, pair, or_bias, rng)
            assert np.array_equal(got, want), (k_cells, zc, cx, cy, pair)
            # inside the pair's counters: [pair][cell][lane] words of kCells * 256 bytes a pair
            assert got.max() < (pair + 1) * k_cells * 256
        n += 1
    assert n > 1000


def test_ored_96_cell_bias_is_caught():
    """The form this build had before: the 96-cell bias ORed into Y is wrong for y >= 32."""
    rng = np.random.default_rng(7)
    got, want = replay((), 2, 40, 96, 1, True, rng)
    assert not np.array_equal(got, want)
    got, want = replay((), 2, 31, 96, 1, True, rng)  # y < 32: the OR happens to be right
    assert np.array_equal(got, want)
