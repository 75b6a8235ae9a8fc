"""Swizzle: the oracle pinned against the GS block/column arrangement tables, and the
product's host/device swizzle checked against the oracle.

The tables below are the PS2 GS local-memory arrangements (block order inside a page and
word/byte/nibble order inside a column) as published in the GS hardware documentation;
they are typed here independently of any code under test.
"""
import numpy as np
import pytest

from oracle_lib import swizzle as oswz
import pgs

P = pgs.PSM

# Block number at (block column, block row) inside one page.
BLOCKS_32 = [[0, 1, 4, 5, 16, 17, 20, 21],
             [2, 3, 6, 7, 18, 19, 22, 23],
             [8, 9, 12, 13, 24, 25, 28, 29],
             [10, 11, 14, 15, 26, 27, 30, 31]]
BLOCKS_16 = [[0, 2, 8, 10], [1, 3, 9, 11], [4, 6, 12, 14], [5, 7, 13, 15],
             [16, 18, 24, 26], [17, 19, 25, 27], [20, 22, 28, 30], [21, 23, 29, 31]]
BLOCKS_16S = [[0, 2, 16, 18], [1, 3, 17, 19], [8, 10, 24, 26], [9, 11, 25, 27],
              [4, 6, 20, 22], [5, 7, 21, 23], [12, 14, 28, 30], [13, 15, 29, 31]]
BLOCKS_Z32 = [[24, 25, 28, 29, 8, 9, 12, 13],
              [26, 27, 30, 31, 10, 11, 14, 15],
              [16, 17, 20, 21, 0, 1, 4, 5],
              [18, 19, 22, 23, 2, 3, 6, 7]]

# Element order inside one column (column 0; odd columns of 8/4-bit swap the halves).
COLUMN_32 = [[0, 1, 4, 5, 8, 9, 12, 13],
             [2, 3, 6, 7, 10, 11, 14, 15]]
COLUMN_16 = [[0, 2, 8, 10, 16, 18, 24, 26, 1, 3, 9, 11, 17, 19, 25, 27],
             [4, 6, 12, 14, 20, 22, 28, 30, 5, 7, 13, 15, 21, 23, 29, 31]]
COLUMN_8 = [[0, 4, 16, 20, 32, 36, 48, 52, 2, 6, 18, 22, 34, 38, 50, 54],
            [8, 12, 24, 28, 40, 44, 56, 60, 10, 14, 26, 30, 42, 46, 58, 62],
            [33, 37, 49, 53, 1, 5, 17, 21, 35, 39, 51, 55, 3, 7, 19, 23],
            [41, 45, 57, 61, 9, 13, 25, 29, 43, 47, 59, 63, 11, 15, 27, 31]]
COLUMN_4 = [[0, 8, 32, 40, 64, 72, 96, 104, 2, 10, 34, 42, 66, 74, 98, 106,
             4, 12, 36, 44, 68, 76, 100, 108, 6, 14, 38, 46, 70, 78, 102, 110],
            [16, 24, 48, 56, 80, 88, 112, 120, 18, 26, 50, 58, 82, 90, 114, 122,
             20, 28, 52, 60, 84, 92, 116, 124, 22, 30, 54, 62, 86, 94, 118, 126],
            [65, 73, 97, 105, 1, 9, 33, 41, 67, 75, 99, 107, 3, 11, 35, 43,
             69, 77, 101, 109, 5, 13, 37, 45, 71, 79, 103, 111, 7, 15, 39, 47],
            [81, 89, 113, 121, 17, 25, 49, 57, 83, 91, 115, 123, 19, 27, 51, 59,
             85, 93, 117, 125, 21, 29, 53, 61, 87, 95, 119, 127, 23, 31, 55, 63]]

# (block w, block h, elements per byte-address unit) per family.
GEOM = {
    P["CT32"]: (8, 8, BLOCKS_32), P["CT24"]: (8, 8, BLOCKS_32), P["Z32"]: (8, 8, BLOCKS_Z32),
    P["CT16"]: (16, 8, BLOCKS_16), P["CT16S"]: (16, 8, BLOCKS_16S),
    P["T8"]: (16, 16, BLOCKS_32), P["T4"]: (32, 16, BLOCKS_16),
}


@pytest.mark.parametrize("psm", list(GEOM))
def test_oracle_block_arrangement(psm):
    bw, bh, table = GEOM[psm]
    # Element size in the oracle's address unit: 32-bit formats address words, 16-bit
    # halfwords, PSMT8 bytes, PSMT4 nibbles; a block is 256 bytes.
    per_block = {P["CT32"]: 64, P["CT24"]: 64, P["Z32"]: 64, P["CT16"]: 128, P["CT16S"]: 128,
                 P["T8"]: 256, P["T4"]: 512}[psm]
    for by, row in enumerate(table):
        for bx, expected in enumerate(row):
            a = oswz(bx * bw, by * bh, 0, 64, psm)
            assert a // per_block == expected, (hex(psm), bx, by)


def test_oracle_column_layouts():
    for y in range(2):
        for x in range(8):
            assert oswz(x, y, 0, 1, P["CT32"]) == COLUMN_32[y][x]
    for y in range(2):
        for x in range(16):
            assert oswz(x, y, 0, 1, P["CT16"]) == COLUMN_16[y][x]
    for y in range(4):
        for x in range(16):
            assert oswz(x, y, 0, 2, P["T8"]) == COLUMN_8[y][x]
    for y in range(4):
        for x in range(32):
            assert oswz(x, y, 0, 2, P["T4"]) == COLUMN_4[y][x]


def test_oracle_t8_odd_column_swaps_halves():
    # In odd columns the upper and lower 32-byte halves of the column are exchanged.
    for y in range(4):
        for x in range(16):
            a0 = oswz(x, y, 0, 2, P["T8"])
            a1 = oswz(x, y + 4, 0, 2, P["T8"]) - 64
            assert a1 == (a0 ^ 32)


ALL_PSMS = list(P.values())


@pytest.mark.parametrize("psm", ALL_PSMS)
def test_oracle_page_bijection(psm):
    """Every pixel of one page maps to a distinct element inside that page's 8 KiB."""
    geom = {0x02: (64, 64), 0x0A: (64, 64), 0x32: (64, 64), 0x3A: (64, 64), 0x13: (128, 64), 0x14: (128, 128)}
    pw, ph = geom.get(psm, (64, 32))
    units = {0x13: 8192, 0x14: 16384}.get(psm, 4096 if psm in (0x02, 0x0A, 0x32, 0x3A) else 2048)
    seen = set()
    for y in range(ph):
        for x in range(pw):
            seen.add(oswz(x, y, 0, 2, psm))
    if psm in (0x1B, 0x24, 0x2C, 0x01, 0x31):  # share the 32-bit layout
        units = 2048
    assert len(seen) == pw * ph
    assert min(seen) == 0 and max(seen) == units - 1


@pytest.mark.parametrize("pair", [(0x00, 0x30), (0x01, 0x31), (0x02, 0x32), (0x0A, 0x3A)])
def test_oracle_z_is_color_xor_0x18(pair):
    c, z = pair
    per_block = 64 if c in (0x00, 0x01) else 128
    for base in (0, 5, 32, 100):
        for by in range(8):
            for bx in range(8):
                x, y = bx * 8, by * 8
                bc = oswz(x, y, base, 4, c) // per_block
                bz = oswz(x, y, base, 4, z) // per_block
                assert bz == bc ^ 0x18, (base, bx, by)


def _element_shift(psm):
    if psm in (0x02, 0x0A, 0x32, 0x3A):
        return 1
    if psm == 0x13:
        return 0
    if psm == 0x14:
        return 0  # nibble address in both
    return 2


@pytest.mark.parametrize("psm", ALL_PSMS)
def test_product_swizzle_matches_oracle(psm):
    rng = np.random.default_rng(psm)
    sh = _element_shift(psm)
    for _ in range(4000):
        x, y = int(rng.integers(0, 2048)), int(rng.integers(0, 2048))
        base, stride = int(rng.integers(0, 16384)), int(rng.integers(0, 64))
        a = pgs.host_swizzle_addr(x, y, base, stride, psm)
        assert (a >> sh) == oswz(x, y, base, stride, psm), (hex(psm), x, y, base, stride)


@pytest.mark.parametrize("psm", ALL_PSMS)
def test_block_lane_inverse(psm):
    """The copy kernel's lane -> pixel map inverts the forward swizzle inside a block."""
    sc = 1 if psm in (0x02, 0x0A, 0x32, 0x3A) else 2 if psm == 0x13 else 3 if psm == 0x14 else 0
    npix = [1, 2, 4, 8][sc]
    elem_bits = [32, 16, 8, 4][sc]
    seen = set()
    for lane in range(64):
        for k in range(npix):
            dx, dy = pgs.host_block_lane_pixel(psm, lane, k)
            # Offset inside the block (Z formats relocate the block itself).
            a = pgs.host_swizzle_addr(dx, dy, 0, 2, psm) % (512 if sc == 3 else 256)
            # byte address (nibble address for T4) of element k of dword `lane`
            if sc == 3:
                expect = lane * 8 + k
            else:
                expect = lane * 4 + k * (elem_bits // 8)
            assert a == expect, (hex(psm), lane, k, dx, dy, a, expect)
            seen.add((dx, dy))
    assert len(seen) == 64 * npix
