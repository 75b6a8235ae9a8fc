"""Oracle checks of AA1 edge anti-aliasing (no GPU): intersect.h:156-244 and
ubershader.comp:848-851, 1497-1505.

A triangle sample whose centre fails exactly one edge gets coverage 16 * (8 - d), d being its
distance to that edge in 1/8 pixel along the edge's major axis; covered samples get 0x80. The
coverage replaces the source alpha when it is 0x80, blending is forced on, and partial coverage
does not write depth. Lines are two-pixel-wide bands whose coverage falls off linearly from the
centre, so each column's coverages add up to one pixel (0x80).
"""
import numpy as np

import gif
import render_streams as rs
from oracle_lib import Oracle, swizzle

OFS = 2048 * 16
W, H = 64, 48


def ct32(vram, x, y, base=0):
    a = swizzle(x, y, base, 1, 0x00)
    return int(vram[a * 4:(a + 1) * 4].view(np.uint32)[0])


def z32(vram, x, y):
    a = swizzle(x, y, 32 * 32, 1, 0x30)
    return int(vram[a * 4:(a + 1) * 4].view(np.uint32)[0])


def draw(prim, verts, rgba=0x80204060, z=0x1000, zte=0):
    o = Oracle()
    regs = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 1, 0x00)), (0x4E, rs.zbuf(32, 0x0)),
            (0x40, rs.scissor(0, W - 1, 0, H - 1)), (0x47, rs.test_reg(zte=zte, ztst=1)),
            (0x42, rs.alpha_reg(0, 1, 0, 1, 0)), (0x00, prim)]
    for (x, y) in verts:
        regs += [(0x01, rs.rgbaq(rgba)), (0x05, rs.xyz2(OFS + x, OFS + y, z))]
    o.gif_transfer(gif.arr(gif.ad(regs)))
    return o.vram()


def test_triangle_vertical_edge_feather():
    """Left edge at x = 10.5: pixel 10's centre is 4/8 px outside (coverage 64, colour halved),
    pixel 9 is out of reach, pixels 11.. are covered (full colour); AA1 forces blending with
    As = coverage: floor(Cs * cov / 128) over a zero frame buffer."""
    x0 = 168  # 10.5 px in 1/16
    v = draw(rs.prim_reg(3, iip=0, aa1=1), [(x0, 0), (x0, 40 * 16), (60 * 16, 20 * 16)])
    cs = (0x60, 0x40, 0x20)
    for y in (10, 20, 30):
        full = ct32(v, 12, y)
        assert full == 0x80204060
        half = ct32(v, 10, y)
        cov = half >> 24
        assert cov in (64, 63), hex(half)
        for ch in range(3):
            assert (half >> (8 * ch)) & 0xFF == (cs[ch] * cov) // 128
        assert ct32(v, 9, y) == 0


def test_partial_coverage_does_not_write_depth():
    x0 = 168
    v = draw(rs.prim_reg(3, iip=0, aa1=1), [(x0, 0), (x0, 40 * 16), (60 * 16, 20 * 16)], z=0x12345, zte=1)
    assert z32(v, 12, 20) == 0x12345
    assert z32(v, 10, 20) == 0, "a partially covered sample keeps its depth"
    assert ct32(v, 10, 20) != 0, "but its colour is blended"


def test_without_aa1_the_edge_is_hard():
    x0 = 168
    v = draw(rs.prim_reg(3, iip=0), [(x0, 0), (x0, 40 * 16), (60 * 16, 20 * 16)])
    assert ct32(v, 10, 20) == 0 and ct32(v, 11, 20) == 0x80204060


def test_feather_falls_off_with_distance():
    """Edge at x = 10 + k/8: pixel 10 lies k/8 outside; coverage 16 * (8 - k)."""
    for k in range(1, 8):
        x0 = 160 + 2 * k
        v = draw(rs.prim_reg(3, iip=0, aa1=1), [(x0, 0), (x0, 40 * 16), (60 * 16, 20 * 16)])
        cov = ct32(v, 10, 20) >> 24
        assert cov in (16 * (8 - k), 16 * (8 - k) - 1), (k, cov)


def test_aa1_line_columns_add_to_one_pixel():
    """An x-major AA1 line: in every column its coverages (alpha) add up to 0x80 (a linear
    falloff over a two-pixel-wide band), within rounding."""
    v = draw(rs.prim_reg(1, iip=0, aa1=1), [(4 * 16, 20 * 16 + 5), (50 * 16, 27 * 16 + 11)])
    sums = []
    for x in range(8, 46):
        col = [ct32(v, x, y) >> 24 for y in range(H)]
        assert max(col) <= 0x80
        sums.append(sum(col))
    assert all(124 <= s <= 130 for s in sums), sums
    assert len({s for s in sums}) > 1 or sums[0] == 128
