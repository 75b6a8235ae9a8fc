"""Oracle checks of mipmapped sampling (no GPU): ubershader.comp:947-953, 1066-1258,
gs_interface.cpp:1539-1580 and 1834-1863, gs_renderer.cpp:1235-1265 and 3458-3598.

* The Q-derived level of detail ceil(K - 2^(L+4) log2|q|) is evaluated exactly. The threshold
  table the kernels use (smallest float >= 2^(k/128)) and the oracle's long-double evaluation
  are both checked against integer arithmetic: for q = M * 2^(e-23), floor(2^(L+4) log2 m) is
  bit_length(M^(2^(L+4))) - 1 - 23 * 2^(L+4), exact in Python integers.
* Rendering: levels filled with distinct solid colours, so a sprite's colour names the level
  read; trilinear blends two levels with weights floor(fract * 16) / 16, rounding half to even.
"""
import ctypes
import math
import os
import re
import struct

import numpy as np

import gif
import render_streams as rs
from oracle_lib import Oracle, oracle_lib, swizzle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OFS = 2048 * 16
LEVEL_COLOURS = [0x80102030, 0x80406080, 0x80C0A050, 0x8020E0F0, 0x80FF8000, 0x80008080, 0x80FFFFFF]
LEVEL_TBP = [64 * 32 + 32 * i for i in range(7)]  # one page per level


def lod_table():
    text = open(os.path.join(ROOT, "parallel-gs_amd", "csrc", "lod_table.hpp")).read()
    vals = [float.fromhex(v) for v in re.findall(r"(0x[0-9a-fA-F.]+p[+-]\d+)f", text)]
    assert len(vals) == 129
    return vals


def mantissa(x):
    m, e = math.frexp(x)  # x = m 2^e, m in [0.5, 1)
    return int(m * (1 << 24)), e - 1  # 24-bit mantissa M, x = M * 2^(e' - 23)


def exact_lod(q, k, l):
    """ceil(K - 2^(L+4) log2|q|) with integers only; 0 for zero, subnormal, inf and NaN."""
    a = abs(q)
    if not (a >= 2.0 ** -126) or math.isinf(a):
        return 0
    M, e = mantissa(a)
    s = 16 << l
    return k - s * e - ((M ** s).bit_length() - 1 - 23 * s)


def f32(x):
    return struct.unpack("<f", struct.pack("<f", x))[0]


def prev_float(x):
    return struct.unpack("<f", struct.pack("<I", struct.unpack("<I", struct.pack("<f", x))[0] - 1))[0]


def oracle_lod(q, k, l):
    L = oracle_lib()
    L.orc_lod_from_q.restype = ctypes.c_int32
    L.orc_lod_from_q.argtypes = [ctypes.c_float, ctypes.c_int32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)]
    pos = ctypes.c_int32()
    n = L.orc_lod_from_q(q, k, l, ctypes.byref(pos))
    assert bool(pos.value) == (n > 0)
    return n


def test_threshold_table_is_exact():
    """Entry k is the smallest float >= 2^(k/128): (M / 2^23)^128 >= 2^k > ((M - 1) / 2^23)^128."""
    t = lod_table()
    assert t[0] == 1.0 and t[128] == 2.0
    for k in range(1, 128):
        M = int(t[k] * (1 << 23))
        assert M / float(1 << 23) == t[k]
        assert M ** 128 >= 1 << (k + 23 * 128)
        assert (M - 1) ** 128 < 1 << (k + 23 * 128)
        assert t[k - 1] < t[k]


def test_oracle_lod_exact_at_every_threshold():
    """The oracle's floor of a long-double log2 matches the integer evaluation on both sides of
    every step of floor(128 log2 m) (and so of every coarser 2^(L+4) grid), at several exponents."""
    t = lod_table()
    for k in range(1, 128):
        for m in (t[k], prev_float(t[k])):
            for e in (-20, -1, 0, 3):
                q = f32(math.ldexp(m, e))
                for l in range(4):
                    for kk in (0, -37, 100):
                        assert oracle_lod(q, kk, l) == exact_lod(q, kk, l), (k, m, e, l, kk)


def test_oracle_lod_random_and_closed_forms():
    rng = np.random.default_rng(5)
    for bits in rng.integers(0x00800000, 0x7F800000, size=3000, dtype=np.uint32):
        q = struct.unpack("<f", struct.pack("<I", int(bits)))[0]
        l = int(rng.integers(0, 4))
        k = int(rng.integers(-2048, 2048))
        assert oracle_lod(q, k, l) == exact_lod(q, k, l)
        assert oracle_lod(-q, k, l) == exact_lod(q, k, l)
    # Powers of two: the formula is an integer, K + 2^(L+4) i for q = 2^-i.
    for i in range(-10, 11):
        for l in range(4):
            assert oracle_lod(2.0 ** -i, 5, l) == 5 + (16 << l) * i
    # Non-finite and flushed inputs give 0.
    for q in (0.0, -0.0, 1e-40, float("inf"), float("-inf"), float("nan")):
        assert oracle_lod(q, 77, 2) == 0


# ------------------------------------------------------------------------------------------
# Rendering with level colours.

def mip_setup(tw=4, th=4, levels=5):
    """Levels 0..levels-1 of a PSMCT32 2^tw x 2^th texture, level l solid LEVEL_COLOURS[l] at
    LEVEL_TBP[l] (TBW 1)."""
    out = []
    for lv in range(levels):
        # At least 2 x 2: narrower 32-bit host transfers move no pixels.
        w, h = max((1 << tw) >> lv, 2), max((1 << th) >> lv, 2)
        px = np.full(w * h, LEVEL_COLOURS[lv], np.uint32).tobytes()
        out.append(gif.upload(LEVEL_TBP[lv], 1, 0x00, 0, 0, w, h, px + b"\0" * (-len(px) % 16)))
    return out


def draw(t1, q=1.0, tw=4, th=4, levels=5, clamp=0, prim=None, size=16):
    """A 16x16 DECAL sprite sampling UV 0..16 with TEX1 = t1 and RGBAQ.Q = q; returns the oracle's
    frame buffer pixels (64 px stride)."""
    o = Oracle()
    out = [gif.ad([(0x1A, 1), (0x18, OFS | OFS << 32)])] + mip_setup(tw, th, levels)
    regs = [(0x4C, rs.frame(0, 1, 0x00)), (0x40, rs.scissor(0, 63, 0, 63)), (0x47, rs.test_reg(zte=1, ztst=1)),
            (0x14, t1), (0x08, clamp),
            (0x34, rs.miptbp([(LEVEL_TBP[i], 1) for i in (1, 2, 3)])),
            (0x36, rs.miptbp([(LEVEL_TBP[i], 1) for i in (4, 5, 6)])),
            (0x06, rs.tex0(LEVEL_TBP[0], 1, 0x00, tw, th, tcc=1, tfx=1)),
            (0x00, prim if prim is not None else rs.prim_reg(6, iip=0, tme=1, fst=1)),
            (0x01, rs.rgbaq(0x80808080, q)), (0x03, rs.uv(8, 8)), (0x05, rs.xyz2(OFS, OFS, 0)),
            (0x01, rs.rgbaq(0x80808080, q)), (0x03, rs.uv(size * 16 + 8, size * 16 + 8)),
            (0x05, rs.xyz2(OFS + size * 16, OFS + size * 16, 0))]
    out.append(gif.ad(regs))
    o.gif_transfer(gif.arr(b"".join(out)))
    v = o.vram()
    return np.array([[int(v[4 * swizzle(x, y, 0, 1, 0x00):][:4].view(np.uint32)[0]) for x in range(size)]
                     for y in range(size)], np.uint64)


def only(px):
    vals = set(int(v) for v in px.ravel())
    assert len(vals) == 1, [hex(v) for v in sorted(vals)][:8]
    return vals.pop()


def blend(a, b, w):
    """Per channel round-half-even(((16 - w) a + w b) / 16)."""
    out = 0
    for ch in range(4):
        x = ((a >> (8 * ch)) & 0xFF) * (16 - w) + ((b >> (8 * ch)) & 0xFF) * w
        q, r = divmod(x, 16)
        if r > 8 or (r == 8 and q & 1):
            q += 1
        out |= q << (8 * ch)
    return out


def test_fixed_lod_selects_level():
    """LCM = 1: LOD = K >> 4 clamped to [0, MXL]."""
    for lv in range(5):
        assert only(draw(rs.tex1(0, 2, lcm=1, mxl=4, k=16 * lv))) == LEVEL_COLOURS[lv]
    assert only(draw(rs.tex1(0, 2, lcm=1, mxl=4, k=16 * 6 + 5))) == LEVEL_COLOURS[4]
    assert only(draw(rs.tex1(0, 2, lcm=1, mxl=2, k=16 * 4))) == LEVEL_COLOURS[2], "clamped to MXL"
    assert only(draw(rs.tex1(0, 2, lcm=1, mxl=4, k=-40))) == LEVEL_COLOURS[0]


def test_without_a_mipmapped_filter_level_zero_is_read():
    for mmin in (0, 1, 6, 7):
        assert only(draw(rs.tex1(0, mmin, lcm=1, mxl=4, k=16 * 3))) == LEVEL_COLOURS[0]


def test_q_lod_selects_level():
    """LCM = 0: LOD = ceil(K - 2^(L+4) log2 Q) / 16; Q = 2^-i gives K / 16 + 2^L i."""
    for i in range(5):
        assert only(draw(rs.tex1(0, 2, mxl=4), q=2.0 ** -i)) == LEVEL_COLOURS[i]
    assert only(draw(rs.tex1(0, 2, mxl=4, l=1), q=0.5)) == LEVEL_COLOURS[2]
    assert only(draw(rs.tex1(0, 2, mxl=4, k=32), q=2.0)) == LEVEL_COLOURS[1]
    assert only(draw(rs.tex1(0, 2, mxl=4), q=4.0)) == LEVEL_COLOURS[0], "negative LOD clamps to 0"


def test_nearest_level_rounds_half_up():
    """Without MIPMAP_LINEAR the level is int(LOD + 0.5)."""
    assert only(draw(rs.tex1(0, 2, mxl=4, k=16 + 8))) == LEVEL_COLOURS[2]
    assert only(draw(rs.tex1(0, 2, mxl=4, k=16 + 7))) == LEVEL_COLOURS[1]
    # ceil of a non-integer formula: Q slightly below 1 makes the LOD just above K / 16.
    assert only(draw(rs.tex1(0, 2, mxl=4, k=16 + 7), q=0.999)) == LEVEL_COLOURS[2]


def test_trilinear_blends_two_levels():
    """NEAREST_MIPMAP_LINEAR at LOD 1 + w/16 blends levels 1 and 2 with weight w."""
    for w in range(16):
        got = only(draw(rs.tex1(0, 3, mxl=4, k=16 + w)))
        assert got == blend(LEVEL_COLOURS[1], LEVEL_COLOURS[2], w), (w, hex(got))


def test_lod_past_the_last_image_level_reads_the_last():
    """A 16x4 texture holds 3 levels (16x4, 8x2, 4x1) whatever MXL says; LOD 5 reads level 2."""
    assert only(draw(rs.tex1(0, 2, lcm=1, mxl=6, k=16 * 5), tw=4, th=2, levels=3)) == LEVEL_COLOURS[2]


def test_min_and_mag_filters_follow_the_lod():
    """A checkerboard level 1 sampled at half-texel offsets: MMIN = LINEAR_MIPMAP_NEAREST blends
    it once the LOD is above 0, MMAG = LINEAR alone does not."""
    o_lin, o_near = [], []
    for mmag, mmin, dst in ((0, 4, o_lin), (1, 2, o_near)):
        o = Oracle()
        cb = np.array([[0x80FFFFFF if (x + y) & 1 else 0x80000000 for x in range(8)] for y in range(8)],
                      np.uint32).tobytes()
        out = [gif.ad([(0x1A, 1), (0x18, OFS | OFS << 32)]), gif.upload(LEVEL_TBP[1], 1, 0x00, 0, 0, 8, 8, cb)]
        regs = [(0x4C, rs.frame(0, 1, 0x00)), (0x40, rs.scissor(0, 63, 0, 63)), (0x47, rs.test_reg(zte=1, ztst=1)),
                (0x14, rs.tex1(mmag, mmin, lcm=1, mxl=1, k=16)), (0x08, 0),
                (0x34, rs.miptbp([(LEVEL_TBP[1], 1), (0, 1), (0, 1)])),
                (0x06, rs.tex0(LEVEL_TBP[0], 1, 0x00, 4, 4, tcc=1, tfx=1)),
                (0x00, rs.prim_reg(6, iip=0, tme=1, fst=1)),
                (0x03, rs.uv(0, 0)), (0x05, rs.xyz2(OFS, OFS, 0)),
                (0x03, rs.uv(16 * 16, 16 * 16)), (0x05, rs.xyz2(OFS + 16 * 16, OFS + 16 * 16, 0))]
        out.append(gif.ad(regs))
        o.gif_transfer(gif.arr(b"".join(out)))
        v = o.vram()
        dst.extend(int(v[4 * swizzle(x, y, 0, 1, 0x00):][:4].view(np.uint32)[0]) for y in range(16) for x in range(16))
    assert set(o_near) <= {0x80FFFFFF, 0x80000000}
    assert len(set(o_lin) - {0x80FFFFFF, 0x80000000}) > 0
