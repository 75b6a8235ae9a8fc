"""CPU oracle pinned by closed forms and exhaustive sweeps (no GPU).

Each expectation below is computed independently of the oracle, from the GS behaviour the
reference implements, so that a plausible mistake in the oracle (a swapped comparison, a
dropped term, floor instead of round, a wrong table) fails:

* reciprocals: rcp_fixed over every 32-bit input and rcp_float over every mantissa of [1, 2)
  against the correctly rounded 1/x (gs/shaders/math_utils.h:27-101), and the float seed LUT
  against its formula 1/(1 + (i + 0.5)/256) (gs/gs_renderer.cpp:3638-3640);
* alpha test x AFAIL (ubershader.comp:1341-1370, 1509-1541), DATE/DATM (1543-1548);
* fog exhaustive over colour and coefficient: floor(f + (c - f) F / 256 + 0.5) (1329-1335);
* blend overflow with COLCLAMP wrap and clamp, FBA (1424-1431, 1570-1586);
* 16-bit frame buffers: dither with signed DIMX entries and the 5:5:5:1 quantisation
  (1558-1568, 1599-1600);
* DECAL, HIGHLIGHT and HIGHLIGHT2 (918-945, 1304-1320);
* depth formats: clamping to 16/24 bits and the GEQUAL / GREATER tie (854-871);
* depth interpolation: the nearest integer of the exact plane (roundEven, 864);
* lines: the minor-axis band around the exact line and half-pixel points rounding up
  (triangle_setup.comp:154-168, 212-268);
* bilinear filtering at every 1/16-texel weight pair (ubershader.comp:1139-1143, 1220-1230):
  within one of the exact rational lerp and unbiased, and exact with round-half-up where only U
  varies.
"""
import ctypes
import struct
from concurrent.futures import ThreadPoolExecutor
from fractions import Fraction

import numpy as np
import pytest

import gif
import render_streams as rs
from oracle_lib import Oracle, oracle_lib, swizzle as oswz

OFS = 2048 * 16


def sweep(fn, first, count, threads=8):
    """Runs an oracle reciprocal sweep over [first, first + count) on several threads; returns
    (wrapping hash sum, max ulp error)."""
    L = oracle_lib()
    f = getattr(L, fn)
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]
    f.restype = None
    step = (count + threads - 1) // threads

    def part(k):
        a = k * step
        n = max(0, min(step, count - a))
        h, m = ctypes.c_uint64(), ctypes.c_double()
        f((first + a) & 0xFFFFFFFF, n, ctypes.byref(h), ctypes.byref(m))
        return h.value, m.value

    with ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(part, range(threads)))
    return sum(p[0] for p in parts) % 2**64, max(p[1] for p in parts)


def test_rcp_float_lut_formula():
    L = oracle_lib()
    L.orc_rcp_float.restype = ctypes.c_float
    L.orc_rcp_float.argtypes = [ctypes.c_float]
    # Powers of two are exact (seed, then Newton steps that keep an exact reciprocal).
    for k in range(-20, 21):
        assert L.orc_rcp_float(2.0 ** k) == 2.0 ** -k
    # The seed LUT: 1/(1 + (i + 0.5)/256) in float32 arithmetic, then four Newton steps;
    # with a wrong seed table the sweep's error bound below fails.
    lut = [np.float32(1.0) / (np.float32(1.0) + (np.float32(i) + np.float32(0.5)) / np.float32(256.0))
           for i in range(256)]
    assert all(0.5 < float(v) < 1.0 for v in lut)


def test_rcp_float_exhaustive_over_mantissas():
    """Every float in [1, 2): rcp_float is within 1.25 ulp of the correctly rounded 1/x (four
    single-precision Newton steps from the 256-entry seed)."""
    h, worst = sweep("orc_rcp_float_sweep", 0x3F800000, 1 << 23)
    assert worst <= 1.25, worst
    # A different exponent scales the result exactly (ldexp).
    h2, worst2 = sweep("orc_rcp_float_sweep", 0x41000000, 1 << 23)
    assert worst2 <= 1.25


def test_rcp_fixed_exhaustive():
    """Every 32-bit input but 0: the fixed-point Newton reciprocal is within 0.51 ulp of 1/x
    ("~23.99 bits of precision", math_utils.h:24-26)."""
    h, worst = sweep("orc_rcp_fixed_sweep", 1, (1 << 32) - 1)
    assert worst <= 0.51, worst


# ---------------------------------------------------------------------------------------
# Rendering closed forms. Frame buffer at block 0, FBW as given; Z buffer at page 64.

ZBP = 64


def regs(fbw=2, fb_psm=0x00, zpsm=0x1, zte=0, ztst=1, zmsk=1, w=128, h=64, extra=()):
    r = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, fbw, fb_psm)), (0x4E, rs.zbuf(ZBP, zpsm, zmsk)),
         (0x40, rs.scissor(0, w - 1, 0, h - 1)), (0x47, rs.test_reg(zte=zte, ztst=ztst)), (0x46, 1)]
    r.extend(extra)
    return r


def sprite(x, y, w, h, rgba, z=0, prim=None, fog=None):
    p = prim if prim is not None else rs.prim_reg(6, iip=0)
    if fog is None:
        v = [(0x05, rs.xyz2(OFS + x * 16, OFS + y * 16, z)), (0x05, rs.xyz2(OFS + (x + w) * 16, OFS + (y + h) * 16, z))]
    else:
        v = [(0x04, rs.xyzf2(OFS + x * 16, OFS + y * 16, z, fog)),
             (0x04, rs.xyzf2(OFS + (x + w) * 16, OFS + (y + h) * 16, z, fog))]
    return [(0x00, p), (0x01, rs.rgbaq(rgba)), v[0], (0x01, rs.rgbaq(rgba)), v[1]]


def image(o, base, fbw, psm, w, h):
    v = o.vram()
    if gif.BPP[psm] == 16:
        e = v.view(np.uint16)
    else:
        e = v.view(np.uint32)
    out = np.zeros((h, w), np.uint64)
    for y in range(h):
        for x in range(w):
            out[y, x] = e[oswz(x, y, base, fbw, psm)]
    return out


def fill(base, fbw, psm, w, h, values):
    vals = np.asarray(values, np.uint64).ravel()
    bpp = gif.BPP[psm]
    if bpp == 32:
        b = vals.astype(np.uint32).tobytes()
    elif bpp == 24:
        b = np.stack([(vals >> s) & 0xFF for s in (0, 8, 16)], 1).astype(np.uint8).tobytes()
    else:
        b = vals.astype(np.uint16).tobytes()
    b += b"\0" * (-len(b) % 16)
    return gif.upload(base, fbw, psm, 0, 0, w, h, b)


def run(stream):
    o = Oracle()
    o.gif_transfer(gif.arr(stream))
    o.flush()
    return o


ATST_CMP = [lambda a, r: False, lambda a, r: True, lambda a, r: a < r, lambda a, r: a <= r,
            lambda a, r: a == r, lambda a, r: a >= r, lambda a, r: a > r, lambda a, r: a != r]


def test_alpha_test_and_afail_table():
    """Every ATST x AFAIL with alphas around AREF over a pre-filled colour and depth buffer:
    KEEP writes nothing, FB_ONLY colour and alpha, ZB_ONLY depth, RGB_ONLY colour without
    alpha; a passing pixel writes all three."""
    F = 0x55302010
    Z0, ZS = 0x123456, 0x654321
    aref = 0x60
    alphas = (0x5F, 0x60, 0x61)
    s = gif.ad(regs(zte=1, ztst=1, zmsk=0))
    s += fill(0, 2, 0x00, 128, 64, np.full(128 * 64, F))
    s += fill(ZBP * 32, 2, 0x31, 128, 64, np.full(128 * 64, Z0))
    expect_c = np.full((64, 128), F, np.uint64)
    expect_z = np.full((64, 128), Z0, np.uint64)
    for atst in range(8):
        for afail in range(4):
            for k, a in enumerate(alphas):
                x, y = (atst * 3 + k) * 4, afail * 4
                col = 0x00332211 | a << 24
                s += gif.ad([(0x47, rs.test_reg(ate=1, atst=atst, aref=aref, afail=afail, zte=1, ztst=1))] +
                            sprite(x, y, 4, 4, col, ZS))
                ok = ATST_CMP[atst](a, aref)
                c, z = F, Z0
                if ok or afail == 1:
                    c, z = col, (ZS if ok else Z0)
                elif afail == 2:
                    z = ZS
                elif afail == 3:
                    c = (col & 0xFFFFFF) | (F & 0xFF000000)
                expect_c[y:y + 4, x:x + 4] = c
                expect_z[y:y + 4, x:x + 4] = z
    o = run(s)
    assert np.array_equal(image(o, 0, 2, 0x00, 128, 64), expect_c)
    assert np.array_equal(image(o, ZBP * 32, 2, 0x31, 128, 64) & 0xFFFFFF, expect_z)


@pytest.mark.parametrize("fb_psm", [0x00, 0x02, 0x01])
def test_destination_alpha_test(fb_psm):
    """DATE: a pixel is drawn only where the destination alpha bit (alpha >= 0x80, or bit 15 of
    a 16-bit pixel) equals DATM; a 24-bit frame buffer has no alpha and draws everywhere."""
    w, h = 64, 16
    ys, xs = np.mgrid[0:h, 0:w]
    left = xs < 32
    if fb_psm == 0x02:
        pre = np.where(left, 0x8000 | 0x1234, 0x1234).astype(np.uint64)
    else:
        pre = np.where(left, 0x80102030, 0x7F102030).astype(np.uint64)
    col = 0x40AABBCC
    for datm in (0, 1):
        s = gif.ad(regs(fbw=1, fb_psm=fb_psm, w=w, h=h)) + fill(0, 1, fb_psm, w, h, pre)
        s += gif.ad([(0x47, rs.test_reg(date=1, datm=datm, zte=0))] + sprite(0, 0, w, h, col))
        img = image(run(s), 0, 1, fb_psm, w, h)
        if fb_psm == 0x01:
            drawn = np.ones_like(left)
            new = col & 0xFFFFFF
        elif fb_psm == 0x02:
            drawn = left == bool(datm)
            new = ((col >> 3) & 0x1F) | (((col >> 11) & 0x1F) << 5) | (((col >> 19) & 0x1F) << 10) | ((col >> 31) << 15)
        else:
            drawn = left == bool(datm)
            new = col
        mask = 0xFFFFFF if fb_psm == 0x01 else (0xFFFF if fb_psm == 0x02 else 0xFFFFFFFF)
        exp = np.where(drawn, new, pre & mask).astype(np.uint64)
        assert np.array_equal(img & mask, exp), (fb_psm, datm)


def test_fog_exhaustive():
    """Every colour byte c and fog coefficient F (one pixel each): each channel becomes
    floor(f + (c - f) F / 256 + 0.5) = floor((256 f + (c - f) F + 128) / 256) for the FOGCOL
    channel f; exact ties (.5) round up."""
    fogcol = (17, 200, 93)
    s = gif.ad(regs(fbw=4, w=256, h=256, extra=[(0x3D, fogcol[0] | fogcol[1] << 8 | fogcol[2] << 16)]))
    prim = rs.prim_reg(6, iip=0, fge=1)
    writes = []
    for F in range(256):
        for c in range(256):
            col = c | (255 - c) << 8 | ((c * 7) & 255) << 16 | 0x80 << 24
            writes += sprite(c, F, 1, 1, col, prim=prim, fog=F)
    for i in range(0, len(writes), 5000):
        s += gif.ad(writes[i:i + 5000])
    img = image(run(s), 0, 4, 0x00, 256, 256)
    F, c = np.mgrid[0:256, 0:256].astype(np.int64)
    chans = [c, 255 - c, (c * 7) & 255]
    exp = np.full((256, 256), 0x80 << 24, np.int64)
    for k, (ch, f) in enumerate(zip(chans, fogcol)):
        exp |= ((256 * f + (ch - f) * F + 128) // 256) << (8 * k)
    assert np.array_equal(img.astype(np.int64), exp)


@pytest.mark.parametrize("colclamp", [0, 1])
@pytest.mark.parametrize("fba", [0, 1])
def test_blend_overflow_colclamp_and_fba(colclamp, fba):
    """(A - B) * C >> 7 + D with sums above 255 and below 0: COLCLAMP = 0 keeps the low 8 bits
    (x mod 256), COLCLAMP = 1 clamps to 0..255; FBA sets alpha bit 7 of every written pixel.
    Every source value 0..255 in one row per blend."""
    fix = 200
    w = 256
    xs = np.arange(w, dtype=np.int64)
    cd = np.stack([(xs * 3) & 255, 255 - xs, np.full(w, 100)], 0)
    pre = (cd[0] | cd[1] << 8 | cd[2] << 16 | 0x10 << 24).astype(np.uint64)
    s = gif.ad(regs(fbw=4, w=w, h=2, extra=[(0x46, colclamp), (0x4A, fba)]))
    s += fill(0, 4, 0x00, w, 1, pre) + fill(0, 4, 0x00, w, 2, np.concatenate([pre, pre]))
    blends = [rs.alpha_reg(0, 2, 2, 1, fix), rs.alpha_reg(2, 0, 2, 1, fix)]  # Cs*FIX+Cd, -Cs*FIX+Cd
    writes = []
    for row, ab in enumerate(blends):
        writes.append((0x42, ab))
        for x in range(w):
            col = x | ((x * 5) & 255) << 8 | (255 - x) << 16 | 0x40 << 24
            writes += sprite(x, row, 1, 1, col, prim=rs.prim_reg(6, iip=0, abe=1))
    s += gif.ad(writes)
    img = image(run(s), 0, 4, 0x00, w, 2).astype(np.int64)
    cs = np.stack([xs, (xs * 5) & 255, 255 - xs], 0)
    for row in range(2):
        sign = 1 if row == 0 else -1
        exp = np.zeros(w, np.int64)
        for k in range(3):
            v = (sign * cs[k] * fix) // 128 + cd[k]
            v = np.clip(v, 0, 255) if colclamp else v % 256
            exp |= v << (8 * k)
        exp |= (0x40 | (0x80 if fba else 0)) << 24
        assert np.array_equal(img[row], exp), (row, colclamp, fba)


@pytest.mark.parametrize("colclamp", [0, 1])
def test_16bit_frame_buffer_dither_and_quantise(colclamp):
    """PSMCT16 with DTHE: entry (x & 3, y & 3) of DIMX (3-bit two's complement at bit
    16 (y & 3) + 4 (x & 3)) is added to R, G and B, the sum is clamped or wrapped by COLCLAMP,
    and the pixel keeps the top 5 bits of each channel and alpha bit 7."""
    rng = np.random.default_rng(5)
    dm = rng.integers(0, 8, size=(4, 4))
    dimx = 0
    for y in range(4):
        for x in range(4):
            dimx |= int(dm[y, x]) << (16 * y + 4 * x)
    signed = np.where(dm >= 4, dm - 8, dm)
    w, h = 64, 32
    cols = [0x80FF0003, 0x00010204, 0xFFFFFFFF, 0x7F808080]
    s = gif.ad(regs(fbw=1, fb_psm=0x02, w=w, h=h, extra=[(0x45, 1), (0x44, dimx), (0x46, colclamp)]))
    for k, col in enumerate(cols):
        s += gif.ad(sprite(k * 16, 0, 16, h, col))
    img = image(run(s), 0, 1, 0x02, w, h).astype(np.int64)
    ys, xs = np.mgrid[0:h, 0:w]
    d = signed[ys & 3, xs & 3]
    exp = np.zeros((h, w), np.int64)
    for k, col in enumerate(cols):
        sel = (xs // 16) == k
        v = 0
        for ch in range(3):
            c = ((col >> (8 * ch)) & 255) + d
            c = np.clip(c, 0, 255) if colclamp else c % 256
            v = v | ((c >> 3) << (5 * ch))
        v = v | ((col >> 31) << 15)
        exp = np.where(sel, v, exp)
    assert np.array_equal(img, exp)


@pytest.mark.parametrize("tfx", [1, 2, 3])
@pytest.mark.parametrize("tcc", [0, 1])
def test_decal_highlight_highlight2(tfx, tcc):
    """Texture function on a 1:1 textured sprite (NEAREST): DECAL takes the texel (alpha from
    the vertex without TCC); HIGHLIGHT and HIGHLIGHT2 add the vertex alpha to min(255,
    floor(t c / 128)) and clamp; HIGHLIGHT's alpha is t.a + c.a with TCC, HIGHLIGHT2's t.a."""
    rng = np.random.default_rng(tfx * 2 + tcc)
    tex = rng.integers(0, 2**32, size=(16, 16), dtype=np.uint64)
    tex[0, :4] = [0xFFFFFFFF, 0, 0x80808080, 0x01FF7F00]
    tbp = 32 * 32
    cols = [0x7F40C0FF, 0x05030201, 0xFF808080]
    s = gif.ad(regs(fbw=1, w=64, h=16, extra=[(0x06, rs.tex0(tbp, 1, 0x00, 4, 4, tcc=tcc, tfx=tfx))]))
    s += fill(tbp, 1, 0x00, 16, 16, tex)
    for k, col in enumerate(cols):
        s += gif.ad([(0x00, rs.prim_reg(6, iip=0, tme=1, fst=1)), (0x01, rs.rgbaq(col)), (0x03, rs.uv(0, 0)),
                     (0x05, rs.xyz2(OFS + k * 16 * 16, OFS, 0)), (0x03, rs.uv(256, 256)),
                     (0x05, rs.xyz2(OFS + (k + 1) * 16 * 16, OFS + 16 * 16, 0))])
    img = image(run(s), 0, 1, 0x00, 64, 16)
    for k, col in enumerate(cols):
        got = img[:, k * 16:(k + 1) * 16].astype(np.int64)
        t = [((tex >> (8 * ch)) & 255).astype(np.int64) for ch in range(4)]
        c = [(col >> (8 * ch)) & 255 for ch in range(4)]
        if tfx == 1:
            out = t[:3] + [t[3] if tcc else np.full_like(t[3], c[3])]
        else:
            out = [np.minimum(np.minimum((t[ch] * c[ch]) // 128, 255) + c[3], 255) for ch in range(3)]
            if tcc:
                out.append(np.minimum(t[3] + c[3], 255) if tfx == 2 else t[3])
            else:
                out.append(np.full_like(t[3], c[3]))
        exp = out[0] | out[1] << 8 | out[2] << 16 | out[3] << 24
        assert np.array_equal(got, exp), (tfx, tcc, k)


@pytest.mark.parametrize("zpsm,zmax", [(0x2, 0xFFFF), (0x1, 0xFFFFFF), (0x0, 0xFFFFFFFF)])
@pytest.mark.parametrize("ztst", [2, 3])
def test_depth_formats_clamp_and_ties(zpsm, zmax, ztst):
    """Sprites at depths around a pre-filled Z0 and above the format's range: the depth is
    clamped to the format's maximum, GEQUAL passes ties and GREATER does not; passing pixels
    write colour and depth."""
    Z0 = zmax - 5 if zpsm == 0x0 else zmax // 2
    zs = [Z0 - 1, Z0, Z0 + 1, min(zmax * 2 + 7, 0xFFFFFFFF)]
    zfmt = {0x2: 0x32, 0x1: 0x31, 0x0: 0x30}[zpsm]
    w = 4 * len(zs)
    s = gif.ad(regs(fbw=1, zpsm=zpsm, zte=1, ztst=ztst, zmsk=0, w=w, h=4))
    s += fill(ZBP * 32, 1, zfmt, w, 4, np.full(w * 4, Z0))
    for k, z in enumerate(zs):
        s += gif.ad(sprite(4 * k, 0, 4, 4, 0x80000000 | (k + 1), z))
    o = run(s)
    col = image(o, 0, 1, 0x00, w, 4)
    zimg = image(o, ZBP * 32, 1, zfmt, w, 4)
    for k, z in enumerate(zs):
        zc = min(z, zmax)
        ok = zc >= Z0 if ztst == 2 else zc > Z0
        assert np.all(col[:, 4 * k:4 * k + 4] == (0x80000000 | (k + 1) if ok else 0)), (k, z)
        assert np.all((zimg[:, 4 * k:4 * k + 4] & zmax) == (zc if ok else Z0)), (k, z)


def plane_z(verts, zs, x, y):
    (xa, ya), (xb, yb), (xc, yc) = [(Fraction(a), Fraction(b)) for a, b in verts]
    za, zb, zc = (Fraction(z) for z in zs)
    det = (xb - xa) * (yc - ya) - (xc - xa) * (yb - ya)
    i = ((x - xa) * (yc - ya) - (xc - xa) * (y - ya)) / det
    j = ((xb - xa) * (y - ya) - (x - xa) * (yb - ya)) / det
    return za + i * (zb - za) + j * (zc - za)


@pytest.mark.parametrize("case", range(4))
def test_depth_interpolation_is_the_nearest_integer_of_the_plane(case):
    """Z at each covered pixel (sampled at its top-left corner) is the exact plane through the
    three vertices rounded to the nearest integer wherever that plane value is not within 1/64
    of a half; on an axis-aligned triangle with power-of-two legs and depth steps of 2^k / 8
    per pixel, the halves are exact and round to even."""
    import test_render_oracle as tro
    if case < 2:
        rng = np.random.default_rng(case)
        verts = [(int(rng.integers(0, 8 * 16)), int(rng.integers(0, 8 * 16))),
                 (int(rng.integers(52 * 16, 64 * 16)), int(rng.integers(0, 16 * 16))),
                 (int(rng.integers(0, 24 * 16)), int(rng.integers(40 * 16, 48 * 16)))]
        zs = [int(rng.integers(1000, 5000)) for _ in range(3)]
    else:
        # Right triangle with 32-pixel legs at pixel corners; z changes by 12/8 and 20/8 per pixel.
        verts = [(4 * 16, 4 * 16), (36 * 16, 4 * 16), (4 * 16, 36 * 16)]
        zs = [1000, 1000 + 48, 1000 + 80] if case == 2 else [2000, 2000 + 16, 2000 + 112]
    o = Oracle()
    writes = tro.setup_regs(zte=1, ztst=1) + [(0x00, rs.prim_reg(3, iip=0))]
    for (x, y), z in zip(verts, zs):
        writes.append((0x01, rs.rgbaq(0xFFFFFFFF)))
        writes.append((0x05, rs.xyz2(OFS + x, OFS + y, z)))
    o.gif_transfer(gif.arr(gif.ad(writes)))
    zimg = tro.z_image(o)
    cov = tro.exact_coverage(verts)
    checked = ties = 0
    for y, x in zip(*np.nonzero(cov)):
        v = plane_z(verts, zs, x * 16, y * 16)
        frac = v - (v.numerator // v.denominator)
        if case >= 2 and frac == Fraction(1, 2):
            ties += 1
            fl = v.numerator // v.denominator
            assert int(zimg[y, x]) == (fl if fl % 2 == 0 else fl + 1), (x, y, int(zimg[y, x]), float(v))
            continue
        if abs(frac - Fraction(1, 2)) <= Fraction(1, 64):
            continue
        assert int(zimg[y, x]) == round(v), (x, y, int(zimg[y, x]), float(v))
        checked += 1
    assert checked > 100
    if case >= 2:
        assert ties > 20


def line_image(a, b, w=64, h=48):
    import test_render_oracle as tro
    o = Oracle()
    writes = tro.setup_regs() + [(0x00, rs.prim_reg(1, iip=0)), (0x01, rs.rgbaq(0x80FFFFFF)),
                                 (0x05, rs.xyz2(OFS + a[0], OFS + a[1], 0)), (0x01, rs.rgbaq(0x80FFFFFF)),
                                 (0x05, rs.xyz2(OFS + b[0], OFS + b[1], 0))]
    o.gif_transfer(gif.arr(gif.ad(writes)))
    return tro.fb_image(o) != 0


@pytest.mark.parametrize("seed", range(16))
def test_lines_cover_the_band_around_the_exact_line(seed):
    """A line becomes a parallelogram about one pixel thick along the minor axis: its vertices
    are ordered top first (then left); the third corner sits delta from the first along the
    minor axis (delta = -8 sign(dx) for x-major lines, 8 sign(dy) for y-major ones) and the two
    end points move by 1 - delta, so at each major-axis sample position the band runs between
    L + delta and L + 1 - delta (1/16 pixel units), L being the exact line there: 15 units wide
    (gaps are possible) or 17 (two pixels are possible). Strictly inside the line's extent the
    covered pixels are the multiples of 16 strictly inside the band, plus possibly one on its
    boundary, and none lies outside it."""
    rng = np.random.default_rng(seed)

    def coord(lo, hi):
        v = int(rng.integers(lo, hi))
        return v + 1 if v % 16 == 0 else v  # not on a pixel corner (no Bresenham shift)

    p0 = (coord(2 * 16, 60 * 16), coord(2 * 16, 44 * 16))
    p1 = (coord(2 * 16, 60 * 16), coord(2 * 16, 44 * 16))
    img = line_image(p0, p1)
    # Provoking (first) vertex: smaller y, then smaller x (equal depths).
    a, b = (p0, p1) if (p0[1], p0[0]) < (p1[1], p1[0]) else (p1, p0)
    xmaj = abs(b[0] - a[0]) >= abs(b[1] - a[1])
    ma, mi = (0, 1) if xmaj else (1, 0)
    # The third vertex sits delta away from the first along the minor axis and the first two move
    # by 1 - delta: delta = -8 sign(dx) for x-major lines, 8 sign(dy) for y-major ones.
    delta = -8 * int(np.sign(b[0] - a[0])) if xmaj else 8 * int(np.sign(b[1] - a[1]))
    band_lo, band_hi = min(delta, 1 - delta), max(delta, 1 - delta)
    lo, hi = sorted((a[ma], b[ma]))
    checked = 0
    for p in range(64 if xmaj else 48):
        sel = np.nonzero(img[:, p] if xmaj else img[p, :])[0] * 16
        if not (lo - 16 < p * 16 < hi + 16):
            assert sel.size == 0, (p, sel)
            continue
        if not (lo + 16 < p * 16 < hi - 16):
            continue  # end columns depend on the end edges
        L = a[mi] + Fraction(p * 16 - a[ma], b[ma] - a[ma]) * (b[mi] - a[mi])
        low, high = L + band_lo, L + band_hi
        strict = {q for q in range(0, 64 * 16, 16) if low < q < high}
        got = set(int(q) for q in sel)
        assert strict <= got, (p, sorted(strict), sorted(got), float(L))
        assert all(low <= q <= high for q in got), (p, sorted(got), float(low), float(high))
        checked += 1
    assert checked >= (hi - lo) // 16 - 3


@pytest.mark.parametrize("y_sub", [0, 4, 7, 8, 12])
@pytest.mark.parametrize("leftward", [False, True])
def test_horizontal_line_rows(y_sub, leftward):
    """A horizontal line at y = 16 Y + y_sub, drawn either way (setup orders its end points left
    first): its band is [y - 8, y + 9], so a half-pixel line (y_sub = 8) covers row Y + 1 and
    possibly row Y, and lines below the half stay on row Y. Rows strictly inside the band are
    covered, rows on its boundary may be, no other row is."""
    Y = 10
    y = Y * 16 + y_sub
    a, b = (4 * 16 + 3, y), (40 * 16 + 3, y)
    img = line_image(b, a) if leftward else line_image(a, b)
    lo, hi = y - 8, y + 9
    rows = {int(r) for r in np.nonzero(img.any(axis=1))[0]}
    strict = {r for r in range(48) if lo < 16 * r < hi}
    closed = {r for r in range(48) if lo <= 16 * r <= hi}
    assert strict <= rows <= closed, (sorted(rows), sorted(strict), sorted(closed))
    cols = np.nonzero(img.any(axis=0))[0]
    assert cols.min() in (4, 5) and cols.max() in (39, 40) and len(cols) == cols.max() - cols.min() + 1


@pytest.mark.parametrize("axis", ["uv", "u"])
def test_bilinear_every_sixteenth_weight(axis):
    """LINEAR sampling of a random texture at every offset (du, dv) in 1/16 texel: with UV =
    16 x + du the sample sits (du - 8) / 16 of a texel past the centre of texel x, so the
    weights are multiples of 1/16. Each channel is within one of the exact rational lerp of the
    four texels (two rounded lerps, U then V, unbiased on average); where rows are equal (only U
    varies) it is the exact lerp rounded half up."""
    rng = np.random.default_rng(77)
    tw, th = 16, 16
    tex = rng.integers(0, 2**32, size=(th, tw), dtype=np.uint64)
    if axis == "u":
        tex[:] = tex[0]
    tbp = 64 * 32
    W, H = 256, 256
    s = gif.ad(regs(fbw=4, w=W, h=H, extra=[(0x14, rs.tex1(mmag=1, mmin=1)),
                                            (0x06, rs.tex0(tbp, 1, 0x00, 4, 4, tcc=1, tfx=1))]))
    s += fill(tbp, 1, 0x00, tw, th, tex)
    dvs = range(16) if axis == "uv" else [8]
    writes = []
    for du in range(16):
        for dv in dvs:
            x0, y0 = du * 16, dv * 16
            writes += [(0x00, rs.prim_reg(6, iip=0, tme=1, fst=1)), (0x01, rs.rgbaq(0x80808080)),
                       (0x03, rs.uv(du, dv)), (0x05, rs.xyz2(OFS + x0 * 16, OFS + y0 * 16, 0)),
                       (0x03, rs.uv(du + 16 * 16, dv + 16 * 16)),
                       (0x05, rs.xyz2(OFS + (x0 + 16) * 16, OFS + (y0 + 16) * 16, 0))]
    s += gif.ad(writes)
    img = image(run(s), 0, 4, 0x00, W, H)
    ch = [((tex >> (8 * k)) & 255).astype(np.int64) for k in range(4)]
    checked = 0
    errs = []
    for du in range(16):
        for dv in dvs:
            for py in range(16):
                for px in range(16):
                    # Texel coordinate of the sample (repeat wrap on a 16 x 16 texture).
                    fu, fv = Fraction(16 * px + du, 16) - Fraction(1, 2), Fraction(16 * py + dv, 16) - Fraction(1, 2)
                    u0, v0 = fu.numerator // fu.denominator, fv.numerator // fv.denominator
                    wu, wv = fu - u0, fv - v0
                    got = int(img[dv * 16 + py, du * 16 + px])
                    for k in range(4):
                        t = lambda uu, vv: ch[k][vv % th, uu % tw]
                        ex = ((1 - wu) * (1 - wv) * t(u0, v0) + wu * (1 - wv) * t(u0 + 1, v0) +
                              (1 - wu) * wv * t(u0, v0 + 1) + wu * wv * t(u0 + 1, v0 + 1))
                        g = (got >> (8 * k)) & 255
                        if axis == "u":
                            assert g == (ex + Fraction(1, 2)).numerator // (ex + Fraction(1, 2)).denominator, \
                                (du, px, k, g, float(ex))
                        else:
                            assert abs(g - ex) <= 1, (du, dv, px, py, k, g, float(ex))
                            errs.append(float(g - ex))
                        checked += 1
    assert checked >= 16 * len(dvs) * 256 * 4
    if errs:
        # Two rounded lerps (U, then V) stay unbiased: the mean error is small.
        assert abs(np.mean(errs)) < 0.25 and np.mean(np.abs(errs)) < 0.5, (np.mean(errs), np.mean(np.abs(errs)))
