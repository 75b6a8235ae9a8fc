"""CPU oracle texturing pinned by closed forms (no GPU).

A 1:1 sprite (pixel (x, y) at UV (16x, 16y)) with NEAREST sampling and DECAL/TCC=1 must
reproduce the decoded texture exactly, so each expected image below is built directly from the
uploaded texel data: raw 32-bit texels, 24-bit texels with TEXA alpha, 16-bit texels expanded
per the GS TEXA rules, and palette indices looked up in a CLUT laid out by the GS CSM1/CSM2
arrangement (written here from the layout description, not from the oracle's formulas).
"""
import struct

import numpy as np
import pytest

import gif
import render_streams as rs
from oracle_lib import Oracle, swizzle as oswz

OFS = 2048 * 16
FW, FH = 128, 64          # frame: 128 x 64 PSMCT32 at block 0, FBW = 2
TBP, CBP, CBP2 = 1024, 2048, 2112


def base_regs(extra=()):
    regs = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 2, 0x00)), (0x4E, rs.zbuf(256, 0x1, 1)),
            (0x40, rs.scissor(0, FW - 1, 0, FH - 1)), (0x47, rs.test_reg(zte=0)), (0x46, 1)]
    regs.extend(extra)
    return regs


def pack(psm, pixels):
    p = np.asarray(pixels, dtype=np.uint64).ravel()
    bpp = gif.BPP[psm]
    if bpp == 32:
        b = p.astype(np.uint32).tobytes()
    elif bpp == 24:
        b = np.stack([(p >> s) & 0xFF for s in (0, 8, 16)], axis=1).astype(np.uint8).tobytes()
    elif bpp == 16:
        b = p.astype(np.uint16).tobytes()
    elif bpp == 8:
        b = p.astype(np.uint8).tobytes()
    else:
        b = (p[0::2] | (p[1::2] << 4)).astype(np.uint8).tobytes()
    return b + b"\0" * (-len(b) % 16)


def upload(bp, bw, psm, w, h, pixels, x=0, y=0):
    return gif.upload(bp, bw, psm, x, y, w, h, pack(psm, pixels))


def sprite(x0, y0, x1, y1, u0, v0, u1, v1, rgba=0x80808080, prim_extra=None):
    prim = rs.prim_reg(6, iip=0, tme=1, fst=1)
    return [(0x00, prim), (0x01, rs.rgbaq(rgba)), (0x03, rs.uv(u0, v0)), (0x05, rs.xyz2(OFS + x0, OFS + y0, 0)),
            (0x03, rs.uv(u1, v1)), (0x05, rs.xyz2(OFS + x1, OFS + y1, 0))]


def frame_image(o, w=FW, h=FH):
    w32 = o.vram().view(np.uint32)
    return np.array([[w32[oswz(x, y, 0, 2, 0x00)] for x in range(w)] for y in range(h)], dtype=np.uint64)


def run(stream_bytes):
    o = Oracle()
    o.gif_transfer(gif.arr(stream_bytes))
    o.flush()
    return o


def expand16(c, aem=0, ta0=0, ta1=0x80):
    """GS 16-bit texel to RGBA8: 5-bit channels shifted up by 3, alpha from TA0/TA1 by bit 15,
    and AEM forcing alpha 0 for an all-zero texel."""
    c = np.asarray(c, dtype=np.uint64)
    r, g, b = (c & 31) << 3, ((c >> 5) & 31) << 3, ((c >> 10) & 31) << 3
    a = np.where((c >> 15) & 1, ta1, ta0).astype(np.uint64)
    if aem:
        a = np.where(c == 0, 0, a).astype(np.uint64)
    return r | g << 8 | b << 16 | a << 24


def draw_identity(psm, texels, tw, th, extra_regs=(), clut=b"", sprite_w=None, sprite_h=None, tex_regs=None):
    w, h = 1 << tw, 1 << th
    sw, sh = sprite_w or w, sprite_h or h
    tbw = max(1, w // 64)
    if psm in (0x13, 0x14):
        tbw = max(2, tbw + (tbw & 1))
    regs = base_regs(extra_regs)
    regs += tex_regs if tex_regs is not None else [(0x06, rs.tex0(TBP, tbw, psm, tw, th))]
    s = gif.ad(regs[:7]) + clut + upload(TBP, tbw, psm, w, h, texels) + gif.ad(regs[7:])
    s += gif.ad(sprite(0, 0, sw * 16, sh * 16, 0, 0, sw * 16, sh * 16))
    return run(s)


@pytest.mark.parametrize("tw,th", [(6, 5), (5, 6), (3, 4)])
def test_nearest_identity_ct32(tw, th):
    rng = np.random.default_rng(tw * 16 + th)
    w, h = 1 << tw, 1 << th
    tex = rng.integers(0, 2**32, size=(h, w), dtype=np.uint64)
    o = draw_identity(0x00, tex, tw, th)
    img = frame_image(o)
    assert np.array_equal(img[:h, :w], tex)
    assert np.all(img[h:, :] == 0) and np.all(img[:, w:] == 0)


@pytest.mark.parametrize("aem", [0, 1])
def test_nearest_identity_ct24_texa(aem):
    rng = np.random.default_rng(7 + aem)
    tex = rng.integers(0, 2**24, size=(32, 64), dtype=np.uint64)
    tex[3, 5] = 0
    ta0 = 0x5A
    o = draw_identity(0x01, tex, 6, 5, extra_regs=[(0x3B, rs.texa(ta0=ta0, aem=aem))])
    img = frame_image(o)[:32, :64]
    alpha = np.where((tex == 0) & bool(aem), 0, ta0).astype(np.uint64)
    assert np.array_equal(img, tex | alpha << 24)


@pytest.mark.parametrize("psm,aem", [(0x02, 0), (0x02, 1), (0x0A, 1)])
def test_nearest_identity_ct16_texa(psm, aem):
    rng = np.random.default_rng(psm + aem)
    tex = rng.integers(0, 2**16, size=(64, 64), dtype=np.uint64)
    tex[0, :4] = [0, 0x8000, 0x7FFF, 0xFFFF]
    o = draw_identity(psm, tex, 6, 6, extra_regs=[(0x3B, rs.texa(ta0=0x11, aem=aem, ta1=0xEE))])
    img = frame_image(o)[:64, :64]
    assert np.array_equal(img, expand16(tex, aem, 0x11, 0xEE))


def csm1_clut_upload(bp, colours32):
    """Writes 256 CT32 palette entries as a 16x16 CSM1 CLUT rectangle at block `bp`."""
    img = np.zeros((16, 16), dtype=np.uint64)
    for i, c in enumerate(colours32):
        x, y = rs.csm1_position(i)
        img[y, x] = c
    return upload(bp, 1, 0x00, 16, 16, img)


def test_palette_t8_csm1_ct32():
    rng = np.random.default_rng(11)
    pal = rng.integers(0, 2**32, size=256, dtype=np.uint64)
    idx = rng.integers(0, 256, size=(64, 128), dtype=np.uint64)
    regs = [(0x06, rs.tex0(TBP, 2, 0x13, 7, 6, cbp=CBP, cpsm=0x00, csm=0, cld=1))]
    o = draw_identity(0x13, idx, 7, 6, clut=csm1_clut_upload(CBP, pal), tex_regs=regs)
    img = frame_image(o)[:64, :128]
    assert np.array_equal(img, pal[idx])


def test_palette_t4_ct16_csa_banks():
    """Two 16-colour CT16 palettes loaded into CSA banks 1 and 2; drawing with either bank
    selects its palette, and the second load leaves the first bank intact."""
    rng = np.random.default_rng(12)
    pal1 = rng.integers(0, 2**16, size=16, dtype=np.uint64)
    pal2 = rng.integers(0, 2**16, size=16, dtype=np.uint64)
    idx = rng.integers(0, 16, size=(64, 64), dtype=np.uint64)

    def clut8x2(p):
        img = np.zeros((2, 8), dtype=np.uint64)
        for i, c in enumerate(p):
            img[i // 8, i % 8] = c
        return img

    s = gif.ad(base_regs([(0x3B, rs.texa(ta0=0, ta1=0x80))]))
    s += upload(CBP, 1, 0x02, 8, 2, clut8x2(pal1)) + upload(CBP2, 1, 0x02, 8, 2, clut8x2(pal2))
    s += upload(TBP, 2, 0x14, 64, 64, idx)
    s += gif.ad([(0x06, rs.tex0(TBP, 2, 0x14, 6, 6, cbp=CBP, cpsm=0x02, csa=1, cld=1))])
    s += gif.ad([(0x06, rs.tex0(TBP, 2, 0x14, 6, 6, cbp=CBP2, cpsm=0x02, csa=2, cld=1))])
    s += gif.ad(sprite(0, 0, 64 * 16, 64 * 16, 0, 0, 64 * 16, 64 * 16))
    s += gif.ad([(0x06, rs.tex0(TBP, 2, 0x14, 6, 6, cbp=CBP, cpsm=0x02, csa=1, cld=0))])
    s += gif.ad(sprite(64 * 16, 0, 128 * 16, 64 * 16, 0, 0, 64 * 16, 64 * 16))
    o = run(s)
    img = frame_image(o)
    assert np.array_equal(img[:, :64], expand16(pal2[idx], 0, 0, 0x80))
    assert np.array_equal(img[:, 64:], expand16(pal1[idx], 0, 0, 0x80))


def test_palette_csm2_texclut_offsets():
    """CSM2: a 16-bit palette read as a plain row of the CLUT buffer at (COU*16, COV)."""
    rng = np.random.default_rng(13)
    pal = rng.integers(0, 2**16, size=256, dtype=np.uint64)
    row = np.zeros((8, 320), dtype=np.uint64)
    row[5, 32:288] = pal
    idx = rng.integers(0, 256, size=(32, 64), dtype=np.uint64)
    s = gif.ad(base_regs([(0x1C, rs.texclut(cbw=5, cou=2, cov=5)), (0x3B, rs.texa(ta0=0, ta1=0x80))]))
    s += upload(CBP, 5, 0x02, 320, 8, row)
    s += upload(TBP, 2, 0x13, 64, 32, idx)
    s += gif.ad([(0x06, rs.tex0(TBP, 2, 0x13, 6, 5, cbp=CBP, cpsm=0x02, csm=1, cld=1))])
    s += gif.ad(sprite(0, 0, 64 * 16, 32 * 16, 0, 0, 64 * 16, 32 * 16))
    img = frame_image(run(s))[:32, :64]
    assert np.array_equal(img, expand16(pal[idx], 0, 0, 0x80))


def test_cld_cached_cbp():
    """CLD=2 loads and records CBP0; CLD=4 with the same CBP does not reload even though the
    palette memory changed; CLD=4 with another CBP does."""
    pal_a = np.arange(256, dtype=np.uint64) * 0x01010101
    pal_b = (255 - np.arange(256, dtype=np.uint64)) * 0x00010001
    idx = np.tile(np.arange(64, dtype=np.uint64), (32, 1))
    t0 = lambda cbp, cld: rs.tex0(TBP, 2, 0x13, 6, 5, cbp=cbp, cpsm=0x00, cld=cld)
    s = gif.ad(base_regs()) + csm1_clut_upload(CBP, pal_a) + upload(TBP, 2, 0x13, 64, 32, idx)
    s += gif.ad([(0x06, t0(CBP, 2))])
    s += csm1_clut_upload(CBP, pal_b) + csm1_clut_upload(CBP2, pal_b)
    s += gif.ad([(0x06, t0(CBP, 4))])
    s += gif.ad(sprite(0, 0, 64 * 16, 32 * 16, 0, 0, 64 * 16, 32 * 16))
    s += gif.ad([(0x06, t0(CBP2, 4))])
    s += gif.ad(sprite(64 * 16, 0, 128 * 16, 32 * 16, 0, 0, 64 * 16, 32 * 16))
    img = frame_image(run(s))
    assert np.array_equal(img[:32, :64], pal_a[idx])
    assert np.array_equal(img[:32, 64:], pal_b[idx])


@pytest.mark.parametrize("mode", ["repeat", "clamp", "region_clamp", "region_clamp_reversed"])
def test_wrap_modes(mode):
    rng = np.random.default_rng(21)
    tex = rng.integers(0, 2**32, size=(16, 16), dtype=np.uint64)
    cl = {"repeat": rs.clamp_reg(0, 0), "clamp": rs.clamp_reg(1, 1),
          "region_clamp": rs.clamp_reg(2, 2, minu=2, maxu=9, minv=3, maxv=12),
          # MINU > MAXU: clamp(u, min, max) = min(max(u, min), max) is max everywhere.
          "region_clamp_reversed": rs.clamp_reg(2, 2, minu=9, maxu=2, minv=12, maxv=3)}[mode]
    s = gif.ad(base_regs([(0x08, cl), (0x06, rs.tex0(TBP, 1, 0x00, 4, 4))])) + upload(TBP, 1, 0x00, 16, 16, tex)
    s += gif.ad(sprite(0, 0, 64 * 16, 64 * 16, 0, 0, 64 * 16, 64 * 16))
    img = frame_image(run(s))[:64, :64]
    ys, xs = np.mgrid[0:64, 0:64]
    if mode == "repeat":
        ex = tex[ys % 16, xs % 16]
    elif mode == "clamp":
        ex = tex[np.minimum(ys, 15), np.minimum(xs, 15)]
    elif mode == "region_clamp":
        ex = tex[np.clip(ys, 3, 12), np.clip(xs, 2, 9)]
    else:
        ex = np.broadcast_to(tex[3, 2], (64, 64))
    assert np.array_equal(img, ex)


def test_modulate_closed_form():
    rng = np.random.default_rng(22)
    tex = rng.integers(0, 2**32, size=(32, 64), dtype=np.uint64)
    col = 0x7F40C0FF  # vertex colour: R=0xFF G=0xC0 B=0x40 A=0x7F
    regs = base_regs([(0x06, rs.tex0(TBP, 1, 0x00, 6, 5, tcc=1, tfx=0))])
    s = gif.ad(regs) + upload(TBP, 1, 0x00, 64, 32, tex)
    s += gif.ad(sprite(0, 0, 64 * 16, 32 * 16, 0, 0, 64 * 16, 32 * 16, rgba=col))
    img = frame_image(run(s))[:32, :64]
    ex = np.zeros_like(tex)
    for ch in range(4):
        t = (tex >> (8 * ch)) & 0xFF
        c = (col >> (8 * ch)) & 0xFF
        ex |= np.minimum((t * c) >> 7, 255) << (8 * ch)
    assert np.array_equal(img, ex)


def test_bilinear_at_texel_centres_is_exact():
    """LINEAR sampling at texel centres (UV offset by half a texel) returns the texels."""
    rng = np.random.default_rng(23)
    tex = rng.integers(0, 2**32, size=(32, 64), dtype=np.uint64)
    regs = base_regs([(0x14, rs.tex1(mmag=1, mmin=1)), (0x06, rs.tex0(TBP, 1, 0x00, 6, 5))])
    s = gif.ad(regs) + upload(TBP, 1, 0x00, 64, 32, tex)
    s += gif.ad(sprite(0, 0, 64 * 16, 32 * 16, 8, 8, 64 * 16 + 8, 32 * 16 + 8))
    img = frame_image(run(s))[:32, :64]
    assert np.array_equal(img, tex)


def test_bilinear_midpoints_of_linear_ramp():
    """At texel corners the bilinear filter averages four texels; on a ramp whose channels grow
    by 4 per texel in x and by 8 per texel in y the average is an exact integer."""
    ys, xs = np.mgrid[0:32, 0:64]
    r = (4 * xs) & 0xFF
    g = (8 * ys) & 0xFF
    tex = (r | g << 8 | 0x40 << 16 | 0x80 << 24).astype(np.uint64)
    regs = base_regs([(0x14, rs.tex1(mmag=1, mmin=1)), (0x06, rs.tex0(TBP, 1, 0x00, 6, 5))])
    s = gif.ad(regs) + upload(TBP, 1, 0x00, 64, 32, tex)
    s += gif.ad(sprite(0, 0, 64 * 16, 32 * 16, 0, 0, 64 * 16, 32 * 16))
    img = frame_image(run(s))[:32, :64]
    inner = (xs >= 1) & (xs < 64) & (ys >= 1) & (xs * 4 < 256) & (ys * 8 < 256)
    exp_r = 4 * xs - 2
    exp_g = 8 * ys - 4
    got_r, got_g = img & 0xFF, (img >> 8) & 0xFF
    assert np.array_equal(got_r[inner], exp_r[inner].astype(np.uint64))
    assert np.array_equal(got_g[inner], exp_g[inner].astype(np.uint64))
    assert np.all(((img >> 16) & 0xFF) == 0x40) and np.all((img >> 24) == 0x80)


@pytest.mark.parametrize("q", [1.0, 2.0, 0.5])
def test_perspective_constant_q_matches_uv(q):
    """ST/Q with a constant Q samples the same texels as UV = (S/Q, T/Q) * texture size."""
    rng = np.random.default_rng(24)
    tex = rng.integers(0, 2**32, size=(32, 64), dtype=np.uint64)
    regs = base_regs([(0x06, rs.tex0(TBP, 1, 0x00, 6, 5))])
    prim = rs.prim_reg(6, iip=0, tme=1, fst=0)
    s = gif.ad(regs) + upload(TBP, 1, 0x00, 64, 32, tex)
    s += gif.ad([(0x00, prim), (0x02, rs.st(0.0, 0.0)), (0x01, rs.rgbaq(0x80808080, q)),
                 (0x05, rs.xyz2(OFS, OFS, 0)), (0x02, rs.st(1.0 * q, 1.0 * q)), (0x01, rs.rgbaq(0x80808080, q)),
                 (0x05, rs.xyz2(OFS + 64 * 16, OFS + 32 * 16, 0))])
    img = frame_image(run(s))[:32, :64]
    assert np.array_equal(img, tex)


def test_texture_reads_see_pending_primitives():
    """A sprite textured from frame-buffer pages drawn earlier in the same pass samples the
    drawn pixels: the pass is rendered before the texture is read."""
    red = 0x800000FF
    s = gif.ad(base_regs())
    s += gif.ad([(0x00, rs.prim_reg(6, iip=0)), (0x01, rs.rgbaq(red)), (0x05, rs.xyz2(OFS, OFS, 0)),
                 (0x05, rs.xyz2(OFS + 64 * 16, OFS + 32 * 16, 0))])
    s += gif.ad([(0x06, rs.tex0(0, 2, 0x00, 6, 5))])
    s += gif.ad(sprite(64 * 16, 32 * 16, 128 * 16, 64 * 16, 0, 0, 64 * 16, 32 * 16))
    o = run(s)
    img = frame_image(o)
    assert np.all(img[:32, :64] == red)
    assert np.all(img[32:, 64:] == red)
    assert o.stat(2) == 2  # two render passes


def test_tex2_reloads_clut():
    pal_a = np.arange(256, dtype=np.uint64) * 0x00010101
    pal_b = np.arange(256, dtype=np.uint64) * 0x01000001
    idx = np.tile(np.arange(64, dtype=np.uint64), (32, 1))
    s = gif.ad(base_regs()) + csm1_clut_upload(CBP, pal_a) + csm1_clut_upload(CBP2, pal_b)
    s += upload(TBP, 2, 0x13, 64, 32, idx)
    s += gif.ad([(0x06, rs.tex0(TBP, 2, 0x13, 6, 5, cbp=CBP, cld=1))])
    s += gif.ad(sprite(0, 0, 64 * 16, 32 * 16, 0, 0, 64 * 16, 32 * 16))
    # TEX2 replaces only the palette fields of TEX0 (the texture base and size stay).
    s += gif.ad([(0x16, rs.tex0(0, 0, 0x13, 0, 0, cbp=CBP2, cld=1))])
    s += gif.ad(sprite(64 * 16, 0, 128 * 16, 32 * 16, 0, 0, 64 * 16, 32 * 16))
    img = frame_image(run(s))
    assert np.array_equal(img[:32, :64], pal_a[idx])
    assert np.array_equal(img[:32, 64:], pal_b[idx])
