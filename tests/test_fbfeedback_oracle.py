"""Oracle checks of on-tile frame-buffer feedback textures (no GPU): gs/gs_interface.cpp:2141-2307
(when a texture is the frame or depth buffer), 2388-2470 (pixel mode), 909-933 and
ubershader.comp:955-1040 (the texel is the pixel's own value).

A sprite textured from its own frame buffer whose UVs land on each pixel's own texel (offset in
[0, 1) texel with NEAREST, exactly half a texel with LINEAR) reads the pixel's current colour
without ending the render pass; in a super-sampled pass every sample reads its own colour. Other
UVs read the texture after the pass is rendered (a second pass). A texture that aliases the
frame buffer with another layout is read as it was when the pass started. Depth feedback reads
the depth buffer and renders its pass single-sampled. Palette formats in the top bits of the
word index the CLUT.
"""
import numpy as np

import gif
import render_streams as rs
from oracle_lib import Oracle, swizzle as oswz

OFS = 2048 * 16
FW, FH = 128, 64   # frame: 128 x 64 at block 0, FBW 2
VRAM = 4 << 20


def frame_words(v, psm=0x00, fbp=0, w=FW, h=FH):
    w32 = v[:VRAM].view(np.uint32)
    return np.array([[int(w32[oswz(x, y, fbp * 32, 2, psm)]) for x in range(w)] for y in range(h)], np.uint64)


def background(rng):
    px = rng.integers(0, 2**32, size=FW * FH, dtype=np.uint64).astype(np.uint32)
    return px.reshape(FH, FW), gif.upload(0, 2, 0x00, 0, 0, FW, FH, px.tobytes())


def feedback_sprite(d=8, rgba=0x40404040, tfx=0, mmag=0, psm=0x00, tbp=0, x1=FW, y1=FH, extra=()):
    regs = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 2, 0x00)),
            (0x40, rs.scissor(0, FW - 1, 0, FH - 1)), (0x47, rs.test_reg(zte=0)), (0x14, rs.tex1(mmag, 0)),
            (0x08, 0), (0x06, rs.tex0(tbp, 2, psm, 7, 6, tcc=1, tfx=tfx))]
    regs += list(extra)
    regs += [(0x00, rs.prim_reg(6, iip=0, tme=1, fst=1)), (0x01, rs.rgbaq(rgba)),
             (0x03, rs.uv(d, d)), (0x05, rs.xyz2(OFS, OFS, 0)),
             (0x03, rs.uv(x1 * 16 + d, y1 * 16 + d)), (0x05, rs.xyz2(OFS + x1 * 16, OFS + y1 * 16, 0))]
    return gif.ad(regs)


def halve(words):
    out = np.zeros_like(words)
    for ch in range(4):
        out |= (((words >> (8 * ch)) & 0xFF) // 2) << (8 * ch)
    return out


def run(parts, sample_log2=0):
    o = Oracle(sample_log2)
    o.gif_transfer(gif.arr(b"".join(parts)))
    v = o.vram()
    return o, v


def test_pixel_feedback_reads_own_pixel_without_a_new_pass():
    """MODULATE by 0x40 halves every channel of every pixel: each pixel read itself, in one pass."""
    rng = np.random.default_rng(1)
    img, up = background(rng)
    for d in (0, 8, 15):
        o, v = run([up, feedback_sprite(d=d)])
        assert np.array_equal(frame_words(v), halve(img.astype(np.uint64))), d
        assert o.stat(2) == 1, "the feedback sprite needs no pass of its own"


def test_linear_feedback_needs_texel_centres():
    rng = np.random.default_rng(2)
    img, up = background(rng)
    o, v = run([up, feedback_sprite(d=8, mmag=1)])
    assert np.array_equal(frame_words(v), halve(img.astype(np.uint64)))
    assert o.stat(2) == 1


def test_other_uvs_read_the_texture_after_the_pass():
    """UVs 1.5 texels ahead read the lower right neighbour as a texture of the rendered pass."""
    rng = np.random.default_rng(3)
    img, up = background(rng)
    tri = gif.ad([(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 2, 0x00)),
                  (0x40, rs.scissor(0, FW - 1, 0, FH - 1)), (0x47, rs.test_reg(zte=0)),
                  (0x00, rs.prim_reg(6, iip=0)), (0x01, rs.rgbaq(0x80FF00FF)), (0x05, rs.xyz2(OFS, OFS, 0)),
                  (0x05, rs.xyz2(OFS + 16 * 16, OFS + 16 * 16, 0))])
    o, v = run([up, tri, feedback_sprite(d=24, x1=FW - 2)])
    ref = img.astype(np.uint64).copy()
    ref[:16, :16] = 0x80FF00FF
    got = frame_words(v)
    assert np.array_equal(got[:FH - 1, :FW - 2], halve(ref[1:, 1:FW - 1]))
    assert o.stat(2) == 2


def test_feedback_after_a_draw_in_the_same_pass_sees_it():
    """A flat square drawn first, then the feedback sprite in the same pass: the square's pixels
    are halved from the square's colour."""
    rng = np.random.default_rng(4)
    img, up = background(rng)
    tri = gif.ad([(0x00, rs.prim_reg(6, iip=0)), (0x01, rs.rgbaq(0x80FF00FF)), (0x05, rs.xyz2(OFS, OFS, 0)),
                  (0x05, rs.xyz2(OFS + 16 * 16, OFS + 16 * 16, 0))])
    setup = gif.ad([(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 2, 0x00)),
                    (0x40, rs.scissor(0, FW - 1, 0, FH - 1)), (0x47, rs.test_reg(zte=0))])
    o, v = run([up, setup, tri, feedback_sprite()])
    ref = img.astype(np.uint64).copy()
    ref[:16, :16] = 0x80FF00FF
    assert np.array_equal(frame_words(v), halve(ref))
    assert o.stat(2) == 1


def test_palette_feedback_indexes_the_clut_with_the_top_byte():
    """PSMT8H over a 32-bit frame: the alpha byte indexes a 32-bit CSM1 CLUT (DECAL, TCC)."""
    rng = np.random.default_rng(5)
    img, up = background(rng)
    clut = rng.integers(0, 2**32, size=256, dtype=np.uint64).astype(np.uint32)
    # CSM1 layout of a 256-entry 32-bit palette in a 16 x 16 CT32 block at CBP.
    grid = np.zeros((16, 16), np.uint32)
    for i in range(256):
        x, y = rs.csm1_position(i)
        grid[y, x] = clut[i]
    cbp = 2048
    clut_up = gif.upload(cbp, 1, 0x00, 0, 0, 16, 16, grid.tobytes())
    t0 = rs.tex0(0, 2, 0x1B, 7, 6, tcc=1, tfx=1, cbp=cbp, cpsm=0x00, csm=0, csa=0, cld=1)
    sprite = feedback_sprite(psm=0x1B, tfx=1, extra=[(0x06, t0)])
    o, v = run([up, clut_up, sprite])
    expect = clut[(img >> 24).astype(np.int64)].astype(np.uint64)
    assert np.array_equal(frame_words(v), expect)


def test_depth_feedback_reads_depth_single_sampled():
    """A PSMZ32 texture at the Z buffer: DECAL writes each pixel's depth as its colour; at 4x the
    pass renders single-sampled (no sample plane changes)."""
    zbp_pages = 64
    for sample_log2 in (0, 2):
        o = Oracle(sample_log2)
        setup = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 2, 0x00)),
                 (0x4E, rs.zbuf(zbp_pages, 0x0)), (0x40, rs.scissor(0, FW - 1, 0, FH - 1)),
                 (0x47, rs.test_reg(zte=1, ztst=1))]
        draw = [(0x00, rs.prim_reg(3, iip=0)), (0x01, rs.rgbaq(0x80102030)),
                (0x05, rs.xyz2(OFS, OFS, 0x123456)), (0x05, rs.xyz2(OFS + FW * 16, OFS, 0x123456)),
                (0x05, rs.xyz2(OFS, OFS + FH * 16, 0x654321))]
        o.gif_transfer(gif.arr(gif.ad(setup + draw)))
        o.flush()
        before = o.planes().copy()
        fb = [(0x14, rs.tex1(0, 0)), (0x08, 0), (0x06, rs.tex0(zbp_pages * 32, 2, 0x30, 7, 6, tcc=1, tfx=1)),
              (0x00, rs.prim_reg(6, iip=0, tme=1, fst=1)), (0x01, rs.rgbaq(0x80808080)),
              (0x03, rs.uv(8, 8)), (0x05, rs.xyz2(OFS, OFS, 0)),
              (0x03, rs.uv(FW * 16 + 8, FH * 16 + 8)), (0x05, rs.xyz2(OFS + FW * 16, OFS + FH * 16, 0))]
        o.gif_transfer(gif.arr(gif.ad(setup + fb)))
        v = o.vram()
        depth = frame_words(before, psm=0x30, fbp=zbp_pages)
        assert np.array_equal(frame_words(v), depth)
        if sample_log2:
            after = o.planes()
            assert np.array_equal(after[2 * VRAM:], before[2 * VRAM:]), "single-sampled pass"


def test_other_layout_reads_the_frame_as_the_pass_starts():
    """A PSMCT16 texture over the 32-bit frame buffer (another layout): it is read as it was
    before the pass, so a square drawn earlier in the pass does not show through."""
    rng = np.random.default_rng(6)
    img, up = background(rng)
    setup = gif.ad([(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 2, 0x00)),
                    (0x40, rs.scissor(0, FW - 1, 0, FH - 1)), (0x47, rs.test_reg(zte=0))])
    square = gif.ad([(0x00, rs.prim_reg(6, iip=0)), (0x01, rs.rgbaq(0x80FF00FF)), (0x05, rs.xyz2(OFS, OFS, 0)),
                     (0x05, rs.xyz2(OFS + 16 * 16, OFS + 16 * 16, 0))])
    # The 16-bit view of the same memory, drawn over the right half only.
    sprite = gif.ad([(0x14, rs.tex1(0, 0)), (0x08, 0), (0x06, rs.tex0(0, 2, 0x02, 7, 6, tcc=1, tfx=1)),
                     (0x00, rs.prim_reg(6, iip=0, tme=1, fst=1)), (0x01, rs.rgbaq(0x80808080)),
                     (0x03, rs.uv(0, 0)), (0x05, rs.xyz2(OFS + 64 * 16, OFS, 0)),
                     (0x03, rs.uv(64 * 16, 16 * 16)), (0x05, rs.xyz2(OFS + 128 * 16, OFS + 16 * 16, 0))])
    o, v = run([up, setup, square, sprite])
    # Same stream with the square in a pass of its own before (a transfer renders the pending
    # pass first): the sprite now sees the square.
    far = gif.upload(8192, 1, 0x00, 0, 0, 2, 2, b"\0" * 16)
    o2, v2 = run([up, setup, square, far, sprite])
    got, fresh = frame_words(v), frame_words(v2)
    assert o.stat(2) == 1 and o2.stat(2) == 2
    assert np.array_equal(got[:16, :16], fresh[:16, :16])
    # The sprite's texels at u < 16 (the square's 16-bit halves) differ between the two reads.
    assert not np.array_equal(got[:16, 64:80], fresh[:16, 64:80])


def test_super_sampled_feedback_reads_each_sample():
    """4x: an edge-antialiased triangle and the feedback sprite in one pass. Where the triangle's
    edge made a pixel super-sampled, every sample is halved from its own value (a resolved read
    would give all samples one value); elsewhere the pixel is halved."""
    rng = np.random.default_rng(7)
    img, up = background(rng)
    setup = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 2, 0x00)), (0x4E, rs.zbuf(64, 0x1)),
             (0x40, rs.scissor(0, FW - 1, 0, FH - 1)), (0x47, rs.test_reg(zte=0))]
    # Depth varies over the triangle, so it is not demoted to single-sampled shading.
    tri = [(0x00, rs.prim_reg(3, iip=1)), (0x01, rs.rgbaq(0x80F08010)), (0x05, rs.xyz2(OFS + 3 * 16 + 5, OFS, 0x1000)),
           (0x01, rs.rgbaq(0x8010F080)), (0x05, rs.xyz2(OFS + 100 * 16 + 3, OFS + 20 * 16 + 7, 0x2000)),
           (0x01, rs.rgbaq(0x808010F0)), (0x05, rs.xyz2(OFS + 9 * 16, OFS + 60 * 16 + 9, 0x3000))]
    a = Oracle(2)
    a.gif_transfer(gif.arr(b"".join([up, gif.ad(setup + tri)])))
    a.flush()
    before = a.planes().copy()
    b = Oracle(2)
    b.gif_transfer(gif.arr(b"".join([up, gif.ad(setup + tri), feedback_sprite()])))
    b.flush()
    after = b.planes()
    assert b.stat(2) == 1
    ref = frame_words(after[VRAM:2 * VRAM])
    ss = ref == 0xFFFFFFFF
    assert ss.sum() > 50
    # Super-sampled pixels: per sample.
    distinct = 0
    for s in range(4):
        plane_b = frame_words(before[(2 + s) * VRAM:(3 + s) * VRAM])
        plane_a = frame_words(after[(2 + s) * VRAM:(3 + s) * VRAM])
        assert np.array_equal(plane_a[ss], halve(plane_b[ss]))
        distinct += int((plane_b[ss] != frame_words(before[2 * VRAM:3 * VRAM])[ss]).sum())
    assert distinct > 0, "the samples of edge pixels differ"
    # Single-sampled pixels: the pixel.
    assert np.array_equal(frame_words(after)[~ss], halve(frame_words(before)[~ss]))
