"""Oracle checks of render passes whose depth buffer is their colour buffer (FBP == ZBP), no GPU:
gs/gs_renderer.cpp:2525-2690 (which primitives shade together), 2418-2475 (runs that read a
cached copy of depth) and 2184-2232 (the copy).

The colour and depth formats swizzle differently, so a pixel's colour lands on another pixel's
depth: with PSMCT32 / PSMZ32 in one page, 8x8 block (0, 0)'s colour words are block (4, 2)'s
depth words and the reverse (the layouts are found with the swizzle below, not assumed). The
expected VRAM of each case follows from the shading order alone:

* a primitive that writes depth shades alone and sees every earlier primitive's colour as depth;
  inside it, every read sees VRAM as it was before it, and every colour is stored before any
  depth (where a pixel's colour and another's depth share a word, depth is left);
* primitives that only test depth read depth from a copy taken as their run starts; a run ends
  before a depth-testing sprite whose box overlaps the run's earlier sprites (other primitives
  never end a run), so that sprite sees their colour as depth and the others do not.
"""
import numpy as np
import pytest

import gif
import render_streams as rs
from oracle_lib import Oracle, swizzle as oswz

OFS = 2048 * 16
FBW = 2
W, H = 128, 64

GEQUAL, ALWAYS = 2, 1


def block_words(bx, by, psm):
    """Word (32-bit) or halfword (16-bit) indices of 8x8 block (bx, by) (16x8 for 16-bit formats)."""
    bw = 16 if psm in (0x02, 0x0A, 0x32, 0x3A) else 8
    return sorted({oswz(bx * bw + x, by * 8 + y, 0, FBW, psm) for x in range(bw) for y in range(8)})


def partner(bx, by, cpsm, zpsm):
    """The block whose depth words are block (bx, by)'s colour words."""
    target = set(block_words(bx, by, cpsm))
    bw = 16 if cpsm in (0x02, 0x0A) else 8
    for qy in range(H // 8):
        for qx in range(W // bw):
            if set(block_words(qx, qy, zpsm)) == target:
                return qx, qy
    raise AssertionError("no aliasing block")


def sprite(x0, y0, x1, y1, rgba, z, ztst=GEQUAL, zmsk=0, zte=1, zpsm=0x0):
    """A flat sprite over pixels [x0, x1) x [y0, y1) with its own TEST / ZBUF.ZMSK."""
    return [(0x47, rs.test_reg(zte=zte, ztst=ztst)), (0x4E, rs.zbuf(0, zpsm, zmsk)),
            (0x00, rs.prim_reg(6, iip=0)), (0x01, rs.rgbaq(rgba)),
            (0x05, rs.xyz2(OFS + x0 * 16, OFS + y0 * 16, z)), (0x05, rs.xyz2(OFS + x1 * 16, OFS + y1 * 16, z))]


def setup(fpsm=0x00, zpsm=0x0):
    return [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, FBW, fpsm)), (0x4E, rs.zbuf(0, zpsm)),
            (0x40, rs.scissor(0, W - 1, 0, H - 1))]


def render(regs, sample_log2=0):
    o = Oracle(sample_log2)
    o.gif_transfer(gif.arr(gif.ad(setup() + regs)))
    return o.vram().view(np.uint32)


def expect(**blocks):
    """VRAM words: zero except the listed (words, value) pairs, applied in order."""
    out = np.zeros((4 << 20) // 4, np.uint32)
    for words, value in blocks.values():
        out[words] = value
    return out


A = (0, 0)
Q = partner(*A, 0x00, 0x30)


def px(b):
    return b[0] * 8, b[1] * 8, b[0] * 8 + 8, b[1] * 8 + 8


def test_layouts_swap_blocks_within_a_page():
    assert Q != A
    assert partner(*Q, 0x00, 0x30) == A
    assert set(block_words(*A, 0x30)) == set(block_words(*Q, 0x00))


def test_single_step_sees_earlier_colour_as_depth_and_fails():
    """A writes colour 0xF0000000 and depth 0x1000; B's depth words are A's colour, so B's
    GEQUAL 0x2000 fails everywhere and B writes nothing (shaded together, B would read 0)."""
    v = render(sprite(*px(A), 0xF0000000, 0x1000) + sprite(*px(Q), 0x12345678, 0x2000))
    ref = expect(a=(block_words(*A, 0x00), 0xF0000000), b=(block_words(*A, 0x30), 0x1000))
    assert np.array_equal(v, ref)


def test_single_step_passes_and_depth_lands_on_colour():
    """A's colour 0x10 is below B's 0x2000: B passes, stores its colour over A's depth and its
    depth over A's colour."""
    v = render(sprite(*px(A), 0x00000010, 0x1000) + sprite(*px(Q), 0x12345678, 0x2000))
    ref = expect(a=(block_words(*A, 0x00), 0x2000), b=(block_words(*Q, 0x00), 0x12345678))
    assert np.array_equal(v, ref)


def test_inside_one_primitive_depth_is_stored_after_colour():
    """One depth-writing sprite over blocks (0..4, 0..2): every read sees the cleared VRAM (all
    pass); every word some pixel's depth lands on holds depth, the other colour words colour."""
    v = render(sprite(0, 0, 40, 24, 0x0A0B0C0D, 0x777))
    cwords, zwords = set(), set()
    for by in range(3):
        for bx in range(5):
            cwords |= set(block_words(bx, by, 0x00))
            zwords |= set(block_words(bx, by, 0x30))
    assert cwords & zwords  # blocks (0, 0) and (4, 2) swap
    ref = expect(c=(sorted(cwords), 0x0A0B0C0D), z=(sorted(zwords), 0x777))
    assert np.array_equal(v, ref)


def test_test_only_pass_reads_colour_drawn_before_as_depth():
    """No depth writes at all: A (no depth test) shades first, B (GEQUAL, ZMSK) reads A's colour
    as depth from the copy and fails."""
    v = render(sprite(*px(A), 0xF0000000, 0, ztst=ALWAYS, zmsk=1) + sprite(*px(Q), 0x12345678, 0x2000, zmsk=1))
    ref = expect(a=(block_words(*A, 0x00), 0xF0000000))
    assert np.array_equal(v, ref)


def test_overlapping_sprite_starts_a_run_and_sees_colour():
    """Test-only sprites: the first draws A (depth 0 passes); the second covers blocks
    (0..4, 0..2), overlaps the first, so it reads depth copied after it: its pixels in Q (depth =
    A's colour 0xF0000000) fail, the rest pass."""
    v = render(sprite(*px(A), 0xF0000000, 0x10, zmsk=1) + sprite(0, 0, 40, 24, 0x11223344, 0x2000, zmsk=1))
    words = set()
    for by in range(3):
        for bx in range(5):
            if (bx, by) != Q:
                words |= set(block_words(bx, by, 0x00))
    ref = expect(c=(sorted(words), 0x11223344))
    assert np.array_equal(v, ref)


def test_disjoint_sprite_stays_in_the_run_and_reads_the_old_depth():
    """The second sprite covers Q only, which does not overlap the first sprite's box: same run,
    depth from the copy taken before the first sprite (0), so it passes."""
    v = render(sprite(*px(A), 0xF0000000, 0x10, zmsk=1) + sprite(*px(Q), 0x11223344, 0x2000, zmsk=1))
    ref = expect(a=(block_words(*A, 0x00), 0xF0000000), q=(block_words(*Q, 0x00), 0x11223344))
    assert np.array_equal(v, ref)


def test_triangles_never_end_a_run():
    """The same overlap drawn as two triangles does not end the run (only sprites are checked):
    every pixel reads the old depth and passes."""
    tri = [(0x47, rs.test_reg(zte=1, ztst=GEQUAL)), (0x4E, rs.zbuf(0, 0x0, 1)),
           (0x00, rs.prim_reg(3, iip=0)), (0x01, rs.rgbaq(0x11223344))]
    for (x, y) in ((0, 0), (40, 0), (0, 24), (40, 0), (40, 24), (0, 24)):
        tri.append((0x05, rs.xyz2(OFS + x * 16, OFS + y * 16, 0x2000)))
    v = render(sprite(*px(A), 0xF0000000, 0x10, zmsk=1) + tri)
    words = set()
    for by in range(3):
        for bx in range(5):
            words |= set(block_words(bx, by, 0x00))
    ref = expect(c=(sorted(words), 0x11223344))
    assert np.array_equal(v, ref)


def test_primitives_after_the_last_depth_use_shade_without_depth():
    """After the last depth writer, a primitive without depth test draws normally."""
    v = render(sprite(*px(A), 0x00000010, 0x1000) + sprite(*px(Q), 0x55667788, 0, ztst=ALWAYS, zmsk=1))
    ref = expect(b=(block_words(*A, 0x30), 0x55667788), a=(block_words(*A, 0x00), 0x10))
    assert np.array_equal(v, ref)


@pytest.mark.parametrize("sample_log2", [1, 2, 4])
def test_test_only_runs_super_sampled(sample_log2):
    """Without depth writes the pass keeps the context's rate; plane 0 reads as at 1x."""
    regs = sprite(*px(A), 0xF0000000, 0x10, zmsk=1) + sprite(0, 0, 40, 24, 0x11223344, 0x2000, zmsk=1)
    assert np.array_equal(render(regs, sample_log2), render(regs))


@pytest.mark.parametrize("sample_log2", [2, 4])
def test_depth_writes_render_single_sampled(sample_log2):
    regs = sprite(*px(A), 0xF0000000, 0x1000) + sprite(*px(Q), 0x12345678, 0x2000)
    assert np.array_equal(render(regs, sample_log2), render(regs))


def test_sixteen_bit_layouts():
    """PSMCT16 / PSMZ16: B (GEQUAL) reads A's 16-bit colour 0x8000 (alpha set) as depth and fails."""
    a = (0, 0)
    q = partner(*a, 0x02, 0x32)
    regs = sprite(a[0] * 16, a[1] * 8, a[0] * 16 + 16, a[1] * 8 + 8, 0x80000000, 0, ztst=ALWAYS, zmsk=1, zpsm=0x2)
    b = sprite(q[0] * 16, q[1] * 8, q[0] * 16 + 16, q[1] * 8 + 8, 0x000000F8, 0x4000, zmsk=1, zpsm=0x2)
    o = Oracle(0)
    o.gif_transfer(gif.arr(gif.ad(setup(0x02, 0x2) + regs + b)))
    v = o.vram().view(np.uint16)
    ref = np.zeros_like(v)
    ref[block_words(*a, 0x02)] = 0x8000
    assert np.array_equal(v, ref)
    # With B's depth above 0x8000 it passes and stores red (0x1F).
    b2 = sprite(q[0] * 16, q[1] * 8, q[0] * 16 + 16, q[1] * 8 + 8, 0x000000F8, 0x9000, zmsk=1, zpsm=0x2)
    o = Oracle(0)
    o.gif_transfer(gif.arr(gif.ad(setup(0x02, 0x2) + regs + b2)))
    v = o.vram().view(np.uint16)
    ref[block_words(*q, 0x02)] = 0x001F
    assert np.array_equal(v, ref)
