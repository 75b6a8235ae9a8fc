"""CPU oracle rendering pinned by closed forms (no GPU).

Each expected image is computed here independently with exact integer arithmetic:
coverage from the top-left-rule edge test on 12.4 fixed-point vertices, flat colours and
depths, the GS blend formula ((A - B) * C >> 7) + D, dither offsets and FBMSK merges.
"""
import hashlib
import struct

import numpy as np
import pytest

import gif
import render_streams as rs
from oracle_lib import Oracle, oracle_lib, swizzle as oswz

OFS = 2048 * 16
W, H = 64, 48


def setup_regs(fb_psm=0x00, zte=0, ztst=1, fbmsk=0, extra=()):
    regs = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 1, fb_psm, fbmsk)),
            (0x4E, rs.zbuf(16, 0x1)), (0x40, rs.scissor(0, W - 1, 0, H - 1)),
            (0x47, rs.test_reg(zte=zte, ztst=ztst)), (0x46, 1)]
    regs.extend(extra)
    return regs


def fb_image(o, psm=0x00):
    v = o.vram()
    w32 = v.view(np.uint32)
    img = np.zeros((H, W), dtype=np.uint64)
    for y in range(H):
        for x in range(W):
            img[y, x] = w32[oswz(x, y, 0, 1, psm)]
    return img


def z_image(o):
    v = o.vram()
    w32 = v.view(np.uint32)
    img = np.zeros((H, W), dtype=np.uint64)
    for y in range(H):
        for x in range(W):
            img[y, x] = w32[oswz(x, y, 16 * 32, 1, 0x31)] & 0xFFFFFF
    return img


def exact_coverage(verts):
    """Pixels whose top-left corner passes the edge test with the top-left tie-break rule."""
    (ax, ay), (bx, by), (cx, cy) = verts
    # Orientation with positive area in y-down window coordinates.
    area = (by - ay) * (cx - ax) - (bx - ax) * (cy - ay)
    if area == 0:
        return np.zeros((H, W), dtype=bool)
    if area < 0:
        (bx, by), (cx, cy) = (cx, cy), (bx, by)
    edges = [((ax, ay), (bx, by)), ((bx, by), (cx, cy)), ((cx, cy), (ax, ay))]
    cov = np.ones((H, W), dtype=bool)
    ys, xs = np.mgrid[0:H, 0:W]
    px, py = xs * 16, ys * 16
    for (x0, y0), (x1, y1) in edges:
        ex, ey = x1 - x0, y1 - y0
        # Inside is where the edge function ey * (x - x0) - ex * (y - y0) is positive; points
        # exactly on an edge belong to it only when the edge is a top or left edge.
        e = ey * (px - x0) - ex * (py - y0)
        top_left = ey > 0 or (ey == 0 and ex < 0)
        cov &= (e > 0) | ((e == 0) & top_left)
    return cov


def draw_flat_triangle(verts, rgba, z=0, fb_psm=0x00):
    o = Oracle()
    writes = setup_regs(fb_psm) + [(0x00, rs.prim_reg(3, iip=0))]
    for (x, y) in verts:
        writes.append((0x01, rs.rgbaq(rgba)))
        writes.append((0x05, rs.xyz2(OFS + x, OFS + y, z)))
    o.gif_transfer(gif.arr(gif.ad(writes)))
    return o


@pytest.mark.parametrize("seed", range(40))
def test_triangle_coverage_matches_exact_edge_test(seed):
    rng = np.random.default_rng(seed)
    verts = [(int(rng.integers(-64, W * 16 + 64)), int(rng.integers(-64, H * 16 + 64))) for _ in range(3)]
    if seed % 5 == 0:
        # axis-aligned and shared-coordinate edges exercise the tie-break rule
        verts[1] = (verts[0][0], verts[1][1])
        verts[2] = (verts[2][0], verts[0][1])
    o = draw_flat_triangle(verts, 0x80FFFFFF)
    img = fb_image(o)
    cov = exact_coverage(verts)
    # The GS vertex order is reversed on assembly (newest first); coverage is order independent.
    assert np.array_equal(img != 0, cov), (verts, int((img != 0).sum()), int(cov.sum()))


def test_flat_colour_and_depth_are_exact():
    verts = [(10 * 16 + 3, 5 * 16), (50 * 16, 9 * 16 + 7), (20 * 16, 40 * 16)]
    o = Oracle()
    writes = setup_regs(zte=1, ztst=1) + [(0x00, rs.prim_reg(3, iip=0))]
    for (x, y) in verts:
        writes.append((0x01, rs.rgbaq(0x11223344)))
        writes.append((0x05, rs.xyz2(OFS + x, OFS + y, 0x00ABCDEF)))
    o.gif_transfer(gif.arr(gif.ad(writes)))
    cov = exact_coverage(verts)
    img, zimg = fb_image(o), z_image(o)
    assert np.all(img[cov] == 0x11223344) and np.all(img[~cov] == 0)
    assert np.all(zimg[cov] == 0xABCDEF) and np.all(zimg[~cov] == 0)


def test_depth_ramp_within_float_rounding_of_the_plane():
    verts = [(4 * 16, 4 * 16), (60 * 16, 8 * 16), (12 * 16, 44 * 16)]
    zs = [1000, 9000000, 16000000]
    o = Oracle()
    writes = setup_regs(zte=1, ztst=1) + [(0x00, rs.prim_reg(3, iip=0))]
    for (x, y), z in zip(verts, zs):
        writes.append((0x01, rs.rgbaq(0xFFFFFFFF)))
        writes.append((0x05, rs.xyz2(OFS + x, OFS + y, z)))
    o.gif_transfer(gif.arr(gif.ad(writes)))
    zimg = z_image(o)
    cov = exact_coverage(verts)
    (ax, ay), (bx, by), (cx, cy) = verts
    m = np.array([[bx - ax, cx - ax], [by - ay, cy - ay]], dtype=np.float64)
    inv = np.linalg.inv(m)
    ys, xs = np.nonzero(cov)
    for x, y in zip(xs, ys):
        i, j = inv @ np.array([x * 16 - ax, y * 16 - ay])
        exact = zs[0] + i * (zs[1] - zs[0]) + j * (zs[2] - zs[0])
        assert abs(int(zimg[y, x]) - exact) <= 4, (x, y, int(zimg[y, x]), exact)


def test_gouraud_colour_within_one_of_linear_interpolation():
    verts = [(2 * 16, 2 * 16), (62 * 16, 6 * 16), (30 * 16, 46 * 16)]
    cols = [0x000000FF, 0x00FF0000, 0x0000FF00]
    o = Oracle()
    writes = setup_regs() + [(0x00, rs.prim_reg(3, iip=1))]
    for (x, y), c in zip(verts, cols):
        writes.append((0x01, rs.rgbaq(c)))
        writes.append((0x05, rs.xyz2(OFS + x, OFS + y, 0)))
    o.gif_transfer(gif.arr(gif.ad(writes)))
    img = fb_image(o)
    cov = exact_coverage(verts)
    (ax, ay), (bx, by), (cx, cy) = verts
    inv = np.linalg.inv(np.array([[bx - ax, cx - ax], [by - ay, cy - ay]], dtype=np.float64))
    ys, xs = np.nonzero(cov)
    # The newest vertex (last kicked) is the provoking one after assembly; the plane is the same.
    for x, y in zip(xs, ys):
        i, j = inv @ np.array([x * 16 - ax, y * 16 - ay])
        for ch in range(3):
            c = [(cc >> (8 * ch)) & 0xFF for cc in cols]
            exact = c[0] + i * (c[1] - c[0]) + j * (c[2] - c[0])
            got = (int(img[y, x]) >> (8 * ch)) & 0xFF
            assert abs(got - exact) <= 1.0, (x, y, ch, got, exact)


def sprite(o_writes, x0, y0, x1, y1, rgba, z=0):
    o_writes.append((0x01, rs.rgbaq(rgba)))
    o_writes.append((0x05, rs.xyz2(OFS + x0, OFS + y0, z)))
    o_writes.append((0x01, rs.rgbaq(rgba)))
    o_writes.append((0x05, rs.xyz2(OFS + x1, OFS + y1, z)))


@pytest.mark.parametrize("a,b,c,d", [(0, 1, 0, 1), (1, 0, 2, 0), (0, 2, 0, 2), (0, 1, 2, 1), (1, 2, 1, 0)])
def test_alpha_blend_integer_formula(a, b, c, d):
    """Cv = ((A - B) * C >> 7) + D on integers, wrapped to 8 bits when COLCLAMP=0."""
    cd, cs, afix = 0x40C08020, 0x60305090, 77
    o = Oracle()
    writes = setup_regs() + [(0x46, 0), (0x42, rs.alpha_reg(a, b, c, d, afix)), (0x00, rs.prim_reg(6, abe=0))]
    sprite(writes, 0, 0, W * 16, H * 16, cd)
    writes.append((0x00, rs.prim_reg(6, abe=1)))
    sprite(writes, 0, 0, W * 16, H * 16, cs)
    o.gif_transfer(gif.arr(gif.ad(writes)))
    img = fb_image(o)
    sel = {0: cs, 1: cd, 2: 0}
    cval = {0: cs >> 24, 1: cd >> 24, 2: afix}[c]
    expect = 0
    for ch in range(3):
        A = (sel[a] >> (8 * ch)) & 0xFF if a != 2 else 0
        B = (sel[b] >> (8 * ch)) & 0xFF if b != 2 else 0
        D = (sel[d] >> (8 * ch)) & 0xFF if d != 2 else 0
        v = ((A - B) * cval >> 7) + D
        expect |= (v & 0xFF) << (8 * ch)
    expect |= cs & 0xFF000000
    assert np.all(img == expect), (hex(int(img[0, 0])), hex(expect))


def test_dither_adds_signed_matrix_entry():
    dimx = 0
    vals = [-4, -3, -2, -1, 0, 1, 2, 3, 3, 2, 1, 0, -1, -2, -3, -4]
    for i, v in enumerate(vals):
        dimx |= (v & 7) << (4 * i)
    o = Oracle()
    writes = setup_regs() + [(0x45, 1), (0x44, dimx), (0x00, rs.prim_reg(6))]
    sprite(writes, 0, 0, W * 16, H * 16, 0x80808080)
    o.gif_transfer(gif.arr(gif.ad(writes)))
    img = fb_image(o)
    for y in range(4):
        for x in range(4):
            v = 0x80 + vals[y * 4 + x]
            assert int(img[y, x]) & 0xFFFFFF == v | v << 8 | v << 16


def test_fbmsk_merges_bitwise():
    o = Oracle()
    writes = setup_regs() + [(0x00, rs.prim_reg(6))]
    sprite(writes, 0, 0, W * 16, H * 16, 0x11223344)
    writes.append((0x4C, rs.frame(0, 1, 0x00, 0x0F0F00FF)))
    sprite(writes, 0, 0, W * 16, H * 16, 0xAABBCCDD)
    o.gif_transfer(gif.arr(gif.ad(writes)))
    img = fb_image(o)
    assert np.all(img == ((0xAABBCCDD & ~0x0F0F00FF) | (0x11223344 & 0x0F0F00FF)) & 0xFFFFFFFF)


def test_sprite_covers_top_left_inclusive_rectangle():
    o = Oracle()
    writes = setup_regs() + [(0x00, rs.prim_reg(6))]
    x0, y0, x1, y1 = 5 * 16 + 8, 3 * 16, 20 * 16 + 1, 17 * 16 + 15
    sprite(writes, x0, y0, x1, y1, 0xFFFFFFFF)
    o.gif_transfer(gif.arr(gif.ad(writes)))
    img = fb_image(o) != 0
    ys, xs = np.mgrid[0:H, 0:W]
    expect = (xs * 16 >= x0) & (xs * 16 < x1) & (ys * 16 >= y0) & (ys * 16 < y1)
    assert np.array_equal(img, expect)


@pytest.mark.parametrize("second,fused", [
    # Right angle of the second triangle kicked first: both triangles are candidates with the
    # provoking corner last after assembly, so they fuse (gs/gs_util.cpp:237-377).
    ("fuse", 1),
    # Right angle kicked second: the candidate test checks the angle at A for the "B provoking"
    # order (gs_util.cpp:287-289), so this order does not fuse; coverage is identical anyway.
    ("no_fuse", 2),
])
def test_two_triangles_forming_a_rectangle(second, fused):
    o = Oracle()
    writes = setup_regs() + [(0x00, rs.prim_reg(3, iip=0))]
    x0, y0, x1, y1 = 3 * 16, 2 * 16, 40 * 16, 30 * 16
    tri2 = [(x1, y1), (x1, y0), (x0, y1)] if second == "fuse" else [(x1, y0), (x1, y1), (x0, y1)]
    for (x, y) in [(x0, y0), (x1, y0), (x0, y1)] + tri2:
        writes.append((0x01, rs.rgbaq(0xFFFFFFFF)))
        writes.append((0x05, rs.xyz2(OFS + x, OFS + y, 0)))
    o.gif_transfer(gif.arr(gif.ad(writes)))
    img = fb_image(o) != 0
    assert o.stat(1) == fused  # primitives that reached the raster stage
    ys, xs = np.mgrid[0:H, 0:W]
    expect = (xs * 16 >= x0) & (xs * 16 < x1) & (ys * 16 >= y0) & (ys * 16 < y1)
    assert np.array_equal(img, expect)


def test_rcp_seed_table_checksum():
    L = oracle_lib()
    import ctypes
    L.orc_rcp_table.restype = ctypes.c_void_p
    tab = bytes((ctypes.c_uint8 * 256).from_address(L.orc_rcp_table()))
    assert hashlib.sha256(tab).hexdigest() == "9effff71937dcdb904c074dda412250a136d9024f34d921f93efd258324f9c12"


def test_rcp_fixed_error_bound():
    """The fixed-point Newton reciprocal is within ~1 ulp of 1/x over the whole 32-bit range."""
    import ctypes
    L = oracle_lib()
    L.orc_rcp_fixed.restype = ctypes.c_float
    L.orc_rcp_fixed.argtypes = [ctypes.c_uint32]
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.arange(1, 5000), rng.integers(1, 2**32, 20000)]).astype(np.uint64)
    for x in xs:
        r = L.orc_rcp_fixed(int(x))
        exact = 1.0 / float(x)
        assert abs(r - exact) <= 2.0 * np.spacing(np.float32(exact)), (int(x), r, exact)


def test_scissor_change_ends_parallelogram_fusion():
    """Two triangles forming a parallelogram with a SCISSOR write between them are not fused
    (gs/gs_interface.cpp:2546, 2584): the second is clipped by the new scissor, so the right
    half it covers stays untouched, while without the change the whole quad is drawn."""
    import gif
    import render_streams as rs
    from oracle_lib import Oracle, swizzle
    ofs = 2048 * 16

    def run(second_scissor):
        w = [(0x1A, 1), (0x18, ofs | ofs << 32), (0x4C, rs.frame(0, 1, 0)), (0x40, rs.scissor(0, 63, 0, 31)),
             (0x47, rs.test_reg(zte=0)), (0x00, rs.prim_reg(4, iip=0))]  # triangle strip
        quad = [(0, 0), (64, 0), (0, 32), (64, 32)]
        for k, (x, y) in enumerate(quad):
            if k == 3:
                w.append((0x40, second_scissor))
            w += [(0x01, rs.rgbaq(0xFF0000FF)), (0x05, rs.xyz2(ofs + x * 16, ofs + y * 16, 0))]
        o = Oracle(0)
        o.gif_transfer(gif.arr(gif.ad(w)))
        o.flush()
        v = o.vram()
        return np.array([[int.from_bytes(v[swizzle(x, y, 0, 1, 0) * 4:][:4].tobytes(), "little")
                          for x in range(64)] for y in range(32)])

    full = run(rs.scissor(0, 63, 0, 31))
    half = run(rs.scissor(0, 31, 0, 31))
    assert (full == 0xFF0000FF).all()
    # The first triangle (upper left of the diagonal) keeps the full scissor; the second is held
    # to the left half, so the lower right corner stays clear.
    assert (half[:, :32] == 0xFF0000FF).all()
    assert half[0, 62] == 0xFF0000FF and half[31, 63] == 0 and (half[:, 32:] == 0).sum() > 0
