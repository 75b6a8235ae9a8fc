"""CPU oracle super-sampling pinned by closed forms (no GPU).

Planes of a super-sampled context: 0 = single-sampled VRAM, 1 = reference (all ones where a
pixel's super-samples are valid), 2 + s = sample s. Expected images are computed here from
first principles: the top-left-rule edge test at each sample position (ordered 4x grid at
half-pixel offsets, 16x at quarter-pixel offsets, 2x on the diagonal), and the average of the
covered sample colours, which is an exact integer for the colours chosen.
"""
import numpy as np
import pytest

import gif
import render_streams as rs
from oracle_lib import Oracle, swizzle as oswz

OFS = 2048 * 16
W, H = 64, 48
PLANE = 4 << 20


def setup_regs(extra=()):
    regs = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 1, 0x00)), (0x4E, rs.zbuf(16, 0x1)),
            (0x40, rs.scissor(0, W - 1, 0, H - 1)), (0x47, rs.test_reg(zte=0)), (0x46, 1)]
    regs.extend(extra)
    return regs


def plane_image(planes, plane, psm=0x00, base=0):
    w32 = planes[plane * PLANE:(plane + 1) * PLANE].view(np.uint32)
    img = np.zeros((H, W), dtype=np.uint64)
    for y in range(H):
        for x in range(W):
            img[y, x] = w32[oswz(x, y, base, 1, psm)]
    return img


# Sample positions inside a pixel, in 1/16 pixel units, indexed by sample number.
SAMPLES = {
    1: [(0, 8)],                                   # 2x: (0,0), (0.5,0.5) -> listed per s below
    2: None,
    4: None,
}


SPARSE_X = [0, 2, 3, 1]  # quarter-pixel x offsets of the sparse rows (the 0x78 pattern)


def sample_positions(n):
    if n == "4s":   # sparse 1 x 4: quarter rows
        return [(SPARSE_X[s] * 4, s * 4) for s in range(4)]
    if n == "16s":  # sparse 2 x 8: eighth rows, two samples half a pixel apart
        return [((4 * ((s >> 3) & 1) + SPARSE_X[s & 3]) * 2, (s & 7) * 2) for s in range(16)]
    if n == 2:
        return [(0, 0), (8, 8)]
    if n == 4:
        return [((s >> 1) * 8, (s & 1) * 8) for s in range(4)]
    if n == 16:
        return [((((s >> 1) & 1) + 2 * ((s >> 3) & 1)) * 4, ((s & 1) + 2 * ((s >> 2) & 1)) * 4) for s in range(16)]
    raise ValueError(n)


def covered(verts, px, py):
    """Top-left rule edge test of point (px, py) (1/16 units) against a triangle."""
    (ax, ay), (bx, by), (cx, cy) = verts
    area = (by - ay) * (cx - ax) - (bx - ax) * (cy - ay)
    if area < 0:
        (bx, by), (cx, cy) = (cx, cy), (bx, by)
    for (x0, y0), (x1, y1) in (((ax, ay), (bx, by)), ((bx, by), (cx, cy)), ((cx, cy), (ax, ay))):
        ex, ey = x1 - x0, y1 - y0
        e = ey * (px - x0) - ex * (py - y0)
        top_left = ey > 0 or (ey == 0 and ex < 0)
        if not (e > 0 or (e == 0 and top_left)):
            return False
    return True


def draw_triangle(n_samples, verts, rgba, zs=(0x1000, 0x2000, 0x3000), iip=0, extra=()):
    log2 = {1: 0, 2: 1, 4: 2, 16: 4, "4s": 2, "16s": 4}[n_samples]
    o = Oracle(log2, sparse=isinstance(n_samples, str))
    writes = setup_regs(extra) + [(0x00, rs.prim_reg(3, iip=iip))]
    for (x, y), z in zip(verts, zs):
        writes.append((0x01, rs.rgbaq(rgba)))
        writes.append((0x05, rs.xyz2(OFS + x, OFS + y, z)))
    o.gif_transfer(gif.arr(gif.ad(writes)))
    o.flush()
    return o


@pytest.mark.parametrize("n_samples,seed", [(4, 0), (4, 1), (4, 2), (16, 3), (16, 4), (2, 5), (2, 6),
                                             ("4s", 7), ("4s", 8), ("16s", 9), ("16s", 10)])
def test_edge_antialiasing_resolve(n_samples, seed):
    """A 3D triangle (depth varies, so no single-sample demotion) in flat grey 0x80: each
    pixel resolves to 0x80 * k / n for k covered samples, exactly. Sparse grids ("4s", "16s")
    place their samples on the rows of a 4 x 4 / 8 x 8 sub-grid (ubershader.comp:333-352)."""
    rng = np.random.default_rng(seed)
    verts = [(int(rng.integers(0, W * 8)) * 2, int(rng.integers(0, H * 8)) * 2) for _ in range(3)]
    o = draw_triangle(n_samples, verts, 0x80808080)
    planes = o.planes()
    img = plane_image(planes, 0)
    ref = plane_image(planes, 1)
    pos = sample_positions(n_samples)
    n = len(pos)
    for y in range(H):
        for x in range(W):
            k = sum(covered(verts, x * 16 + dx, y * 16 + dy) for dx, dy in pos)
            v = 0x80 * k // n
            assert img[y, x] == v * 0x01010101, (x, y, k, hex(int(img[y, x])))
            assert ref[y, x] == (0xFFFFFFFF if 0 < k < n else 0), (x, y, k)
            if 0 < k < n:
                for s, (dx, dy) in enumerate(pos):
                    smp = plane_image(planes, 2 + s)[y, x] if s < 2 else None
                    if smp is not None:
                        assert smp == (0x80808080 if covered(verts, x * 16 + dx, y * 16 + dy) else 0)


def test_opaque_sprite_renders_single_sampled():
    """An opaque sprite snaps to the pixel grid: every tile it touches is single-sampled, the
    reference plane is cleared and no sample plane is written."""
    o = Oracle(2)
    writes = setup_regs() + [(0x00, rs.prim_reg(6, iip=0)), (0x01, rs.rgbaq(0x11223344)),
                             (0x05, rs.xyz2(OFS + 3 * 16 + 8, OFS + 5 * 16, 0)),
                             (0x05, rs.xyz2(OFS + 40 * 16 + 8, OFS + 30 * 16, 0))]
    o.gif_transfer(gif.arr(gif.ad(writes)))
    o.flush()
    planes = o.planes()
    img = plane_image(planes, 0)
    exp = np.zeros((H, W), dtype=np.uint64)
    exp[5:30, 4:41] = 0x11223344
    assert np.array_equal(img, exp)
    assert not planes[PLANE:].any()


def test_incremental_pass_reads_samples():
    """A later pass over anti-aliased edges starts from the stored samples: an identity blend
    ((Cd - 0) * 128 >> 7) + 0 keeps every sample, so the reference plane, the samples and the
    resolved values all stay (had the pass started from plane 0, the samples would collapse to
    the resolved value and the reference would clear)."""
    verts = [(4 * 16, 4 * 16), (60 * 16 + 6, 10 * 16 + 2), (20 * 16 + 4, 44 * 16 + 8)]
    o = draw_triangle(4, verts, 0x80808080)
    before = o.planes().copy()
    # Alpha is not blended; FBMSK keeps it.
    ident = [(0x4C, rs.frame(0, 1, 0x00, 0xFF000000)), (0x42, rs.alpha_reg(1, 2, 2, 2, 128)),
             (0x00, rs.prim_reg(6, iip=0, abe=1)),
             (0x01, rs.rgbaq(0xFFFFFFFF)), (0x05, rs.xyz2(OFS, OFS, 0)), (0x05, rs.xyz2(OFS + W * 16, OFS + H * 16, 0))]
    o.gif_transfer(gif.arr(gif.ad(ident)))
    o.flush()
    after = o.planes()
    assert np.array_equal(after, before)


def test_flat_ui_triangle_is_demoted():
    """A flat triangle (equal depths) with no 3D geometry in its depth bucket is snapped to
    the pixel grid: single-sampled coverage at pixel corners, reference plane clear."""
    verts = [(4 * 16, 4 * 16), (60 * 16 + 6, 10 * 16 + 2), (20 * 16 + 4, 44 * 16 + 8)]
    o = draw_triangle(4, verts, 0x80808080, zs=(0x5000, 0x5000, 0x5000))
    planes = o.planes()
    img = plane_image(planes, 0)
    for y in range(H):
        for x in range(W):
            assert img[y, x] == (0x80808080 if covered(verts, x * 16, y * 16) else 0), (x, y)
    assert not planes[PLANE:].any()


def test_flat_triangle_sharing_3d_vertex_stays_antialiased():
    """A flat triangle that shares a vertex with a 3D triangle in the same depth bucket is not
    demoted: its coverage is per sample (attributes stay snapped to the pixel)."""
    flat = [(4 * 16, 4 * 16), (60 * 16 + 6, 10 * 16 + 2), (20 * 16 + 4, 44 * 16 + 8)]
    o = Oracle(2)
    writes = setup_regs() + [(0x00, rs.prim_reg(3, iip=0))]
    # 3D triangle sharing flat[0] (same x, y, z) with varying depth elsewhere, off to the side.
    tri3d = [((4 * 16, 4 * 16), 0x5000), ((2 * 16, 2 * 16), 0x5100), ((2 * 16, 6 * 16), 0x5200)]
    for (x, y), z in tri3d:
        writes += [(0x01, rs.rgbaq(0x40404040)), (0x05, rs.xyz2(OFS + x, OFS + y, z))]
    for (x, y) in flat:
        writes += [(0x01, rs.rgbaq(0x80808080)), (0x05, rs.xyz2(OFS + x, OFS + y, 0x5000))]
    o.gif_transfer(gif.arr(gif.ad(writes)))
    o.flush()
    planes = o.planes()
    ref = plane_image(planes, 1)
    pos = sample_positions(4)
    partial = 0
    for y in range(8, H):
        for x in range(8, W):
            k = sum(covered(flat, x * 16 + dx, y * 16 + dy) for dx, dy in pos)
            if 0 < k < 4:
                partial += 1
                assert ref[y, x] == 0xFFFFFFFF, (x, y, k)
    assert partial > 20


def test_depth_resolves_to_minimum():
    """Depth-tested 3D triangle: where samples differ, plane 0 holds the minimum sample depth
    and the sample planes hold each sample's depth."""
    verts = [(4 * 16, 4 * 16), (60 * 16 + 6, 10 * 16 + 2), (20 * 16 + 4, 44 * 16 + 8)]
    o = draw_triangle(4, verts, 0x80808080, zs=(0x100000, 0x900000, 0x400000),
                      extra=[(0x47, rs.test_reg(zte=1, ztst=1))])
    planes = o.planes()
    zimg = [plane_image(planes, p, 0x31, 16 * 32) & 0xFFFFFF for p in range(6)]
    zref = plane_image(planes, 1, 0x31, 16 * 32)
    diff = zref == 0xFFFFFF
    assert diff.sum() > 100
    mins = np.minimum.reduce([zimg[2 + s] for s in range(4)])
    assert np.array_equal(zimg[0][diff], mins[diff])
