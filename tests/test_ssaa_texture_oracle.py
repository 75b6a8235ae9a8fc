"""Pins the oracle's super-sampled textures (gs/gs_interface.cpp:1593-1603, 1783-1831,
gs/gs_renderer.cpp:94-313, triangle_setup.comp:677-789, upload.comp:133-228,
ubershader.comp:1155-1211) on the CPU:

* the phase filter table against a brute force written from its definition: the sample points
  are the rasterizer's own sample offsets, the taps are the four nearest points of the 3 x 3
  texels (Chebyshev distance) with the first in the centre texel, the tent radius is the
  smallest fifth distance over all phases, the weights are the float32 tent values rounded to
  unorm8 with the remainder on the last tap; NEAREST entries are the nearest point;
* closed forms of whole streams: a 1:1 NEAREST sprite blit of a super-sampled frame copies every
  sample (each sample reads its own layer), and without super-sampled textures every sample reads
  the single-sampled value; pages last written by an upload, triangles and 4-bit views never
  sample per sample (the output equals the run without the option)."""
import numpy as np
import pytest

import gif
import render_streams as rs
from oracle_lib import Oracle, oracle_phase_lut, swizzle

GRIDS = [(0, 1), (1, 1), (0, 2), (1, 2), (2, 2), (1, 3)]  # (rate_x, rate_y): 2x, 4x, sparse 4x, 8x, 16x, sparse 16x


def sample_points(rx, ry):
    """The shading kernels' sample offsets (in 1 / 2^ry pixel) of every grid."""
    pts = []
    for s in range(1 << (rx + ry)):
        if (rx, ry) == (0, 1):
            dx, dy = s, s
        elif (rx, ry) == (1, 2):
            dx, dy = 2 * ((s >> 1) & 1) + (s & 1), (s & 1) + 2 * ((s >> 2) & 1)
        elif (rx, ry) == (0, 2):
            dx, dy = [0, 2, 3, 1][s], s
        elif (rx, ry) == (1, 3):
            dx, dy = 4 * ((s >> 3) & 1) + [0, 2, 3, 1][s & 3], s & 7
        else:
            dx, dy = ((s >> 1) & 1) + 2 * ((s >> 3) & 1), (s & 1) + 2 * ((s >> 2) & 1)
        pts.append((dx, dy))
    return pts


def candidates(rx, ry, x, y, nearest=False, reach=1):
    cell = 64 >> ry
    out = []
    for s, (dx, dy) in enumerate(sample_points(rx, ry)):
        px, py = dx * cell, dy * cell
        if nearest:
            px += (cell - 1) >> 1
            py += (cell - 1) >> 1
        for ty in range(-reach, reach + 1):
            for tx in range(-reach, reach + 1):
                ex, ey = abs(px + 64 * tx - x), abs(py + 64 * ty - y)
                out.append((max(ex, ey), ey, s, tx, ty, ex))
    return out


@pytest.mark.parametrize("rx,ry", GRIDS)
def test_phase_lut_brute_force(rx, ry):
    lut = oracle_phase_lut(rx, ry)
    falloff = min(sorted(c[0] for c in candidates(rx, ry, x, y))[4] for y in range(64) for x in range(64))
    for y in range(64):
        for x in range(64):
            w0, w1 = int(lut[2 * (64 * y + x)]), int(lut[2 * (64 * y + x) + 1])
            cands = candidates(rx, ry, x, y)
            dist = {(c[2], c[3], c[4]): c for c in cands}
            taps = []
            for i in range(4):
                s = (w0 >> (4 * i)) & 15
                tx = ((w0 >> (16 + 4 * i)) & 3) - (4 if (w0 >> (17 + 4 * i)) & 1 else 0)
                ty = ((w0 >> (18 + 4 * i)) & 3) - (4 if (w0 >> (19 + 4 * i)) & 1 else 0)
                taps.append(dist[(s, tx, ty)])
            assert taps[0][3] == 0 and taps[0][4] == 0, (x, y)
            assert len({(t[2], t[3], t[4]) for t in taps}) == 4
            order = sorted(c[0] for c in cands)
            centre = min(c[0] for c in cands if c[3] == 0 and c[4] == 0)
            assert taps[0][0] == centre
            if order[0] == centre and sum(1 for c in cands if c[0] <= order[3]) == 4:
                # No ties at the cut and the nearest point is central: the four nearest points.
                assert sorted(t[0] for t in taps) == order[:4]
            wt = [np.float32(max(0.0, 1.0 - np.float32(t[5]) / np.float32(falloff))) *
                  np.float32(max(0.0, 1.0 - np.float32(t[1]) / np.float32(falloff))) for t in taps]
            total = np.float32(0)
            for v in wt:
                total = np.float32(total + v)
            if total < np.float32(0.001):
                wt, total = [np.float32(1)] * 4, np.float32(4)
            used, want = 0, []
            for i in range(3):
                v = int(np.float32(np.float32(wt[i] * np.float32(255)) / total) + np.float32(0.5))
                v = min(v, 255 - used)
                used += v
                want.append(v)
            want.append(255 - used)
            got = [(w1 >> (8 * i)) & 255 for i in range(4)]
            assert got == want and sum(got) == 255, (x, y, got, want)
    for y in range(32):
        for x in range(32):
            cands = candidates(rx, ry, 2 * x, 2 * y, nearest=True, reach=0)
            best = min(c[:2] for c in cands)
            assert (min(c[0] for c in cands), ) and lut[2 * (4096 + 32 * y + x)] in \
                {c[2] for c in cands if c[:2] == best}


def source_stream(seed, src_psm=0x00):
    """A 64 x 32 CT32 source frame at FBP 0 drawn with 3D triangles at the context's rate."""
    rng = np.random.default_rng(seed)
    ofs = 2048 * 16
    out = [gif.ad([(0x1A, 1), (0x18, ofs | ofs << 32)]),
           gif.upload(0, 1, 0x00, 0, 0, 64, 32, rng.integers(0, 256, 64 * 32 * 4, dtype=np.uint8).tobytes()),
           gif.ad([(0x4C, rs.frame(0, 1, src_psm)), (0x40, rs.scissor(0, 63, 0, 31)),
                   (0x47, rs.test_reg(zte=0)), (0x42, rs.alpha_reg(0, 1, 0, 1, 0))])]
    writes = []
    for t in range(12):
        writes.append((0x00, rs.prim_reg(3, iip=1)))
        for v in range(3):
            writes.append((0x01, rs.rgbaq(int(rng.integers(0, 2**32)))))
            # Distinct depths make the triangles 3D, so they shade per sample.
            writes.append((0x05, rs.xyz2(ofs + int(rng.integers(-64, 70 * 16)), ofs + int(rng.integers(-64, 36 * 16)),
                                         int(rng.integers(1, 2**24)))))
    out.append(gif.ad(writes))
    return out


def blit_stream(prim=6, tpsm=0x00, mmag=0, d=0):
    """Sprite (or two triangles) from the source to a 64 x 32 CT32 frame at FBP 8, 1:1, DECAL."""
    ofs = 2048 * 16
    w = [(0x4C, rs.frame(8, 1, 0x00)), (0x40, rs.scissor(0, 63, 0, 31)), (0x47, rs.test_reg(zte=0)),
         (0x14, rs.tex1(mmag, mmag)), (0x08, 0), (0x06, rs.tex0(0, 1, tpsm, 6, 5, tcc=1, tfx=1))]
    if prim == 6:
        w.append((0x00, rs.prim_reg(6, iip=0, tme=1)))
        for (x, y) in [(0, 0), (64, 32)]:
            w += [(0x01, rs.rgbaq(0x80808080)), (0x03, rs.uv(x * 16 + d, y * 16 + d)), (0x05, rs.xyz2(ofs + x * 16, ofs + y * 16, 0))]
    else:
        for tri in [[(0, 0), (64, 0), (0, 32)], [(64, 0), (64, 32), (0, 32)]]:
            w.append((0x00, rs.prim_reg(3, iip=0, tme=1)))
            for (x, y) in tri:
                w += [(0x01, rs.rgbaq(0x80808080)), (0x03, rs.uv(x * 16 + d, y * 16 + d)), (0x05, rs.xyz2(ofs + x * 16, ofs + y * 16, 0))]
    return [gif.ad(w)]


def run(log2, ssaa_tex, parts, sparse=False):
    o = Oracle(log2, sparse, super_sampled_textures=ssaa_tex)
    for p in parts:
        o.gif_transfer(gif.arr(p))
    o.flush()
    return o.planes().reshape(-1, 4 << 20)


def sample_values(planes, fbp, log2):
    """Per-sample CT32 colours of the 64 x 32 frame at `fbp`: the sample plane's word where the
    reference plane marks it super-sampled, else plane 0's."""
    n = 1 << log2
    vals = np.zeros((32, 64, n), np.uint32)
    for y in range(32):
        for x in range(64):
            a = swizzle(x, y, fbp * 32, 1, 0x00) * 4
            w = [int.from_bytes(planes[k][a:a + 4].tobytes(), "little") for k in range(2 + n)]
            vals[y, x] = [w[2 + s] if w[1] == 0xFFFFFFFF else w[0] for s in range(n)]
    return vals


@pytest.mark.parametrize("log2,sparse", [(1, False), (2, False), (2, True), (3, False), (4, True)])
def test_one_to_one_nearest_blit_copies_every_sample(log2, sparse):
    parts = source_stream(10 + log2) + blit_stream()
    on = run(log2, True, parts, sparse)
    src = sample_values(on, 0, log2)
    assert (src != src[:, :, :1]).any(), "the source has no super samples"
    assert np.array_equal(sample_values(on, 8, log2), src)
    off = run(log2, False, parts, sparse)
    single = np.repeat(np.stack([[int.from_bytes(off[0][swizzle(x, y, 0, 1, 0) * 4:][:4].tobytes(), "little")
                                  for x in range(64)] for y in range(32)])[:, :, None], 1 << log2, axis=2)
    assert np.array_equal(sample_values(off, 8, log2), single.astype(np.uint32))


def test_uploaded_pages_and_triangles_do_not_sample_per_sample():
    up = [gif.upload(0, 1, 0x00, 0, 0, 64, 32, bytes(range(256)) * 32)]
    for parts in (source_stream(3) + up + blit_stream(),          # last written by an upload
                  source_stream(4) + blit_stream(prim=3),          # triangles
                  source_stream(5) + blit_stream(tpsm=0x24)):      # a 4-bit view
        assert np.array_equal(run(2, True, parts), run(2, False, parts))


def test_linear_blit_differs_and_one_x_ignores_option():
    parts = source_stream(6) + blit_stream(mmag=1, d=3)
    assert not np.array_equal(run(2, True, parts), run(2, False, parts))
    assert np.array_equal(run(0, True, parts), run(0, False, parts))


@pytest.mark.parametrize("seed", [0, 1])
def test_random_stream_runs(seed):
    """The random generator's streams exercise every sampling path (smoke of the oracle)."""
    o = Oracle(2, super_sampled_textures=True)
    o.gif_transfer(rs.random_ssaa_texture_stream(700 + seed, n_batches=2, sprites_per_batch=6))
    o.flush()
    assert o.planes().size == 6 * (4 << 20)
