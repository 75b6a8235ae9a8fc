"""GPU: sampler feedback (sparse texture decode). With the feedback forced on, every render pass
decodes only the 8x8 texel blocks its primitives can sample, so a block the marking missed would
be sampled stale and show up as a mismatch: the VRAM must equal the oracle's (which decodes
every texel) on random textured streams (all wrap modes, region clamp and repeat, ST/Q,
bilinear, sprites, lines, textures read from the frame buffer) at 1x and 4x, and on a large
texture sampled through small windows, where most blocks stay undecoded (2^20 texels: the
automatic setting turns the feedback on there)."""
import numpy as np
import pytest

import gif
import pgs
import render_streams as R
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


def render(stream, ssaa, mode):
    print(f"[case] stream={len(stream)} bytes ssaa={ssaa} sampler_feedback={mode}", flush=True)
    with pgs.Context(super_sampling=ssaa, sampler_feedback=mode) as ctx:
        ctx.gif_transfer(stream)
        return ctx.vram_planes()


def oracle(stream, ssaa):
    o = Oracle(LOG2[ssaa])
    o.gif_transfer(stream)
    o.flush()
    return o.planes()


def assert_equal(got, ref):
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


@pytest.mark.parametrize("ssaa", [1, 4])
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_feedback_random_textured(hip_device, seed, ssaa):
    stream = R.random_textured_stream(seed)
    assert_equal(render(stream, ssaa, "always"), oracle(stream, ssaa))


def large_texture_windows(seed: int, n_sprites: int = 60, width: int = 256, height: int = 192) -> np.ndarray:
    """A 1024x1024 PSMT8 texture with a 256-entry CLUT, sampled by small sprites and triangles
    whose texture windows (8..48 texels) sit anywhere, including across the wrap edge, with
    REPEAT, CLAMP and REGION_CLAMP, nearest and bilinear, UV and ST/Q."""
    rng = np.random.default_rng(seed)
    ofs = 2048 * 16
    tbp, tbw, clut_bp = 256 * 32, 16, 400 * 32
    out = [gif.ad([(0x1A, 1), (0x18, ofs | ofs << 32)])]
    out.append(gif.upload(tbp, tbw, 0x13, 0, 0, 1024, 1024, rng.integers(0, 256, 1 << 20, dtype=np.uint8).tobytes()))
    out.append(gif.upload(clut_bp, 1, 0x00, 0, 0, 16, 16, rng.integers(0, 2**32, 256, dtype=np.uint32).tobytes()))
    fbw = (width + 63) // 64
    for batch in range(6):
        wms = int(rng.choice([0, 0, 1, 2]))
        cl = R.clamp_reg(wms, wms, 100, 900, 120, 880)
        mmag = int(rng.integers(0, 2))
        fst = int(batch % 3 != 2)
        regs = [
            (0x4C, R.frame(0, fbw, 0x00)),
            (0x4E, R.zbuf(64, 1, 1)),
            (0x40, R.scissor(0, width - 1, 0, height - 1)),
            (0x47, R.test_reg(zte=1, ztst=1)),
            (0x3B, R.texa()),
            (0x14, R.tex1(mmag, mmag)),
            (0x08, cl),
            (0x06, R.tex0(tbp, tbw, 0x13, 10, 10, tcc=1, tfx=1, cbp=clut_bp, cpsm=0, csm=0, cld=1)),
            (0x00, R.prim_reg(6 if batch % 2 == 0 else 3, tme=1, fst=fst)),
        ]
        out.append(gif.ad(regs))
        writes = []
        nv = 2 if batch % 2 == 0 else 3
        for p in range(n_sprites // 6):
            x0, y0 = int(rng.integers(0, width - 48)), int(rng.integers(0, height - 48))
            size = int(rng.integers(8, 48))
            u0, v0 = float(rng.integers(-40, 1060)), float(rng.integers(-40, 1060))
            corners = [(0, 0), (size, size)] if nv == 2 else [(0, 0), (size, 0), (0, size)]
            for dx, dy in corners:
                q = float(rng.uniform(0.5, 2.0)) if not fst else 1.0
                writes.append((0x01, R.rgbaq(int(rng.integers(0, 2**32)), q)))
                if fst:
                    writes.append((0x03, R.uv(int((u0 + dx) * 16) & 0x3FFF, int((v0 + dy) * 16) & 0x3FFF)))
                else:
                    writes.append((0x02, R.st((u0 + dx) / 1024.0 * q, (v0 + dy) / 1024.0 * q)))
                writes.append((0x05, R.xyz2(ofs + (x0 + dx) * 16, ofs + (y0 + dy) * 16, int(rng.integers(0, 2**24)))))
        out.append(gif.ad(writes))
    return gif.arr(b"".join(out))


@pytest.mark.parametrize("ssaa", [1, 4])
def test_feedback_large_texture_windows(hip_device, ssaa):
    stream = large_texture_windows(21 + ssaa)
    ref = oracle(stream, ssaa)
    assert_equal(render(stream, ssaa, "always"), ref)
    assert_equal(render(stream, ssaa, "never"), ref)
    assert_equal(render(stream, ssaa, "auto"), ref)


def texture_cache_stream(seed: int) -> np.ndarray:
    """Passes that reuse a texture and passes after each kind of write that invalidates it:
    - pass 1 decodes a 64x64 CT32 texture, pass 2 (another frame buffer) reuses it;
    - an upload into the texture's pages, then pass 3 must decode again, pass 4 reuses;
    - a T8 texture with CLUT load (pass 5), reused without a load (pass 6), CLD=1 again (a new
      CLUT load) decodes again (pass 7);
    - a pass drawing into the texture's pages (pass 8), after which pass 9 decodes again.
    Expected reuses: passes 2, 4 and 6."""
    rng = np.random.default_rng(seed)
    ofs = 2048 * 16
    out = [gif.ad([(0x1A, 1), (0x18, ofs | ofs << 32)])]
    t32, t8, clut = 300 * 32, 320 * 32, 340 * 32
    out.append(gif.upload(t32, 1, 0x00, 0, 0, 64, 64, rng.integers(0, 2**32, 64 * 64, dtype=np.uint32).tobytes()))
    out.append(gif.upload(t8, 1, 0x13, 0, 0, 64, 64, rng.integers(0, 256, 64 * 64, dtype=np.uint8).tobytes()))
    out.append(gif.upload(clut, 1, 0x00, 0, 0, 16, 16, rng.integers(0, 2**32, 256, dtype=np.uint32).tobytes()))

    def draw(fbp, tex0, textured=True, full=False):
        regs = [(0x4C, R.frame(fbp, 4, 0x00)), (0x4E, R.zbuf(64, 1, 1)), (0x40, R.scissor(0, 255, 0, 191)),
                (0x47, R.test_reg(zte=1, ztst=1)),
                (0x14, R.tex1(1, 1)), (0x08, R.clamp_reg(0, 0)), (0x06, tex0),
                (0x00, R.prim_reg(6, tme=int(textured), fst=1))]
        writes = []
        for k in range(1 if full else 4):
            x0, y0, size = int(rng.integers(0, 200)), int(rng.integers(0, 140)), int(rng.integers(8, 48))
            if full:
                x0, y0, size = 0, 0, 256
            u0, v0 = int(rng.integers(0, 64)), int(rng.integers(0, 64))
            for dx in (0, size):
                writes.append((0x01, R.rgbaq(int(rng.integers(0, 2**32)))))
                writes.append((0x03, R.uv((u0 + dx) * 16, (v0 + dx) * 16)))
                writes.append((0x05, R.xyz2(ofs + (x0 + dx) * 16, ofs + (y0 + dx) * 16, 0)))
        out.append(gif.ad(regs + writes))

    T32 = R.tex0(t32, 1, 0x00, 6, 6, tcc=1, tfx=1)
    T8_load = R.tex0(t8, 1, 0x13, 6, 6, tcc=1, tfx=1, cbp=clut, cpsm=0, csm=0, cld=1)
    T8_keep = R.tex0(t8, 1, 0x13, 6, 6, tcc=1, tfx=1, cbp=clut, cpsm=0, csm=0, cld=0)
    draw(0, T32)                                   # 1: decode
    draw(100, T32)                                 # 2: reuse
    out.append(gif.upload(t32, 1, 0x00, 16, 16, 8, 8, rng.integers(0, 2**32, 64, dtype=np.uint32).tobytes()))
    draw(0, T32)                                   # 3: decode (pages written)
    draw(100, T32)                                 # 4: reuse
    draw(0, T8_load)                               # 5: CLUT load, decode
    draw(100, T8_keep)                             # 6: reuse
    draw(0, T8_load)                               # 7: CLUT load again, decode
    draw(300, T32, textured=False, full=True)      # 8: draws over the T32 pages
    draw(0, T32)                                   # 9: decode
    draw(100, T32)                                 # flushes pass 9
    return gif.arr(b"".join(out))


def test_texture_cache_reuse_and_invalidation(hip_device):
    stream = texture_cache_stream(31)
    print(f"[case] texture cache stream {len(stream)} bytes", flush=True)
    with pgs.Context(sampler_feedback="never") as ctx:
        ctx.gif_transfer(stream)
        got = ctx.vram_planes()
        stats = ctx.consume_flush_stats()
    assert_equal(got, oracle(stream, 1))
    print("decoded", stats["textures_decoded"], "reused", stats["texture_cache_hits"], flush=True)
    assert stats["texture_cache_hits"] == 4  # passes 2, 4, 6 and the last pass (reusing pass 9's decode)
    assert stats["textures_decoded"] == 5    # passes 1, 3, 5, 7, 9
