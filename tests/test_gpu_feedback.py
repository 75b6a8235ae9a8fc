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
