"""GPU parity of the CRTC scanout (pgs_vsync) against the oracle.

GSInterface::vsync / GSRenderer::vsync (gs/gs_interface.cpp:4626-4695, gs/gs_renderer.cpp:
4055-5099): for each display set-up the frame image, its metadata and every VRAM plane after
the call (EXTWRITE writes VRAM) must equal the oracle's bit for bit. Set-ups cover the video
modes, 1:1, magnified and MAGH-adapted circuits, CRTC offsets and overscan with clipped
viewports, both merge modes over BGCOLOR and circuit 2, anti-blur, field and frame interlacing
with a multi-field FastMAD sequence, EXTWRITE conversions, and high-resolution scanout of
rendered 4x / 8x / 16x frames (textured, blended, the 640x448 workload of the SSAA
configuration).
"""
import numpy as np
import pytest

import crtc
import pgs
from oracle_lib import Oracle, VRAM_SIZE

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


class Pair:
    """A GPU context and an oracle fed the same VRAM, streams and registers."""

    def __init__(self, samples=1, sparse=False):
        self.ctx = pgs.Context(super_sampling=samples, ordered=not sparse)
        self.orc = Oracle(LOG2[samples], sparse)

    def close(self):
        self.ctx.close()

    def write_vram(self, data, offset=0):
        self.ctx.write_vram(data, offset)
        self.orc.host_write(data, offset)

    def stream(self, s):
        self.ctx.gif_transfer(s)
        self.orc.gif_transfer(s)

    def priv(self, block):
        self.ctx.set_priv_registers(np.frombuffer(block, np.uint8))
        self.orc.set_priv(block)

    def vsync(self, what, **info):
        print(f"[case] {what} {info}", flush=True)
        rc_g, meta_g, img_g = self.ctx.vsync(**info)
        rc_o, meta_o, img_o = self.orc.vsync(**info)
        assert rc_g == rc_o, f"{what}: status {rc_g} vs oracle {rc_o}"
        assert meta_g == meta_o, f"{what}: {meta_g} vs oracle {meta_o}"
        if img_o is None:
            assert img_g is None, what
        else:
            assert img_g is not None and img_g.shape == img_o.shape, what
            bad = np.argwhere(img_g != img_o)
            assert bad.size == 0, (f"{what}: {len(bad)} pixels differ, first {bad[:4].tolist()}: "
                                   f"{[hex(img_g[tuple(b)]) for b in bad[:4]]} vs "
                                   f"{[hex(img_o[tuple(b)]) for b in bad[:4]]}")
        return img_o, meta_o

    def assert_vram_equal(self, what):
        g = self.ctx.vram_planes()
        self.orc.flush()
        o = self.orc.planes()
        assert np.array_equal(g, o), f"{what}: VRAM differs at {np.flatnonzero(g != o)[:8]}"


def random_vram(seed):
    return np.random.default_rng(seed).integers(0, 256, VRAM_SIZE, dtype=np.uint8)


PROG = dict(force_progressive=1)

SETUPS = [
    ("ct32 frame 1:1", crtc.ntsc_interlaced_640x448(), PROG),
    ("ct24", crtc.ntsc_interlaced_640x448(psm=0x01), PROG),
    ("ct16", crtc.ntsc_interlaced_640x448(psm=0x02), PROG),
    ("ct16s", crtc.ntsc_interlaced_640x448(psm=0x0A), PROG),
    ("raw circuit", crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(dh=399)),
     dict(force_progressive=1, raw_circuit_scanout=1, adapt_to_internal_horizontal_resolution=1)),
    ("field 0", crtc.ntsc_interlaced_640x448(), dict(phase=0, skip_deinterlace=1)),
    ("field 1", crtc.ntsc_interlaced_640x448(), dict(phase=1, skip_deinterlace=1)),
    ("ffmd frame", crtc.ntsc_interlaced_640x448(ffmd=1, DISPLAY1=crtc.display(dh=223)), dict(phase=1)),
    ("background", crtc.priv_block(PMODE=crtc.pmode(en1=0), SMODE1=crtc.smode1(), SMODE2=crtc.smode2(1, 0),
                                   BGCOLOR=crtc.bgcolor(12, 200, 77)), PROG),
    ("constant alpha", crtc.ntsc_interlaced_640x448(PMODE=crtc.pmode(alp=77, slbg=1),
                                                    BGCOLOR=crtc.bgcolor(30, 140, 250)), PROG),
    ("source alpha over circuit 2",
     crtc.priv_block(PMODE=crtc.pmode(en1=1, en2=1, mmod=0, slbg=0), SMODE1=crtc.smode1(), SMODE2=crtc.smode2(1, 0),
                     DISPFB1=crtc.dispfb(0), DISPLAY1=crtc.display(), DISPFB2=crtc.dispfb(200, psm=0x02),
                     DISPLAY2=crtc.display(dx=700, dy=60, dw=2047, dh=299)), PROG),
    ("circuit 2 hidden by SLBG", crtc.priv_block(PMODE=crtc.pmode(en1=1, en2=1, mmod=0, slbg=1),
                                                 SMODE1=crtc.smode1(), SMODE2=crtc.smode2(1, 0),
                                                 DISPFB1=crtc.dispfb(0), DISPLAY1=crtc.display(),
                                                 DISPFB2=crtc.dispfb(200), DISPLAY2=crtc.display(),
                                                 BGCOLOR=crtc.bgcolor(1, 2, 3)), PROG),
    ("magnified 2x", crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(magh=7)), PROG),
    ("magnified 3x vertical", crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(magh=2, magv=2)), PROG),
    ("adapted to MAGH", crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(magh=7)),
     dict(force_progressive=1, adapt_to_internal_horizontal_resolution=1)),
    ("crtc offsets", crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(dx=500, dy=30)),
     dict(force_progressive=1, crtc_offsets=1)),
    ("overscan", crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(dx=700, dy=70)),
     dict(force_progressive=1, overscan=1)),
    ("pal", crtc.ntsc_interlaced_640x448(SMODE1=crtc.smode1(crtc.CMOD_PAL), DISPLAY1=crtc.display(dx=676, dy=72,
                                                                                                  dh=511)), PROG),
    ("vga", crtc.priv_block(PMODE=crtc.pmode(), SMODE1=crtc.smode1(crtc.CMOD_PROGRESSIVE, crtc.LC_VGA),
                            SMODE2=crtc.smode2(0, 0), DISPFB1=crtc.dispfb(), DISPLAY1=crtc.display(
                                dx=276, dy=34, magh=1, dw=1279, dh=479)), {}),
    ("hdtv 720p", crtc.priv_block(PMODE=crtc.pmode(), SMODE1=crtc.smode1(crtc.CMOD_PROGRESSIVE, crtc.LC_HDTV),
                                  SMODE2=crtc.smode2(0, 0), DISPFB1=crtc.dispfb(fbw=20), DISPLAY1=crtc.display(
                                      dx=300, dy=0, magh=0, dw=1279, dh=719)), {}),
    ("double strike", crtc.ntsc_interlaced_640x448(SMODE2=crtc.smode2(0, 0), DISPLAY1=crtc.display(dh=223)), {}),
    ("anti-blur", crtc.priv_block(PMODE=crtc.pmode(en1=1, en2=1, mmod=1, alp=0x80), SMODE1=crtc.smode1(),
                                  SMODE2=crtc.smode2(1, 0), DISPFB1=crtc.dispfb(dby=0), DISPFB2=crtc.dispfb(dby=1),
                                  DISPLAY1=crtc.display(), DISPLAY2=crtc.display()),
     dict(phase=0, skip_deinterlace=1, anti_blur=1)),
    ("unknown mode", crtc.ntsc_interlaced_640x448(SMODE1=crtc.smode1(lc=5)), PROG),
    ("palette format refused", crtc.ntsc_interlaced_640x448(psm=0x13), PROG),
]


def test_scanout_setups_match_oracle(hip_device):
    p = Pair()
    p.write_vram(random_vram(1))
    for name, block, info in SETUPS:
        p.priv(block)
        p.vsync(name, **info)
    p.assert_vram_equal("after scanouts")
    p.close()


def test_deinterlace_sequence(hip_device):
    """Eight fields, alternating phase, with part of the picture changing between fields: the
    FastMAD output (weave where still, bob where moving) equals the oracle's every field."""
    p = Pair()
    rng = np.random.default_rng(7)
    p.write_vram(random_vram(2))
    p.priv(crtc.ntsc_interlaced_640x448())
    for k in range(8):
        band = rng.integers(0, 256, 64 * 1024, dtype=np.uint8)
        p.write_vram(band, int(rng.integers(0, 1_000_000 // 256)) * 256)
        p.vsync(f"field {k}", phase=k & 1)
    # Fewer displayed lines keep the field size (it follows the mode); a PAL field is larger and
    # restarts the history.
    p.priv(crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(dh=399)))
    p.vsync("fewer lines", phase=0)
    p.priv(crtc.ntsc_interlaced_640x448(SMODE1=crtc.smode1(crtc.CMOD_PAL)))
    p.vsync("pal field", phase=1)
    p.priv(crtc.ntsc_interlaced_640x448())
    p.vsync("progressive resets the history", **PROG)
    p.vsync("first field again", phase=1)
    p.close()


@pytest.mark.parametrize("samples", [1, 4])
def test_extwrite_modes(hip_device, samples):
    """EXTWRITE for every EMODC x EMODA pair, from the merged frame and from a circuit, in field
    and frame write modes; VRAM (and, with super-sampling, the cleared reference plane) equals the
    oracle's."""
    p = Pair(samples)
    p.write_vram(random_vram(3))
    p.stream(pgs.gen_textured(0x5EED0031, 300, 1, 640, 448, 0.0))
    for emodc in range(4):
        for emoda in range(4):
            fbin = (emodc + emoda) & 1
            block = crtc.ntsc_interlaced_640x448(
                PMODE=crtc.pmode(en1=1, en2=fbin, mmod=1, alp=0xFF),
                DISPFB2=crtc.dispfb(100), DISPLAY2=crtc.display(),
                EXTBUF=crtc.extbuf(exbp=(300 + emodc * 8) * 32, exbw=4, fbin=fbin, wffmd=emoda & 1, emoda=emoda,
                                   emodc=emodc, wdx=emodc * 3, wdy=emoda * 5),
                EXTDATA=crtc.extdata(smph=3 if emodc != 2 else 1, smpv=emoda & 1, ww=239, wh=127), EXTWRITE=1)
            p.priv(block)
            p.vsync(f"extwrite emodc {emodc} emoda {emoda} fbin {fbin}", **PROG)
    p.assert_vram_equal("after EXTWRITE")
    # A later draw reads what EXTWRITE wrote through the texture cache: same again.
    p.stream(pgs.gen_textured(0x5EED0032, 200, 1, 640, 448, 0.0))
    p.assert_vram_equal("after a draw over the written pages")
    p.close()


@pytest.mark.parametrize("samples", [4, 8, 16])
def test_high_resolution_scanout_of_rendered_frames(hip_device, samples):
    """A textured, blended frame rendered at 4x / 8x / 16x, scanned out at twice the resolution
    from its super-samples (progressive), and at single resolution (interlaced fields)."""
    p = Pair(samples)
    p.stream(pgs.gen_textured(0x5EED0040 + samples, 1500, 1, 640, 448, 0.0))
    p.priv(crtc.ntsc_interlaced_640x448())
    img, meta = p.vsync(f"{samples}x high resolution", force_progressive=1, high_resolution_scanout=1)
    assert meta["high_resolution_scanout"] == 1 and img.shape == (896, 1280)
    p.vsync(f"{samples}x raw high resolution", force_progressive=1, high_resolution_scanout=1,
            raw_circuit_scanout=1, adapt_to_internal_horizontal_resolution=1)
    p.vsync(f"{samples}x field", phase=1, high_resolution_scanout=1)
    p.vsync(f"{samples}x progressive", force_progressive=1)
    p.close()


def test_high_resolution_scanout_of_a_sparse_16x_frame(hip_device):
    """The sparse 2 x 8 grid at 16x: high-resolution scanout averages each quad's four sample
    planes as for the ordered grid (sample_circuit.frag:110-140)."""
    p = Pair(16, sparse=True)
    p.stream(pgs.gen_textured(0x5EED0051, 1500, 1, 640, 448, 0.0))
    p.priv(crtc.ntsc_interlaced_640x448())
    img, meta = p.vsync("sparse 16x high resolution", force_progressive=1, high_resolution_scanout=1)
    assert meta["high_resolution_scanout"] == 1 and img.shape == (896, 1280)
    p.close()
    p = Pair(4, sparse=True)  # 1 x 4: no x samples, so no high-resolution scanout
    p.stream(pgs.gen_textured(0x5EED0052, 800, 1, 640, 448, 0.0))
    p.priv(crtc.ntsc_interlaced_640x448())
    img, meta = p.vsync("sparse 4x", force_progressive=1, high_resolution_scanout=1)
    assert meta["high_resolution_scanout"] == 0
    p.close()


def test_field_aware_request_is_refused_and_can_skip(hip_device):
    p = Pair(4)
    p.priv(crtc.ntsc_interlaced_640x448(ffmd=1))
    rc, _, img = p.ctx.vsync(force_progressive=1, high_resolution_scanout=1)
    assert rc == pgs.ERROR_UNSUPPORTED and img is None
    for info in ({}, PROG, dict(phase=1)):
        assert p.ctx.vsync_can_skip(**info) == p.orc.vsync_can_skip(**info)
    p.close()


def test_vsync_timestamps(hip_device):
    p = Pair()
    assert p.ctx.set_debug_mode(timestamps=1) == 0
    p.priv(crtc.ntsc_interlaced_640x448())
    p.vsync("timed", phase=0)
    ts = p.ctx.accumulated_timestamps()
    assert ts["vsync"] > 0.0
    p.close()
