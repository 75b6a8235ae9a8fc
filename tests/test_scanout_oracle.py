"""Oracle checks of the CRTC scanout (vsync) against closed forms (no GPU).

GSRenderer::vsync (gs/gs_renderer.cpp:4114-5099) with sample_circuit.frag, blit_circuit.frag,
extwrite.comp and weave.frag. Each test states what the mathematics fixes for a set-up the
reference handles: 1:1 scanout is a copy of the frame buffer; field mode reads every other line;
the deinterlacer keeps the current field's lines; constant and source-alpha merges are the
rational blends; the linear filter's weights at a 2x magnification are 1/4 and 3/4; EXTWRITE
with a matched sampling rate copies the frame back; the 4x quad layout reads the sample planes
in the order the renderer writes them.
"""
import functools

import numpy as np
import pytest

import crtc
from oracle_lib import Oracle, VRAM_SIZE, swizzle

W, H = 640, 448


@functools.lru_cache(maxsize=64)
def element_map(w, h, fbp, fbw, psm):
    return np.array([[swizzle(x, y, fbp * 32, fbw, psm) for x in range(w)] for y in range(h)], np.int64)


def frame_from_vram(vram, w, h, fbp=0, fbw=10, psm=0, plane=0):
    """Reads a w x h region of a frame buffer as RGBA8 words the way the CRTC expands it."""
    e = element_map(w, h, fbp, fbw, psm)
    p = np.asarray(vram[plane * VRAM_SIZE:(plane + 1) * VRAM_SIZE])
    if psm in (0x02, 0x0A, 0x32, 0x3A):
        v = p.view(np.uint16)[e].astype(np.uint32)
        return ((v & 31) << 3) | (((v >> 5) & 31) << 11) | (((v >> 10) & 31) << 19) | \
            np.where(v & 0x8000, np.uint32(0xFF000000), np.uint32(0))
    c = p.view(np.uint32)[e]
    if psm != 0x00:
        c = (c & np.uint32(0xFFFFFF)) | np.uint32(0x80000000)
    return c


def random_vram(seed, planes=1):
    return np.random.default_rng(seed).integers(0, 256, planes * VRAM_SIZE, dtype=np.uint8)


def make(seed=1, sample_log2=0, planes=None):
    o = Oracle(sample_log2)
    n = planes if planes is not None else (1 if sample_log2 == 0 else 2 + (1 << sample_log2))
    v = random_vram(seed, n)
    o.set_vram(v)
    return o, v


PROGRESSIVE = dict(force_progressive=1)


@pytest.mark.parametrize("psm", [0x00, 0x01, 0x02, 0x0A])
def test_one_to_one_scanout_copies_the_frame_buffer(psm):
    """NTSC frame mode, DX/DY at the scan offsets, MAGH+1 = clock divider: the merged frame is the
    frame buffer itself (16-bit: 5:5:5 expanded, alpha 0 / 255; 24-bit: alpha 0x80)."""
    o, v = make(psm + 3)
    o.set_priv(crtc.ntsc_interlaced_640x448(psm=psm))
    rc, meta, img = o.vsync(**PROGRESSIVE)
    assert rc == 0
    assert (meta["width"], meta["height"]) == (W, H)
    assert (meta["mode_width"], meta["mode_height"], meta["internal_width"]) == (640, 448, 640)
    assert meta["interlaced"] == 0 and meta["double_strike"] == 0
    assert abs(meta["sampling_rate_mhz"] - 27.0) < 1e-6  # composite clock 54/4, doubled when progressive
    expect = frame_from_vram(v, W, H, psm=psm)
    assert np.array_equal(img, expect)


def test_raw_circuit_returns_the_circuit():
    o, v = make(7)
    o.set_priv(crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(dh=399)))
    rc, meta, img = o.vsync(force_progressive=1, raw_circuit_scanout=1,
                            adapt_to_internal_horizontal_resolution=1)
    assert rc == 0 and img.shape == (400, 640)
    assert (meta["internal_width"], meta["internal_height"], meta["mode_height"]) == (640, 400, 448)
    assert np.array_equal(img, frame_from_vram(v, 640, 400))


def test_background_only():
    o, _ = make(2)
    o.set_priv(crtc.priv_block(PMODE=crtc.pmode(en1=0), SMODE1=crtc.smode1(), SMODE2=crtc.smode2(1, 0),
                               BGCOLOR=crtc.bgcolor(12, 200, 77)))
    rc, meta, img = o.vsync(**PROGRESSIVE)
    assert rc == 0 and img.shape == (448, 640)
    assert (img == (12 | 200 << 8 | 77 << 16)).all()  # clear alpha is 0


def test_unknown_video_mode_has_no_frame():
    o, _ = make(3)
    o.set_priv(crtc.ntsc_interlaced_640x448(SMODE1=crtc.smode1(lc=5)))
    rc, meta, img = o.vsync(**PROGRESSIVE)
    assert rc == 0 and img is None and meta["mode_width"] == 0


def channels(a):
    return np.stack([(a >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int64)


def check_rational(got, num, den):
    """got == round(num / den) except within 1e-3 of a tie, where it may be either neighbour."""
    exact = num / den
    lo, hi = np.floor(exact), np.floor(exact) + 1
    near_tie = np.abs(exact - lo - 0.5) < 1e-3
    ok = (got == np.rint(exact)) | (near_tie & ((got == lo) | (got == hi)))
    assert ok.all(), f"{np.count_nonzero(~ok)} pixels off the rational blend"


@pytest.mark.parametrize("alp", [0, 1, 77, 128, 254])
def test_constant_alpha_merge(alp):
    """MMOD = 1: circuit 1 * ALP/255 + background * (1 - ALP/255) on RGB; alpha is circuit 1's."""
    o, v = make(11)
    bg = (30, 140, 250)
    o.set_priv(crtc.ntsc_interlaced_640x448(PMODE=crtc.pmode(alp=alp, mmod=1, slbg=1),
                                            BGCOLOR=crtc.bgcolor(*bg)))
    rc, _, img = o.vsync(**PROGRESSIVE)
    src = channels(frame_from_vram(v, W, H)[:64, :64])
    got = channels(img[:64, :64])
    for ch in range(3):
        check_rational(got[..., ch], src[..., ch] * alp + bg[ch] * (255 - alp), 255)
    assert np.array_equal(got[..., 3], src[..., 3])


def test_source_alpha_merge_over_circuit2():
    """MMOD = 0, SLBG = 0: circuit 2 is copied first, circuit 1 blends over it with its own alpha."""
    o, v = make(12)
    o.set_priv(crtc.priv_block(PMODE=crtc.pmode(en1=1, en2=1, mmod=0, slbg=0), SMODE1=crtc.smode1(),
                               SMODE2=crtc.smode2(1, 0), DISPFB1=crtc.dispfb(0), DISPLAY1=crtc.display(),
                               DISPFB2=crtc.dispfb(200), DISPLAY2=crtc.display()))
    rc, _, img = o.vsync(**PROGRESSIVE)
    c1 = channels(frame_from_vram(v, 64, 64, fbp=0))
    c2 = channels(frame_from_vram(v, 64, 64, fbp=200))
    got = channels(img[:64, :64])
    a = c1[..., 3]
    for ch in range(3):
        check_rational(got[..., ch], c1[..., ch] * a + c2[..., ch] * (255 - a), 255)
    assert np.array_equal(got[..., 3], c1[..., 3])


def test_field_mode_reads_every_other_line():
    """INT = 1, FFMD = 0 without force_progressive: each field reads lines phase, phase + 2, ...
    (DY = 50 is even) into a 640 x 224 field."""
    for phase in (0, 1):
        o, v = make(21)
        o.set_priv(crtc.ntsc_interlaced_640x448())
        rc, meta, img = o.vsync(phase=phase, skip_deinterlace=1)
        assert img.shape == (224, 640) and meta["interlaced"] == 1 and meta["interlace_phase"] == phase
        assert abs(meta["sampling_rate_mhz"] - 13.5) < 1e-6
        full = frame_from_vram(v, W, H)
        assert np.array_equal(img, full[phase::2])


def test_deinterlace_keeps_field_lines_and_weaves_a_still_picture():
    """First field: the history is the field itself, so the missing lines weave in the same
    field's lines (no motion): line y of the phase's parity is frame line y; the others are
    field line y >> 1 with alpha 255, except the first and last, which copy the previous field."""
    o, v = make(22)
    o.set_priv(crtc.ntsc_interlaced_640x448())
    rc, meta, img = o.vsync(phase=0)
    assert img.shape == (448, 640)
    full = frame_from_vram(v, W, H)
    assert np.array_equal(img[0::2], full[0::2])
    woven = full[0::2][:-1] | np.uint32(0xFF000000)
    assert np.array_equal(img[1:-1:2], woven[: img[1:-1:2].shape[0]])
    assert np.array_equal(img[-1], full[-2])


def test_deinterlace_bobs_where_fields_differ():
    """FastMAD compares each field with the one two fields back (same parity). Three black frames
    then a white one: the current field differs from its predecessor by luma 1 > 0.06, so every
    interpolated line is the average of the current field's lines above and below (white, alpha
    255); with an unchanged picture the previous field is woven in instead."""
    o = Oracle(0)
    o.set_priv(crtc.ntsc_interlaced_640x448())
    for k, value in enumerate((0, 0, 0, 255)):
        o.set_vram(np.full(VRAM_SIZE, value, np.uint8))
        rc, meta, img = o.vsync(phase=k & 1)
    assert (img[1::2] == 0xFFFFFFFF).all()          # current field (phase 1)
    assert (img[2::2] == 0xFFFFFFFF).all()          # bobbed
    assert (img[0] == 0).all()                      # first line: previous field as is
    # Alternating black / white fields: same-parity fields agree, so nothing is bobbed and the
    # interpolated lines are the previous (black) field's with alpha 255.
    o = Oracle(0)
    o.set_priv(crtc.ntsc_interlaced_640x448())
    for k in range(4):
        o.set_vram(np.full(VRAM_SIZE, 255 if k & 1 else 0, np.uint8))
        rc, meta, img = o.vsync(phase=k & 1)
    assert (img[1::2] == 0xFFFFFFFF).all()
    assert (img[2::2] == 0xFF000000).all()


def test_linear_filter_weights_at_2x_magnification():
    """MAGH + 1 = 8 with the composite divider 4: a 320-wide circuit drawn over 640 pixels. Pixel
    2k + 1 samples texel coordinate k + 0.25 (weights 3/4, 1/4), pixel 2k samples k - 0.25; the
    ends clamp."""
    o, v = make(31)
    o.set_priv(crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(magh=7)))
    rc, meta, img = o.vsync(**PROGRESSIVE)
    src = channels(frame_from_vram(v, 320, 8))
    got = channels(img[:8])
    k = np.arange(320)
    left = np.clip(k - 1, 0, 319)
    right = np.clip(k + 1, 0, 319)
    even = 64 * src[:, left] + 192 * src[:, k]
    odd = 192 * src[:, k] + 64 * src[:, right]
    check_rational(got[:, 0::2], even, 256)
    check_rational(got[:, 1::2], odd, 256)


def test_adapting_to_magh_removes_the_magnification():
    o, v = make(32)
    o.set_priv(crtc.ntsc_interlaced_640x448(DISPLAY1=crtc.display(magh=7)))
    rc, meta, img = o.vsync(force_progressive=1, adapt_to_internal_horizontal_resolution=1)
    assert (meta["mode_width"], img.shape[1]) == (320, 320)
    assert abs(meta["sampling_rate_mhz"] - 13.5) < 1e-6
    assert np.array_equal(img, frame_from_vram(v, 320, 448))


def test_extwrite_copies_the_frame_back():
    """EXTWRITE with SMPH + 1 = the clock divider samples the merged frame pixel for pixel and
    writes it as PSMCT32 at EXBP / WDX / WDY (EMODC = EMODA = 0)."""
    o, v = make(41)
    o.set_priv(crtc.ntsc_interlaced_640x448(
        EXTBUF=crtc.extbuf(exbp=320 * 32, exbw=10, wdx=0, wdy=0), EXTDATA=crtc.extdata(smph=3, ww=639, wh=447),
        EXTWRITE=1))
    rc, meta, img = o.vsync(force_progressive=1)
    # EXTWRITE disables force_progressive's high-resolution path only; the frame is 640 x 448.
    assert img.shape == (448, 640)
    vram = o.vram()
    back = frame_from_vram(vram, W, H, fbp=320)
    assert np.array_equal(back, img)


def test_extwrite_luma():
    """EMODC = 1 writes Y = (219 Ey + 16) / 255 with Ey = 0.587 G + 0.114 B + 0.299 R in RGB,
    EMODA = 2 writes Y / 2 in alpha."""
    o, v = make(42)
    o.set_priv(crtc.ntsc_interlaced_640x448(
        EXTBUF=crtc.extbuf(exbp=320 * 32, exbw=10, emodc=1, emoda=2), EXTDATA=crtc.extdata(smph=3, ww=639, wh=447),
        EXTWRITE=1))
    rc, meta, img = o.vsync(force_progressive=1)
    src = channels(img[:64, :64]).astype(np.float64) / 255
    ey = 0.587 * src[..., 1] + 0.114 * src[..., 2] + 0.299 * src[..., 0]
    y = (219 * ey + 16) / 255 * 255
    got = channels(frame_from_vram(o.vram(), 64, 64, fbp=320))
    for ch in range(3):
        assert np.abs(got[..., ch] - y).max() <= 0.5 + 1e-3
    assert np.abs(got[..., 3] - y / 2).max() <= 0.5 + 1e-3


def test_high_resolution_quads_read_the_sample_planes():
    """4x (2 x 2 ordered grid), high-resolution scanout: image pixel (2x + i, 2y + j) is sample
    plane 2 + (j + 2 i) where the reference plane marks the pixel's samples valid, else the
    single-sampled pixel."""
    o, v = make(51, sample_log2=2)
    vv = v.copy()
    ref = vv[VRAM_SIZE:2 * VRAM_SIZE].view(np.uint32)
    rng = np.random.default_rng(5)
    ref[:] = np.where(rng.random(ref.size) < 0.5, 0xFFFFFFFF, 0x12345678).astype(np.uint32)
    o.set_vram(vv)
    o.set_priv(crtc.ntsc_interlaced_640x448())
    rc, meta, img = o.vsync(force_progressive=1, high_resolution_scanout=1)
    assert meta["high_resolution_scanout"] == 1 and img.shape == (896, 1280)
    w, h = 32, 16
    single = frame_from_vram(vv, w, h)
    for i in range(2):
        for j in range(2):
            plane = frame_from_vram(vv, w, h, plane=2 + j + 2 * i)
            valid = np.array([[ref[swizzle(x, y, 0, 10, 0)] == 0xFFFFFFFF for x in range(w)] for y in range(h)])
            expect = np.where(valid, plane, single)
            assert np.array_equal(img[j:2 * h:2, i:2 * w:2], expect)


def test_high_resolution_16x_averages_four_samples():
    o, v = make(52, sample_log2=4)
    vv = v.copy()
    vv[VRAM_SIZE:2 * VRAM_SIZE] = 0xFF  # every pixel valid
    o.set_vram(vv)
    o.set_priv(crtc.ntsc_interlaced_640x448())
    rc, meta, img = o.vsync(force_progressive=1, high_resolution_scanout=1)
    w, h = 16, 8
    for i in range(2):
        for j in range(2):
            q = j + 2 * i
            acc = sum(channels(frame_from_vram(vv, w, h, plane=2 + 4 * q + k)) for k in range(4))
            check_rational(channels(img[j:2 * h:2, i:2 * w:2]), acc, 4)


def test_anti_blur_drops_the_offset_circuit():
    """Both circuits show one buffer with DBY one line apart in field mode: anti-blur keeps
    circuit 1 alone, replacing (the result equals circuit 1 without circuit 2)."""
    o, v = make(61)
    common = dict(SMODE1=crtc.smode1(), SMODE2=crtc.smode2(1, 0), DISPLAY1=crtc.display(), DISPLAY2=crtc.display())
    o.set_priv(crtc.priv_block(PMODE=crtc.pmode(en1=1, en2=1, mmod=1, alp=0x80), DISPFB1=crtc.dispfb(dby=0),
                               DISPFB2=crtc.dispfb(dby=1), **common))
    _, _, blurred = o.vsync(phase=0, skip_deinterlace=1)
    _, _, kept = o.vsync(phase=0, skip_deinterlace=1, anti_blur=1)
    o.set_priv(crtc.priv_block(PMODE=crtc.pmode(en1=1, en2=0), DISPFB1=crtc.dispfb(dby=0), **common))
    _, _, single = o.vsync(phase=0, skip_deinterlace=1)
    assert np.array_equal(kept, single) and not np.array_equal(blurred, single)


def test_vsync_can_skip_and_scanmsk_timeout():
    import gif
    import render_streams as rs
    o = Oracle(0)
    o.set_priv(crtc.ntsc_interlaced_640x448())
    assert not o.vsync_can_skip()                    # field mode is interlaced
    assert o.vsync_can_skip(force_progressive=1)     # unless forced progressive
    ofs = 2048 * 16
    regs = [(0x1A, 1), (0x18, ofs | ofs << 32), (0x4C, rs.frame(0, 10, 0x00)), (0x40, rs.scissor(0, 639, 0, 447)),
            (0x47, rs.test_reg()), (0x22, 2), (0x00, rs.prim_reg(6)), (0x01, rs.rgbaq(0x80FFFFFF)),
            (0x05, rs.xyz2(ofs, ofs, 0)), (0x05, rs.xyz2(ofs + 64 * 16, ofs + 440 * 16, 0))]
    o.gif_transfer(gif.arr(gif.ad(regs)))
    assert o.complex_scanmsk_timeout() == 4
    assert not o.vsync_can_skip(force_progressive=1)
    for left in (3, 2, 1, 0):
        o.vsync(force_progressive=1)
        assert o.complex_scanmsk_timeout() == left
    assert o.vsync_can_skip(force_progressive=1)


def test_field_aware_request_is_refused():
    o = Oracle(2)
    o.set_priv(crtc.ntsc_interlaced_640x448(ffmd=1))
    rc, _, img = o.vsync(force_progressive=1, high_resolution_scanout=1)
    assert rc == -1000 and img is None
    rc, _, img = o.vsync(force_progressive=1)
    assert rc == 0 and img is not None


def test_mode_table():
    """Mode sizes and pixel clocks: NTSC 640x224 fields / 448 frames at 13.5 / 27 MHz, PAL 640x256
    fields, VGA 640x480 at 27 MHz, HDTV 720p at 54 MHz; overscan widens to 712."""
    cases = [
        (crtc.smode1(crtc.CMOD_NTSC), crtc.smode2(1, 1), {}, (640, 224, 13.5, 1)),
        (crtc.smode1(crtc.CMOD_NTSC), crtc.smode2(1, 1), {"overscan": 1}, (712, 240, 13.5, 1)),
        (crtc.smode1(crtc.CMOD_PAL), crtc.smode2(1, 1), {}, (640, 256, 13.5, 1)),
        (crtc.smode1(crtc.CMOD_PROGRESSIVE), crtc.smode2(0, 0), {}, (640, 448, 27.0, 0)),
        (crtc.smode1(crtc.CMOD_PROGRESSIVE, crtc.LC_VGA), crtc.smode2(0, 0), {}, (640, 480, 27.0, 0)),
        (crtc.smode1(crtc.CMOD_PROGRESSIVE, crtc.LC_HDTV), crtc.smode2(0, 0), {}, (1280, 1440, 54.0, 0)),
        (crtc.smode1(crtc.CMOD_NTSC), crtc.smode2(0, 0), {}, (640, 224, 27.0, 0)),  # double strike 240p
    ]
    for s1, s2, info, (mw, mh, mhz, interlaced) in cases:
        o = Oracle(0)
        o.set_priv(crtc.priv_block(PMODE=crtc.pmode(en1=0), SMODE1=s1, SMODE2=s2))
        rc, meta, img = o.vsync(skip_deinterlace=1, **info)
        assert (meta["mode_width"], meta["mode_height"], meta["interlaced"]) == (mw, mh, interlaced), (s1, s2, info)
        assert abs(meta["sampling_rate_mhz"] - mhz) < 1e-4
