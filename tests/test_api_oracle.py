"""Oracle checks of front-end register semantics and host access (no GPU).

PRMODECONT / PRMODE (gs/gs_interface.cpp:3507-3556, 3686-3722), LOCAL->HOST readback
(gs/gs_interface.cpp:3213-3275, gs/gs_util.hpp:49-110) and host VRAM writes
(gs/gs_interface.cpp:4184-4228, gs/gs_renderer.cpp:1478-1492), each against a closed form.
"""
import numpy as np
import pytest

import gif
import render_streams as rs
from oracle_lib import Oracle, VRAM_SIZE

OFS = 2048 * 16


def ct32_pixel(vram, x, y, fbp, fbw):
    from oracle_lib import swizzle
    a = swizzle(x, y, fbp * 32, fbw, 0x00)
    return int(vram[a * 4:(a + 1) * 4].view(np.uint32)[0])


def sprite_stream(prmodecont, prmode, prim, color=0x80402010):
    """Two contexts with frame buffers at pages 0 and 64 (64 px wide), then one 8x8 sprite
    at (4, 4)."""
    regs = [(0x1A, prmodecont), (0x18, OFS | OFS << 32), (0x19, OFS | OFS << 32),
            (0x4C, rs.frame(0, 1, 0x00)), (0x4D, rs.frame(64, 1, 0x00)),
            (0x40, rs.scissor(0, 63, 0, 63)), (0x41, rs.scissor(0, 63, 0, 63)),
            (0x47, rs.test_reg(zte=1, ztst=1)), (0x48, rs.test_reg(zte=1, ztst=1)),
            (0x4E, rs.zbuf(32, 0, 1)), (0x4F, rs.zbuf(96, 0, 1))]
    if prmode is not None:
        regs.append((0x1B, prmode))
    regs += [(0x00, prim), (0x01, rs.rgbaq(color)), (0x05, rs.xyz2(OFS + 4 * 16, OFS + 4 * 16, 0)),
             (0x05, rs.xyz2(OFS + 12 * 16, OFS + 12 * 16, 0))]
    return gif.arr(gif.ad(regs))


def test_prmodecont_ac1_prim_selects_context():
    o = Oracle()
    o.gif_transfer(sprite_stream(1, None, rs.prim_reg(6, ctxt=1)))
    v = o.vram()
    assert ct32_pixel(v, 5, 5, 64, 1) == 0x80402010
    assert ct32_pixel(v, 5, 5, 0, 1) == 0


def test_prmodecont_ac0_takes_attributes_from_prmode():
    """With AC = 0 the PRIM write only sets the primitive type: the context bit comes from
    PRMODE, and a PRIM write's other bits are ignored."""
    o = Oracle()
    o.gif_transfer(sprite_stream(0, rs.prim_reg(3, ctxt=1), rs.prim_reg(6, ctxt=0)))
    v = o.vram()
    assert ct32_pixel(v, 5, 5, 64, 1) == 0x80402010
    assert ct32_pixel(v, 5, 5, 0, 1) == 0
    # The same with PRMODE selecting context 1 but AC = 1: PRMODE is ignored.
    o = Oracle()
    o.gif_transfer(sprite_stream(1, rs.prim_reg(3, ctxt=1), rs.prim_reg(6, ctxt=0)))
    v = o.vram()
    assert ct32_pixel(v, 5, 5, 0, 1) == 0x80402010
    assert ct32_pixel(v, 5, 5, 64, 1) == 0


def test_prmode_written_under_ac1_is_lost():
    """PRMODE is not latched while AC = 1 (the reference writes it into PRIM only when AC = 0)."""
    o = Oracle()
    regs = [(0x1A, 1), (0x1B, rs.prim_reg(0, ctxt=1)), (0x1A, 0)]
    o.gif_transfer(gif.arr(gif.ad(regs)))
    o.gif_transfer(sprite_stream(0, None, rs.prim_reg(6, ctxt=1)))
    v = o.vram()
    assert ct32_pixel(v, 5, 5, 0, 1) == 0x80402010


def test_reset_state_has_ac0():
    """RegisterState is zero after reset (gs/gs_interface.cpp:32-38), so AC = 0: PRIM's context
    bit is ignored until PRMODECONT is written."""
    o = Oracle()
    stream = sprite_stream(1, None, rs.prim_reg(6, ctxt=1))
    # Drop the PRMODECONT write: replace it with a write to an unused register.
    raw = bytearray(stream.tobytes())
    raw[16 + 8] = 0x7F
    o.gif_transfer(np.frombuffer(bytes(raw), np.uint8))
    v = o.vram()
    assert ct32_pixel(v, 5, 5, 0, 1) == 0x80402010


READABLE = [0x00, 0x01, 0x02, 0x0A, 0x13, 0x1B, 0x30, 0x31, 0x32, 0x3A]


@pytest.mark.parametrize("psm", READABLE)
def test_readback_round_trip(psm):
    """Upload w x h pixels of a format, read them back: the FIFO holds the payload (24-bit
    formats: 3 bytes per pixel), for 8H the byte written."""
    rng = np.random.default_rng(psm)
    w, h = 40, 24
    bpp = gif.BPP[psm]
    payload = rng.integers(0, 256, w * h * bpp // 8, dtype=np.uint8).tobytes()
    o = Oracle()
    o.gif_transfer(gif.arr(gif.upload(96, 2, psm, 8, 5, w, h, payload)))
    o.gif_transfer(gif.arr(gif.ad([(0x50, gif.bitbltbuf(sbp=96, sbw=2, spsm=psm)), (0x51, gif.trxpos(8, 5)),
                                   (0x52, gif.trxreg(w, h)), (0x53, 1)])))
    fifo = o.read_transfer_fifo(len(payload) // 16)
    assert fifo.tobytes() == payload
    # Reading past the end gives zeros.
    assert not o.read_transfer_fifo(4).any()


def test_readback_sees_pending_draw():
    o = Oracle()
    o.gif_transfer(sprite_stream(1, None, rs.prim_reg(6)))
    o.gif_transfer(gif.arr(gif.ad([(0x50, gif.bitbltbuf(sbp=0, sbw=1, spsm=0x00)), (0x51, gif.trxpos(0, 0)),
                                   (0x52, gif.trxreg(16, 16)), (0x53, 1)])))
    px = o.read_transfer_fifo(16 * 16 * 4 // 16).view(np.uint32).reshape(16, 16)
    assert (px[4:12, 4:12] == 0x80402010).all()
    assert px[:4].sum() == 0 and px[12:].sum() == 0


def test_readback_4bit_formats_read_zero():
    o = Oracle()
    o.gif_transfer(gif.arr(gif.upload(0, 1, 0x00, 0, 0, 64, 32, bytes([0xFF]) * (64 * 32 * 4))))
    o.gif_transfer(gif.arr(gif.ad([(0x50, gif.bitbltbuf(sbp=0, sbw=1, spsm=0x14)), (0x51, gif.trxpos(0, 0)),
                                   (0x52, gif.trxreg(32, 32)), (0x53, 1)])))
    assert not o.read_transfer_fifo(32).any()


def test_host_write_clears_reference_plane_pages():
    """A host write lands in plane 0 and clears the super-sample reference plane of every 8 KiB
    page it touches, nothing else."""
    o = Oracle(2)
    tri = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 1, 0x00)), (0x40, rs.scissor(0, 63, 0, 63)),
           (0x47, rs.test_reg(zte=1, ztst=1)), (0x4E, rs.zbuf(32, 0, 1)), (0x00, rs.prim_reg(3, iip=0)),
           (0x01, rs.rgbaq(0x80FFFFFF)), (0x05, rs.xyz2(OFS + 3, OFS + 5, 100)),
           (0x05, rs.xyz2(OFS + 30 * 16 + 7, OFS + 2 * 16, 5000)),
           (0x05, rs.xyz2(OFS + 9 * 16, OFS + 20 * 16 + 3, 90000))]
    o.gif_transfer(gif.arr(gif.ad(tri)))
    o.flush()
    before = o.planes().copy()
    ref = before[VRAM_SIZE:2 * VRAM_SIZE]
    assert ref.any(), "the triangle must leave reference bits"
    data = np.arange(100, dtype=np.uint8)
    o.host_write(data, 8192 * 0 + 300)
    after = o.planes()
    assert (after[300:400] == data).all()
    assert not after[VRAM_SIZE:VRAM_SIZE + 8192].any()
    assert (after[VRAM_SIZE + 8192:] == before[VRAM_SIZE + 8192:]).all()
    assert (after[:300] == before[:300]).all() and (after[400:VRAM_SIZE] == before[400:VRAM_SIZE]).all()


def test_rate_change_clears_sample_planes():
    o = Oracle(0)
    o.gif_transfer(gif.arr(sprite_stream(1, None, rs.prim_reg(6))))
    o.set_sample_log2(2)
    p = o.planes()
    assert p.size == 6 * VRAM_SIZE
    assert not p[VRAM_SIZE:].any()
    assert ct32_pixel(p[:VRAM_SIZE], 5, 5, 0, 1) == 0x80402010
