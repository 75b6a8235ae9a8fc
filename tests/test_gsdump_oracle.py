"""GS dump format, oracle replay and the replayer's image helpers (no GPU).

The format follows dump/gs_dump_parser.cpp:6-207 and dump/gs_dump_generator.cpp:39-237; the
replayer's PNG and PSNR follow tools/gs_repro_replayer.cpp:29-63 (stb_image PNGs, 40 dB per
channel).
"""
import os
import struct
import zlib

import numpy as np
import pytest

import crtc
import gif
import gsdump
import pgs
from oracle_lib import Oracle
from pgs import replay


def textured_dump(seed, n_prims=200, samples_priv=None, frames=2):
    """A dump from the reset state: random VRAM, then `frames` textured frames, each followed by
    the privileged registers and a vsync."""
    rng = np.random.default_rng(seed)
    w = gsdump.Writer(vram=rng.integers(0, 256, gsdump.VRAM_SIZE, dtype=np.uint8))
    priv = samples_priv or crtc.ntsc_interlaced_640x448()
    for f in range(frames):
        w.transfer(pgs.gen_textured(seed + f, n_prims, 1, 640, 448, 0.0))
        w.vsync(f & 1, priv)
    return w


def test_header_and_state_layout():
    w = gsdump.Writer(regs={"prim": 6, "frame1": 0x1234, "uv": 0x00050007, "fog": 0xAB << 56 | 0x11},
                      paths=[(0, 0, 0), (5 | 1 << 15, 0xE, 1), (0, 0, 0), (0, 0, 0)], internal_q=0.5)
    d = gsdump.parse(bytes(w.data))
    assert d["header"][0] == 9 and d["header"][1] == gsdump.STATE_SIZE == 4194304 + 336 + 28 + 80 + 4
    assert d["version"] == 8
    assert d["regs"]["prim"] == 6 and d["regs"]["frame1"] == 0x1234 and d["regs"]["uv"] == 0x00050007
    assert d["regs"]["fog"] == 0x11  # only FOG's low word travels (its value sits in the high byte)
    assert d["paths"][1] == (5 | 1 << 15, 0xE, 1) and d["internal_q"] == 0.5
    # Start state: 8 + 36 + 4 + 15*8 + 24*8 + 8 + 8 + 4 + 4 + 8 + 16 bytes of registers, VRAM,
    # 4 x 20 bytes of paths, Q, priv.
    assert len(w.data) == 8 + 36 + 4 + 120 + 192 + 24 + 24 + gsdump.VRAM_SIZE + 80 + 4 + 8192


def test_version9_and_header_extension_parse_alike():
    a = gsdump.Writer(regs={"prim": 3}, state_version=8)
    b = gsdump.Writer(regs={"prim": 3}, state_version=9, header_extra=b"SERIAL-0001")
    for w in (a, b):
        w.vsync(1)
    da, db = gsdump.parse(bytes(a.data)), gsdump.parse(bytes(b.data))
    assert da["regs"] == db["regs"] and da["packets"] == db["packets"] == [("vsync", 1)]


def test_oracle_replay_equals_the_direct_run():
    """Replaying a dump drives the oracle exactly as feeding its packets directly."""
    w = textured_dump(0x5EED00D1, frames=2)
    rc, meta, img = gsdump.replay_into_oracle(Oracle(0), bytes(w.data))
    d = gsdump.parse(bytes(w.data))
    o = Oracle(0)
    o.host_write(np.array(d["vram"]), 0)
    for k in range(2):
        o.gif_transfer(pgs.gen_textured(0x5EED00D1 + k, 200, 1, 640, 448, 0.0))
        o.set_priv(crtc.ntsc_interlaced_640x448())
        rc2, meta2, img2 = o.vsync(phase=k & 1, adapt_to_internal_horizontal_resolution=1, force_progressive=1,
                                   anti_blur=1)
    assert rc == rc2 == 0 and meta == meta2 and np.array_equal(img, img2)


def test_read_fifo_stops_the_replay():
    w = gsdump.Writer()
    w.transfer(gif.arr(gif.ad([(0x3F, 0)])))
    w.read_fifo(64)
    w.vsync(0, crtc.ntsc_interlaced_640x448())
    rc, meta, img = gsdump.replay_into_oracle(Oracle(0), bytes(w.data))
    assert img is None


def _png_with_filters(img, filters):
    """Encodes (h, w, 4) bytes with the given filter type per row (PNG filters 0-4)."""
    h, w, c = img.shape
    stride = w * c
    rows = []
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        cur = img[y].reshape(-1).astype(np.int32)
        ft = filters[y % len(filters)]
        a = np.concatenate([np.zeros(c, np.int32), cur[:-c]])
        cc = np.concatenate([np.zeros(c, np.int32), prev[:-c]])
        if ft == 0:
            f = cur
        elif ft == 1:
            f = cur - a
        elif ft == 2:
            f = cur - prev
        elif ft == 3:
            f = cur - ((a + prev) >> 1)
        else:
            pa, pb, pc = np.abs(prev - cc), np.abs(a - cc), np.abs(a + prev - 2 * cc)
            p = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, cc))
            f = cur - p
        rows.append(bytes([ft]) + (f & 255).astype(np.uint8).tobytes())
        prev = cur
    return b"\x89PNG\r\n\x1a\n" + replay._chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) + \
        replay._chunk(b"IDAT", zlib.compress(b"".join(rows))) + replay._chunk(b"IEND", b"")


def test_png_round_trip_and_every_filter(tmp_path):
    img = np.random.default_rng(3).integers(0, 256, (13, 17, 4), dtype=np.uint8)
    p = str(tmp_path / "a.png")
    replay.write_png(p, img)
    assert np.array_equal(replay.read_png(p), img)
    q = str(tmp_path / "b.png")
    with open(q, "wb") as f:
        f.write(_png_with_filters(img, [0, 1, 2, 3, 4]))
    assert np.array_equal(replay.read_png(q), img)


def test_psnr():
    a = np.zeros((10, 10, 4), np.uint8)
    assert replay.psnr(a, a) == [float("inf")] * 3
    b = a.copy()
    b[0, 0, 0] = 255  # one full-scale error in 100 pixels: 10 log10(100) = 20 dB
    assert abs(replay.psnr(b, a)[0] - 20.0) < 1e-9
