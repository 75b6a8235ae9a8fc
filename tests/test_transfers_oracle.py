"""CPU oracle: transfer semantics pinned by closed forms (no GPU).

Each check constructs the expected VRAM contents independently (numpy pixel gathers through
the separately pinned swizzle) and compares with what the oracle wrote.
"""
import numpy as np
import pytest

import gif
from oracle_lib import Oracle, swizzle as oswz

PSMS = [0x00, 0x01, 0x02, 0x0A, 0x13, 0x14, 0x1B, 0x24, 0x2C, 0x30, 0x31, 0x32, 0x3A]


def read_pixel(vram: np.ndarray, psm: int, x: int, y: int, base: int, stride: int) -> int:
    a = oswz(x, y, base, stride, psm)
    w32 = vram.view(np.uint32)
    if psm in (0x02, 0x0A, 0x32, 0x3A):
        return int(vram.view(np.uint16)[a])
    if psm == 0x13:
        return int(vram[a])
    if psm == 0x14:
        return int((vram[a >> 1] >> ((a & 1) * 4)) & 0xF)
    if psm == 0x1B:
        return int(vram[a * 4 + 3])
    if psm == 0x24:
        return int((w32[a] >> 24) & 0xF)
    if psm == 0x2C:
        return int(w32[a] >> 28)
    if psm in (0x01, 0x31):
        return int(w32[a] & 0xFFFFFF)
    return int(w32[a])


@pytest.mark.parametrize("psm", PSMS)
def test_upload_round_trip(psm):
    rng = np.random.default_rng(1000 + psm)
    w, h, dsax, dsay, dbw, dbp = 40, 24, 16, 8, 2, 96
    nbytes = w * h * gif.BPP[psm] // 8
    payload = rng.integers(0, 256, nbytes, dtype=np.uint8).tobytes()
    o = Oracle()
    o.gif_transfer(gif.arr(gif.upload(dbp, dbw, psm, dsax, dsay, w, h, payload)))
    vram = o.vram()
    expect = gif.unpack_pixels(payload, gif.BPP[psm], w * h)
    got = np.array([read_pixel(vram, psm, (dsax + i % w) & 2047, (dsay + i // w) & 2047, dbp, dbw)
                    for i in range(w * h)], dtype=np.uint64)
    assert np.array_equal(got, expect)


@pytest.mark.parametrize("psm,keep", [(0x01, 0xFF000000), (0x31, 0xFF000000), (0x1B, 0x00FFFFFF),
                                      (0x24, 0xF0FFFFFF), (0x2C, 0x0FFFFFFF)])
def test_upload_preserves_other_lanes(psm, keep):
    """24-bit, 8H and 4HL/HH uploads leave the other bits of each 32-bit word untouched."""
    rng = np.random.default_rng(psm)
    o = Oracle()
    fill = rng.integers(0, 2**32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    o.set_vram(fill.view(np.uint8))
    w, h = 64, 32
    payload = rng.integers(0, 256, w * h * gif.BPP[psm] // 8, dtype=np.uint8).tobytes()
    o.gif_transfer(gif.arr(gif.upload(0, 1, psm, 0, 0, w, h, payload)))
    after = o.vram().view(np.uint32)
    # one page touched; everything else identical, and the kept lanes identical everywhere
    assert np.array_equal(after & keep, fill & keep)
    assert np.array_equal(after[2048:], fill[2048:])


def test_local_copy_overlapping_reads_whole_source_first():
    """A self-overlapping local copy behaves as read-entire-source, then write."""
    rng = np.random.default_rng(7)
    o = Oracle()
    w, h = 96, 64
    payload = rng.integers(0, 256, w * h * 4, dtype=np.uint8).tobytes()
    o.gif_transfer(gif.arr(gif.upload(0, 2, 0x00, 0, 0, w, h, payload)))
    before = np.array([[read_pixel(o.vram(), 0x00, x, y, 0, 2) for x in range(w)] for y in range(h)], dtype=np.uint64)
    cw, ch, sx, sy, dx, dy = 64, 40, 4, 6, 12, 3
    o.gif_transfer(gif.arr(gif.local_copy(0, 2, 0x00, sx, sy, 0, 2, 0x00, dx, dy, cw, ch)))
    vram = o.vram()
    after = np.array([[read_pixel(vram, 0x00, x, y, 0, 2) for x in range(w)] for y in range(h)], dtype=np.uint64)
    model = before.copy()
    model[dy:dy + ch, dx:dx + cw] = before[sy:sy + ch, sx:sx + cw]
    assert np.array_equal(after, model)


def test_local_copy_format_conversion_rules():
    """PSMT4HH source -> PSMT8 destination moves the 4-bit index into the low nibble."""
    rng = np.random.default_rng(11)
    o = Oracle()
    words = rng.integers(0, 2**32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    o.set_vram(words.view(np.uint8))
    w, h = 16, 16
    o.gif_transfer(gif.arr(gif.local_copy(0, 1, 0x2C, 0, 0, 256, 2, 0x13, 0, 0, w, h)))
    vram = o.vram()
    for y in range(h):
        for x in range(w):
            src = int(words[oswz(x, y, 0, 1, 0x2C)] >> 28)
            assert read_pixel(vram, 0x13, x, y, 256, 2) == src


def test_partial_transfer_equals_whole():
    """An upload delivered in two GIF calls with a flush in between equals one delivery."""
    rng = np.random.default_rng(3)
    w, h = 64, 64
    payload = rng.integers(0, 256, w * h * 2, dtype=np.uint8).tobytes()
    stream = gif.upload(32, 1, 0x02, 8, 8, w, h, payload)
    whole = Oracle()
    whole.gif_transfer(gif.arr(stream))
    part = Oracle()
    cut = 16 * 4 + 16 + 16 * 37  # registers, IMAGE tag, 37 qwords of data
    part.gif_transfer(gif.arr(stream[:cut]))
    part.flush()
    part.gif_transfer(gif.arr(stream[cut:]))
    assert np.array_equal(whole.vram(), part.vram())


def test_later_upload_wins():
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, 32 * 32 * 4, dtype=np.uint8).tobytes()
    b = rng.integers(0, 256, 16 * 16 * 4, dtype=np.uint8).tobytes()
    o = Oracle()
    o.gif_transfer(gif.arr(gif.upload(0, 1, 0x00, 0, 0, 32, 32, a) + gif.upload(0, 1, 0x00, 8, 8, 16, 16, b)))
    vram = o.vram()
    pb = gif.unpack_pixels(b, 32, 256)
    pa = gif.unpack_pixels(a, 32, 1024)
    for y in range(32):
        for x in range(32):
            v = read_pixel(vram, 0x00, x, y, 0, 1)
            if 8 <= x < 24 and 8 <= y < 24:
                assert v == pb[(y - 8) * 16 + (x - 8)]
            else:
                assert v == pa[y * 32 + x]
