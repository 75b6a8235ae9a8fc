"""Exhaustive checks of the float identities the shading kernels use in place of the reference's
float sequences (ubershader.comp shade_resolve and the unorm unpack of colours), evaluated in
IEEE single precision with numpy.

* frame buffer and fog colour: rint((b / 255) * 255) == b and (b / 255) * 255 == b for every
  byte b, so the kernels use the byte itself;
* the colour write: for integer-valued o in -4096..4096,
  rint(clamp(fract(o / 256) * (256 / 255), 0, 1) * 255) == o & 255 (wrap) and
  rint(clamp(o * (1 / 255), 0, 1) * 255) == clamp(o, 0, 255) (COLCLAMP);
* division by a power of two equals multiplication by its reciprocal;
* the texel wrap done in float (floor of v times the reciprocal size, two corrections) is the
  mathematical remainder for every image size up to 1024 and |v| < 2^22;
* b / 255 is not b * fl(1 / 255) for every byte, which is why the kernels keep an exact table;
* the bilinear filter's lerp of two bytes held 16 bits apart in one word equals the per-byte
  lerp of the oracle for every pair of bytes and every weight 0..256.
"""
import numpy as np

F = np.float32


def unorm8(b):
    return np.divide(b.astype(F), F(255))


def test_frame_buffer_and_fog_bytes_round_trip():
    b = np.arange(256, dtype=F)
    assert np.array_equal(np.rint((unorm8(b) * F(255)).astype(F)), b)
    assert np.array_equal((unorm8(b) * F(255)).astype(F), b)


def test_colour_write_wrap_and_clamp_are_integer_ops():
    o = np.arange(-4096, 4097).astype(F)
    q = (o / F(256)).astype(F)
    n = ((q - np.floor(q)).astype(F) * (F(256) / F(255))).astype(F)
    wrap = np.rint((np.clip(n, F(0), F(1)) * F(255)).astype(F)).astype(np.int64)
    assert np.array_equal(wrap, o.astype(np.int64) & 255)
    n2 = (o * (F(1) / F(255))).astype(F)
    clamp = np.rint((np.clip(n2, F(0), F(1)) * F(255)).astype(F)).astype(np.int64)
    assert np.array_equal(clamp, np.clip(o.astype(np.int64), 0, 255))


def test_power_of_two_division_is_reciprocal_multiplication():
    rng = np.random.default_rng(5)
    x = np.concatenate([np.arange(-70000, 70000).astype(F), rng.standard_normal(100000).astype(F) * F(300)])
    for s in (2, 4, 8, 16, 256):
        assert np.array_equal((x / F(s)).astype(F), (x * F(1.0 / s)).astype(F))


def test_reciprocal_multiplication_is_not_exact_for_255():
    b = np.arange(256, dtype=F)
    assert not np.array_equal(unorm8(b), (b * (F(1) / F(255))).astype(F))


def wrap_texel(v, size):
    """The shader's texel wrap in float32 steps (kernels_shade.hip wrap_texel)."""
    vf = v.astype(F)
    inv = F(1) / F(size)
    q = np.floor((vf * inv).astype(F)).astype(F)
    r = (vf - (q * F(size)).astype(F)).astype(F)
    r = np.where(r < 0, (r + F(size)).astype(F), r)
    r = np.where(r >= F(size), (r - F(size)).astype(F), r)
    return r.astype(np.int64)


def test_float_texel_wrap_is_the_remainder():
    """For every image size 1..1024 and coordinates below 2^22 in magnitude (multiples of the
    size and their neighbours, where the float quotient can round across an integer, the
    extremes, and random values), the float wrap equals the mathematical remainder."""
    rng = np.random.default_rng(9)
    lim = (1 << 22) - 1
    for size in range(1, 1025):
        k = np.arange(-(lim // size), lim // size + 1, max(1, (lim // size) // 512), dtype=np.int64)
        mult = np.concatenate([k * size - 1, k * size, k * size + 1])
        v = np.concatenate([mult, rng.integers(-lim, lim, 2048), [-lim, lim, 0, -1, 1]])
        v = v[np.abs(v) <= lim]
        assert np.array_equal(wrap_texel(v, size), np.mod(v, size)), size


def lerp8_ref(a, b, w):
    """Per-byte lerp of the oracle's bilinear filter: a + floor(((b - a) w + 128) / 256)."""
    return a + np.floor_divide((b - a) * w + 128, 256)


def lerp8x4(a, b, w):
    """The shader's packed form (kernels_shade.hip lerp8x4): two bytes per 16-bit field."""
    M = np.uint64(0x00FF00FF)
    a, b, w = a.astype(np.uint64), b.astype(np.uint64), w.astype(np.uint64)
    iw = np.uint64(256) - w
    even = ((a & M) * iw + (b & M) * w + np.uint64(0x00800080)) & np.uint64(0xFFFFFFFF)
    odd = (((a >> np.uint64(8)) & M) * iw + ((b >> np.uint64(8)) & M) * w + np.uint64(0x00800080)) & np.uint64(0xFFFFFFFF)
    return ((even >> np.uint64(8)) & M) | (odd & ~M & np.uint64(0xFFFFFFFF))


def test_packed_bilinear_lerp_matches_per_byte_lerp():
    """Exhaustive over bytes a, b and weights 0..256: a (256 - w) + b w + 128 stays below 2^16
    and its bits 8..15 are the lerp; then the packed four-byte form on random words and all
    weights equals the lerp of each byte."""
    a, b = np.meshgrid(np.arange(256, dtype=np.int64), np.arange(256, dtype=np.int64), indexing="ij")
    for w in range(257):
        s = a * (256 - w) + b * w + 128
        assert s.max() < 1 << 16
        assert np.array_equal(s >> 8, lerp8_ref(a, b, w)), w
    rng = np.random.default_rng(3)
    wa = rng.integers(0, 2**32, 4096, dtype=np.uint64)
    wb = rng.integers(0, 2**32, 4096, dtype=np.uint64)
    for w in range(257):
        ww = np.full(wa.shape, w, dtype=np.uint64)
        got = lerp8x4(wa, wb, ww)
        want = np.zeros_like(wa)
        for ch in range(4):
            sh = np.uint64(8 * ch)
            x = ((wa >> sh) & np.uint64(255)).astype(np.int64)
            y = ((wb >> sh) & np.uint64(255)).astype(np.int64)
            want |= lerp8_ref(x, y, w).astype(np.uint64) << sh
        assert np.array_equal(got, want), w


def test_edge_products_fit_24_bits():
    """The shaders multiply edge slopes by sample coordinates with the 24-bit multiply: slopes
    are differences of vertex coordinates (16-bit register fields minus a 16-bit offset, lines
    moved by at most 17), sample coordinates are pixel positions below 2^11 times 8."""
    coord_max = 0xFFFF + 17
    slope_max = 2 * coord_max
    assert slope_max < 1 << 23
    assert (2047 << 3) | 7 < 1 << 23
