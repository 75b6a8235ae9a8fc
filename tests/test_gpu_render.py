"""GPU parity: rendering through the C ABI against the CPU oracle, bit for bit."""
import numpy as np
import pytest

import pgs
import render_streams
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

SEED_PLAIN = 0x5EED0001  # the plain 1x rendering workload


def render_both(stream, ssaa=1, tile_log2=None):
    log2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}[ssaa]
    print(f"[case] stream={len(stream)} bytes ssaa={ssaa}", flush=True)
    with pgs.Context(super_sampling=ssaa) as ctx:
        ctx.gif_transfer(stream)
        got = ctx.vram_planes()
        stats = ctx.consume_flush_stats()
    o = Oracle(log2)
    o.gif_transfer(stream)
    o.flush()
    return got, o.planes(), stats, o


def assert_vram_equal(got, ref):
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


def test_plain_small(hip_device):
    stream = pgs.gen_plain(SEED_PLAIN, 64, 128, 96)
    got, ref, stats, _ = render_both(stream)
    assert stats["primitives"] > 0
    assert_vram_equal(got, ref)


def test_plain_full_workload(hip_device):
    """2,000 Gouraud triangles at 640x448, PSMCT32 + PSMZ24, GEQUAL: whole VRAM bit-exact."""
    stream = pgs.gen_plain(SEED_PLAIN, 2000, 640, 448)
    got, ref, stats, o = render_both(stream)
    print(stats, "covered samples", o.stat(0), flush=True)
    assert stats["render_passes"] == 1 and stats["primitives"] == 2000
    assert_vram_equal(got, ref)


@pytest.mark.parametrize("seed", list(range(8)))
def test_random_state(hip_device, seed):
    stream = render_streams.random_render_stream(seed)
    got, ref, _, _ = render_both(stream)
    assert_vram_equal(got, ref)


SEED_TEXTURED = 0x5EED0003  # the textured + blended 4x workload


def render_planes_both(stream, ssaa):
    log2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}[ssaa]
    print(f"[case] stream={len(stream)} bytes ssaa={ssaa}", flush=True)
    with pgs.Context(super_sampling=ssaa) as ctx:
        ctx.gif_transfer(stream)
        got = ctx.vram_planes()
        stats = ctx.consume_flush_stats()
    o = Oracle(log2)
    o.gif_transfer(stream)
    o.flush()
    return got, o.planes(), stats, o


@pytest.mark.parametrize("ssaa", [1, 4])
def test_textured_small(hip_device, ssaa):
    stream = pgs.gen_textured(SEED_TEXTURED, 2000, 1, 640, 448, 0.0)
    got, ref, stats, _ = render_planes_both(stream, ssaa)
    assert stats["render_passes"] >= 1
    assert_vram_equal(got, ref)


def test_textured_full_workload(hip_device):
    """8 frames x 10,000 textured, blended primitives at 4x: all six planes bit-exact."""
    stream = pgs.gen_textured(SEED_TEXTURED, 10000, 8, 640, 448, 0.0)
    got, ref, stats, o = render_planes_both(stream, 4)
    print(stats, "covered samples", o.stat(0), flush=True)
    assert stats["render_passes"] == 8
    assert_vram_equal(got, ref)


@pytest.mark.parametrize("ssaa", [16, 2])
def test_textured_pareto(hip_device, ssaa):
    """Heavy-tailed primitive areas (Pareto alpha 1.2) at 16x and 2x."""
    stream = pgs.gen_textured(SEED_TEXTURED + 2, 3000, 2, 640, 448, 1.2)
    got, ref, _, _ = render_planes_both(stream, ssaa)
    assert_vram_equal(got, ref)
