"""GPU parity: VRAM transfers through the C ABI against the CPU oracle, bit for bit."""
import numpy as np
import pytest

import gif
import pgs
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

SEED_TRANSFERS = 0x5EED0002  # the transfer microbenchmark workload


def run_both(stream: np.ndarray, ssaa: int = 1, chunks=None):
    log2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}[ssaa]
    print(f"[case] stream={len(stream)} bytes ssaa={ssaa}", flush=True)
    with pgs.Context(super_sampling=ssaa) as ctx:
        o = Oracle(log2)
        if chunks:
            for a, b in chunks:
                ctx.gif_transfer(stream[a:b])
                o.gif_transfer(stream[a:b])
                ctx.flush()
                o.flush()
        else:
            ctx.gif_transfer(stream)
            o.gif_transfer(stream)
        planes = ctx.vram_planes()
        stats = ctx.consume_flush_stats()
    o.flush()
    ref = o.planes()
    return planes, ref, stats


def assert_vram_equal(got, ref):
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


def test_transfer_small(hip_device):
    stream = pgs.gen_transfers(SEED_TRANSFERS, 48, 12, 128)
    got, ref, stats = run_both(stream)
    assert stats["copies"] == 60
    assert_vram_equal(got, ref)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_transfer_edge_cases(hip_device, seed):
    """Unaligned origins, wrap at 2048, narrow strides, mixed PSM local copies, short payloads."""
    stream = pgs.gen_transfers_wild(seed, 300)
    got, ref, _ = run_both(stream)
    assert_vram_equal(got, ref)


def test_transfer_partial_flushes(hip_device):
    rng = np.random.default_rng(9)
    w, h = 200, 120
    payload = rng.integers(0, 256, w * h * 3, dtype=np.uint8).tobytes()
    stream = gif.arr(gif.upload(100, 4, 0x01, 30, 17, w, h, payload) +
                     gif.upload(120, 4, 0x14, 9, 3, 64, 64, rng.integers(0, 256, 2048, dtype=np.uint8).tobytes()))
    n = len(stream)
    cuts = [0, 16 * 5 + 32 * 16, 16 * 5 + 777 * 16, n]
    got, ref, _ = run_both(stream, chunks=list(zip(cuts[:-1], cuts[1:])))
    assert_vram_equal(got, ref)


def test_transfer_ssaa_context(hip_device):
    """At 4x SSAA all 6 planes exist; uploads write plane 0 and clear reference bits."""
    stream = pgs.gen_transfers(SEED_TRANSFERS + 100, 20, 5, 96)
    got, ref, _ = run_both(stream, ssaa=4)
    assert got.size == 6 * 4 * 1024 * 1024
    assert_vram_equal(got, ref)


def test_transfer_microbenchmark_full_size(hip_device):
    """The full 4,000 + 1,000 transfer workload, whole 4 MiB VRAM bit-exact."""
    stream = pgs.gen_transfers(SEED_TRANSFERS, 4000, 1000, 512)
    got, ref, stats = run_both(stream)
    print(stats, flush=True)
    assert stats["copies"] == 5000
    assert_vram_equal(got, ref)


def test_readback_fifo_matches_vram(hip_device):
    rng = np.random.default_rng(4)
    w, h = 32, 16
    payload = rng.integers(0, 256, w * h * 4, dtype=np.uint8).tobytes()
    with pgs.Context() as ctx:
        ctx.gif_transfer(gif.arr(gif.upload(64, 1, 0x00, 0, 0, w, h, payload)))
        ctx.gif_transfer(gif.arr(gif.ad([(0x50, gif.bitbltbuf(sbp=64, sbw=1, spsm=0x00)),
                                         (0x51, gif.trxpos()), (0x52, gif.trxreg(w, h)), (0x53, 1)])))
        fifo = ctx.read_transfer_fifo(w * h * 4 // 16)
    assert fifo.tobytes() == payload
