"""GPU parity of passes whose depth buffer is their colour buffer (gs/gs_renderer.cpp:2184-2232,
2418-2690): random aliasing streams (every colour / depth format pair, primitives switching
between no depth, test-only and depth writes, runs of test-only sprites) byte for byte against
the oracle at 1x and super-sampled, in a fused single-context replay and a batch replay, and the
closed-form aliasing tests of test_fbz_alias_oracle.py on the GPU."""
import numpy as np
import pytest

import pgs
import render_streams as rs
import test_fbz_alias_oracle as T
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


def oracle_planes(stream, ssaa):
    o = Oracle(LOG2[ssaa])
    o.gif_transfer(stream)
    o.flush()
    return o.planes()


@pytest.mark.parametrize("seed,ssaa", [(0, 1), (1, 1), (2, 1), (3, 1), (4, 2), (5, 4), (6, 8), (7, 16), (8, 4)])
def test_random_alias_streams(hip_device, seed, ssaa):
    stream = rs.random_fbz_alias_stream(1100 + seed)
    print(f"[case] alias stream seed={1100 + seed} ssaa={ssaa}", flush=True)
    with pgs.Context(super_sampling=ssaa) as ctx:
        ctx.gif_transfer(stream)
        got = ctx.vram_planes()
    ref = oracle_planes(stream, ssaa)
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


def test_closed_form_on_gpu(hip_device, monkeypatch):
    """Every closed-form test of the oracle module, its streams also run on the GPU at the same
    sample count: VRAM must equal the oracle's before its assertions run."""

    class Mirrored(Oracle):
        def __init__(self, sample_log2=0, sparse=False):
            super().__init__(sample_log2, sparse)
            self._log2 = sample_log2
            self._streams = []

        def gif_transfer(self, data, path=3):
            self._streams.append(np.array(data, np.uint8))
            super().gif_transfer(data, path)

        def planes(self):
            ref = super().planes()
            with pgs.Context(super_sampling=1 << self._log2) as ctx:
                for s in self._streams:
                    ctx.gif_transfer(s)
                ctx.flush()
                got = ctx.vram_planes()
            bad = np.flatnonzero(got != ref)
            assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"
            return ref

    monkeypatch.setattr(T, "Oracle", Mirrored)
    for name in dir(T):
        if name.startswith("test_") and callable(getattr(T, name)):
            fn = getattr(T, name)
            params = getattr(fn, "pytestmark", [])
            if params:
                for mark in params:
                    if mark.name == "parametrize":
                        for value in mark.args[1]:
                            print(f"[case] {name}[{value}]", flush=True)
                            fn(value)
            else:
                print(f"[case] {name}", flush=True)
                fn()


@pytest.mark.parametrize("ssaa", [1, 4])
def test_alias_context_replay(hip_device, ssaa):
    """A recording replays (fused launches, with and without a graph) to the live result."""
    stream = rs.random_fbz_alias_stream(1150 + ssaa)
    print(f"[case] replay ssaa={ssaa}", flush=True)
    with pgs.Context(super_sampling=ssaa, arena_bytes=32 << 20) as ctx:
        ctx.begin_recording()
        ctx.gif_transfer(stream)
        ctx.end_recording()
        live = ctx.vram_planes()
        assert np.array_equal(live, oracle_planes(stream, ssaa))
        for use_graph in (False, True, True):
            ctx.replay(use_graph)
            ctx.sync()
            assert np.array_equal(ctx.vram_planes(), live), use_graph


def test_alias_batch_replay(hip_device):
    ctxs, live = [], []
    for i, ssaa in enumerate((1, 4, 1, 16)):
        stream = rs.random_fbz_alias_stream(1160 + i)
        print(f"[case] batch context {i} ssaa={ssaa}", flush=True)
        ctx = pgs.Context(super_sampling=ssaa, arena_bytes=32 << 20)
        ctx.begin_recording()
        ctx.gif_transfer(stream)
        ctx.end_recording()
        live.append(ctx.vram_planes())
        assert np.array_equal(live[-1], oracle_planes(stream, ssaa))
        ctxs.append(ctx)
    batch = pgs.Batch(ctxs)
    batch.replay(True)
    batch.replay(True)
    batch.sync()
    for i, ctx in enumerate(ctxs):
        assert np.array_equal(ctx.vram_planes(), live[i]), i
    batch.close()
    for c in ctxs:
        c.close()
