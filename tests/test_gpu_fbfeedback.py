"""GPU parity of on-tile frame-buffer feedback textures (gs/gs_interface.cpp:2141-2470,
ubershader.comp:955-1040): random feedback streams (colour, 24-bit, palette and depth views,
pixel-aligned and not, another layout read as the pass starts) byte for byte against the oracle at
1x and super-sampled, and the closed-form feedback tests of test_fbfeedback_oracle.py on the GPU."""
import numpy as np
import pytest

import pgs
import render_streams as rs
import test_fbfeedback_oracle as T
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


@pytest.mark.parametrize("seed,ssaa", [(0, 1), (1, 1), (2, 1), (3, 4), (4, 4), (5, 16), (6, 2), (7, 8)])
def test_random_feedback_streams(hip_device, seed, ssaa):
    stream = rs.random_fb_feedback_stream(900 + seed)
    print(f"[case] feedback stream seed={900 + seed} ssaa={ssaa}", flush=True)
    with pgs.Context(super_sampling=ssaa) as ctx:
        ctx.gif_transfer(stream)
        got = ctx.vram_planes()
    o = Oracle(LOG2[ssaa])
    o.gif_transfer(stream)
    o.flush()
    ref = o.planes()
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


def test_closed_form_on_gpu(hip_device, monkeypatch):
    """Every closed-form test of the oracle module, its streams also run on the GPU with the same
    sample count: all planes must equal the oracle's before its assertions run."""

    class Mirrored(Oracle):
        def __init__(self, sample_log2=0, sparse=False):
            super().__init__(sample_log2, sparse)
            self._log2 = sample_log2
            self._streams = []

        def gif_transfer(self, data, path=3):
            self._streams.append(np.array(data, np.uint8))
            super().gif_transfer(data, path)

        def flush(self):
            # Pass boundaries decide the sampling rate of a pass: the GPU flushes where the
            # oracle does.
            self._streams.append(None)
            super().flush()

        def planes(self):
            ref = super().planes()
            with pgs.Context(super_sampling=1 << self._log2) as ctx:
                for s in self._streams:
                    if s is None:
                        ctx.flush()
                    else:
                        ctx.gif_transfer(s)
                got = ctx.vram_planes()
            bad = np.flatnonzero(got != ref)
            assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"
            return ref

    monkeypatch.setattr(T, "Oracle", Mirrored)
    for name in dir(T):
        if name.startswith("test_"):
            print(f"[case] {name}", flush=True)
            getattr(T, name)()


def test_feedback_batch_replay(hip_device):
    ctxs, live = [], []
    for i, ssaa in enumerate((1, 4, 16)):
        stream = rs.random_fb_feedback_stream(950 + i)
        print(f"[case] batch context {i} ssaa={ssaa}", flush=True)
        ctx = pgs.Context(super_sampling=ssaa, arena_bytes=32 << 20)
        ctx.begin_recording()
        ctx.gif_transfer(stream)
        ctx.end_recording()
        live.append(ctx.vram_planes())
        o = Oracle(LOG2[ssaa])
        o.gif_transfer(stream)
        o.flush()
        assert np.array_equal(live[-1], o.planes())
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
