"""GPU parity of super-sampled textures (gs/gs_interface.cpp:1593-1603, 1783-1831,
gs/gs_renderer.cpp:94-313 and 1848-1885, triangle_setup.comp:677-789, binning.comp:123-124,
upload.comp:133-228, ubershader.comp:1155-1211): the product's phase filter tables equal the
oracle's, random streams of rendered frames sampled by sprites at every sampling grid are byte
for byte equal to the oracle on every plane, the closed-form tests of
test_ssaa_texture_oracle.py hold on the GPU, and batch replays reproduce the live run."""
import numpy as np
import pytest

import pgs
import render_streams as rs
import test_ssaa_texture_oracle as T
from oracle_lib import Oracle, oracle_phase_lut

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


@pytest.mark.parametrize("rx,ry", T.GRIDS)
def test_phase_lut_equals_oracle(hip_device, rx, ry):
    assert np.array_equal(pgs.phase_lut(rx, ry), oracle_phase_lut(rx, ry))


@pytest.mark.parametrize("seed,ssaa,ordered", [(0, 2, True), (1, 4, True), (2, 4, False), (3, 8, True),
                                               (4, 16, True), (5, 16, False), (6, 4, True), (7, 1, True)])
def test_random_ssaa_texture_streams(hip_device, seed, ssaa, ordered):
    stream = rs.random_ssaa_texture_stream(800 + seed)
    print(f"[case] ssaa texture stream seed={800 + seed} ssaa={ssaa} ordered={ordered}", flush=True)
    with pgs.Context(super_sampling=ssaa, ordered=ordered, super_sampled_textures=True) as ctx:
        ctx.gif_transfer(stream)
        got = ctx.vram_planes()
    o = Oracle(LOG2[ssaa], sparse=not ordered, super_sampled_textures=True)
    o.gif_transfer(stream)
    o.flush()
    ref = o.planes()
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


def test_closed_form_on_gpu(hip_device, monkeypatch):
    """The oracle module's closed-form tests, every stream also run on the GPU: all planes must
    equal the oracle's before the test's own assertions run."""

    class Mirrored(Oracle):
        def __init__(self, sample_log2=0, sparse=False, super_sampled_textures=False):
            super().__init__(sample_log2, sparse, super_sampled_textures)
            self._cfg = (sample_log2, sparse, super_sampled_textures)
            self._streams = []

        def gif_transfer(self, data, path=3):
            self._streams.append(np.array(data, np.uint8))
            super().gif_transfer(data, path)

        def flush(self):
            self._streams.append(None)
            super().flush()

        def planes(self):
            ref = super().planes()
            log2, sparse, tex = self._cfg
            with pgs.Context(super_sampling=1 << log2, ordered=not sparse, super_sampled_textures=tex) as ctx:
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
    for name in ("test_one_to_one_nearest_blit_copies_every_sample",
                 "test_uploaded_pages_and_triangles_do_not_sample_per_sample",
                 "test_linear_blit_differs_and_one_x_ignores_option"):
        fn = getattr(T, name)
        if name == "test_one_to_one_nearest_blit_copies_every_sample":
            for log2, sparse in [(1, False), (2, False), (2, True), (3, False), (4, True)]:
                print(f"[case] {name} log2={log2} sparse={sparse}", flush=True)
                fn(log2, sparse)
        else:
            print(f"[case] {name}", flush=True)
            fn()


def test_runtime_switch_and_batch_replay(hip_device):
    """Super-sampled textures switched on at run time (pgs_set_super_sampling_rate) render as at
    creation, and a batch of recorded contexts replays the live pixels."""
    stream = rs.random_ssaa_texture_stream(850)
    o = Oracle(2, super_sampled_textures=True)
    o.gif_transfer(stream)
    o.flush()
    ref = o.planes()
    with pgs.Context(super_sampling=4) as ctx:
        assert ctx.set_super_sampling_rate(4, True, True) == 0
        ctx.gif_transfer(stream)
        assert np.array_equal(ctx.vram_planes(), ref)
    ctxs, live = [], []
    for i, ssaa in enumerate((4, 16)):
        s = rs.random_ssaa_texture_stream(860 + i)
        print(f"[case] batch context {i} ssaa={ssaa}", flush=True)
        ctx = pgs.Context(super_sampling=ssaa, super_sampled_textures=True, arena_bytes=32 << 20)
        ctx.begin_recording()
        ctx.gif_transfer(s)
        ctx.end_recording()
        live.append(ctx.vram_planes())
        ctxs.append(ctx)
    # A context without super-sampled textures beside them: its passes keep the plain kernels
    # and never share a multi-launch with the others'.
    plain = pgs.Context(super_sampling=4, arena_bytes=32 << 20)
    plain.begin_recording()
    plain.gif_transfer(rs.random_ssaa_texture_stream(862))
    plain.end_recording()
    live.append(plain.vram_planes())
    ctxs.append(plain)
    b = pgs.Batch(ctxs)
    model = pgs.CostModel(4)
    for name in ("hardware", "static", "learned", "dispatch_static", "dispatch_learned"):
        print(f"[case] batch schedule {name}", flush=True)
        b.set_schedule(name, model if "learned" in name else None)
        b.replay(True)
        b.replay(True)
        b.sync()
        for i, (ctx, want) in enumerate(zip(ctxs, live)):
            assert np.array_equal(ctx.vram_planes(), want), (name, i)
    b.close()
    for ctx in ctxs:
        ctx.close()
