"""GPU parity for super-sampled contexts (2x, 4x, 8x, 16x, and the sparse 4x / 16x grids): every
plane (single-sampled VRAM, reference, samples) bit for bit against the CPU oracle.

The closed-form super-sampling tests of test_ssaa_oracle.py are re-run with each oracle
context mirrored on the GPU: the stream the test feeds the oracle is also rendered by a GPU
context and all planes must match before the closed-form assertions run.
"""
import numpy as np
import pytest

import pgs
import render_streams
import test_ssaa_oracle as T
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


def assert_planes_equal(got, ref):
    assert got.size == ref.size
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]} (plane {bad[0] // (4 << 20)})"


def gpu_planes(stream, ssaa, sparse=False):
    print(f"[case] stream={len(stream)} bytes ssaa={ssaa} sparse={sparse}", flush=True)
    with pgs.Context(super_sampling=ssaa, ordered=not sparse) as ctx:
        ctx.gif_transfer(stream)
        planes = ctx.vram_planes()
        stats = ctx.consume_flush_stats()
    return planes, stats


def render_both(stream, ssaa, sparse=False):
    got, stats = gpu_planes(stream, ssaa, sparse)
    o = Oracle(LOG2[ssaa], sparse)
    o.gif_transfer(stream)
    o.flush()
    assert_planes_equal(got, o.planes())
    return o, stats


class MirroredOracle(Oracle):
    """Oracle whose planes() also renders the same stream on the GPU and compares."""

    def __init__(self, sample_log2=0, sparse=False):
        super().__init__(sample_log2, sparse)
        self._log2 = sample_log2
        self._sparse = sparse
        self._chunks = []

    def gif_transfer(self, data, path=3):
        self._chunks.append(np.array(data, dtype=np.uint8))
        super().gif_transfer(data, path)

    def planes(self):
        ref = super().planes()
        got, _ = gpu_planes(np.concatenate(self._chunks), 1 << self._log2, self._sparse)
        assert_planes_equal(got, ref)
        return ref


@pytest.fixture
def mirrored(monkeypatch, hip_device):
    monkeypatch.setattr(T, "Oracle", MirroredOracle)


CLOSED_FORM = [
    (T.test_edge_antialiasing_resolve, dict(n_samples=4, seed=0)),
    (T.test_edge_antialiasing_resolve, dict(n_samples=16, seed=3)),
    (T.test_edge_antialiasing_resolve, dict(n_samples=2, seed=5)),
    (T.test_edge_antialiasing_resolve, dict(n_samples="4s", seed=7)),
    (T.test_edge_antialiasing_resolve, dict(n_samples="16s", seed=9)),
    (T.test_opaque_sprite_renders_single_sampled, {}),
    (T.test_incremental_pass_reads_samples, {}),
    (T.test_flat_ui_triangle_is_demoted, {}),
    (T.test_flat_triangle_sharing_3d_vertex_stays_antialiased, {}),
    (T.test_depth_resolves_to_minimum, {}),
]


@pytest.mark.parametrize("fn,kw", CLOSED_FORM, ids=[f"{f.__name__}-{i}" for i, (f, _) in enumerate(CLOSED_FORM)])
def test_closed_form_on_gpu(mirrored, fn, kw):
    fn(**kw)


@pytest.mark.parametrize("ssaa,seed", [(4, 0), (4, 1), (4, 2), (2, 3), (8, 4), (16, 5), (16, 6), (4, 7)])
def test_random_state_ssaa(hip_device, ssaa, seed):
    render_both(render_streams.random_render_stream(seed, n_batches=8, prims_per_batch=30), ssaa)


@pytest.mark.parametrize("ssaa,seed", [(4, 0), (4, 1), (2, 2), (16, 3)])
def test_random_textured_ssaa(hip_device, ssaa, seed):
    render_both(render_streams.random_textured_stream(seed, n_batches=6, prims_per_batch=25), ssaa)


@pytest.mark.parametrize("ssaa,seed", [(4, 20), (4, 21), (16, 22), (16, 23)])
def test_random_state_sparse_grids(hip_device, ssaa, seed):
    """The sparse 1 x 4 (4x) and 2 x 8 (16x) grids (gs/gs_interface.cpp:84-113)."""
    render_both(render_streams.random_render_stream(seed, n_batches=8, prims_per_batch=30), ssaa, sparse=True)


@pytest.mark.parametrize("ssaa,seed", [(4, 24), (16, 25)])
def test_random_textured_sparse_grids(hip_device, ssaa, seed):
    render_both(render_streams.random_textured_stream(seed, n_batches=6, prims_per_batch=25), ssaa, sparse=True)


@pytest.mark.parametrize("ssaa", [4, 16])
def test_textured_frame_sparse_grids(hip_device, ssaa):
    """A 640x448 textured, blended frame (the SSAA workload's generator) on the sparse grids."""
    render_both(pgs.gen_textured(0x5EED0060 + ssaa, 1500, 1, 640, 448, 0.0), ssaa, sparse=True)


def test_rate_change_to_a_sparse_grid(hip_device):
    """A dynamic context switched from the ordered to the sparse 4x grid and to sparse 16x."""
    s1 = render_streams.random_render_stream(26, n_batches=3, prims_per_batch=20)
    s2 = render_streams.random_render_stream(27, n_batches=3, prims_per_batch=20)
    s3 = render_streams.random_render_stream(28, n_batches=3, prims_per_batch=20)
    o = Oracle(2)
    with pgs.Context(super_sampling=4, dynamic_super_sampling=True) as ctx:
        for st, (samples, ordered) in ((s1, (4, True)), (s2, (4, False)), (s3, (16, False))):
            print(f"[case] rate change to {samples} ordered={ordered}", flush=True)
            if st is not s1:
                assert ctx.set_super_sampling_rate(samples, ordered) == 0
                o.set_sample_log2(LOG2[samples], sparse=not ordered)
            ctx.gif_transfer(st)
            o.gif_transfer(st)
        got = ctx.vram_planes()
    o.flush()
    ref = o.planes()
    assert_planes_equal(got[:ref.size], ref)
