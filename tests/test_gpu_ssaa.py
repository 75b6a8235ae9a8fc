"""GPU parity for super-sampled contexts (2x, 4x, 8x, 16x): every plane (single-sampled VRAM,
reference, samples) bit for bit against the CPU oracle.

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


def gpu_planes(stream, ssaa):
    print(f"[case] stream={len(stream)} bytes ssaa={ssaa}", flush=True)
    with pgs.Context(super_sampling=ssaa) as ctx:
        ctx.gif_transfer(stream)
        planes = ctx.vram_planes()
        stats = ctx.consume_flush_stats()
    return planes, stats


def render_both(stream, ssaa):
    got, stats = gpu_planes(stream, ssaa)
    o = Oracle(LOG2[ssaa])
    o.gif_transfer(stream)
    o.flush()
    assert_planes_equal(got, o.planes())
    return o, stats


class MirroredOracle(Oracle):
    """Oracle whose planes() also renders the same stream on the GPU and compares."""

    def __init__(self, sample_log2=0):
        super().__init__(sample_log2)
        self._log2 = sample_log2
        self._chunks = []

    def gif_transfer(self, data, path=3):
        self._chunks.append(np.array(data, dtype=np.uint8))
        super().gif_transfer(data, path)

    def planes(self):
        ref = super().planes()
        got, _ = gpu_planes(np.concatenate(self._chunks), 1 << self._log2)
        assert_planes_equal(got, ref)
        return ref


@pytest.fixture
def mirrored(monkeypatch, hip_device):
    monkeypatch.setattr(T, "Oracle", MirroredOracle)


CLOSED_FORM = [
    (T.test_edge_antialiasing_resolve, dict(n_samples=4, seed=0)),
    (T.test_edge_antialiasing_resolve, dict(n_samples=16, seed=3)),
    (T.test_edge_antialiasing_resolve, dict(n_samples=2, seed=5)),
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
