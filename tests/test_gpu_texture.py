"""GPU parity for texturing: CLUT loads, texture decode, sampling and combiners, bit for bit
against the CPU oracle.

The closed-form texture tests of test_texture_oracle.py are re-run with their streams also
rendered on the GPU: the whole VRAM must match the oracle's byte for byte, and the closed-form
assertions then hold for the GPU result as well.
"""
import numpy as np
import pytest

import pgs
import render_streams
import test_texture_oracle as T
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def assert_vram_equal(got, ref):
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


def render_both(stream):
    print(f"[case] stream={len(stream)} bytes", flush=True)
    with pgs.Context() as ctx:
        ctx.gif_transfer(stream)
        got = ctx.vram_planes()
        stats = ctx.consume_flush_stats()
    o = Oracle()
    o.gif_transfer(stream)
    o.flush()
    assert_vram_equal(got, o.planes())
    return o, stats


@pytest.fixture
def gpu_checked_run(monkeypatch, hip_device):
    def run(stream_bytes):
        o, _ = render_both(np.frombuffer(stream_bytes, dtype=np.uint8))
        return o
    monkeypatch.setattr(T, "run", run)


CLOSED_FORM = [
    (T.test_nearest_identity_ct32, dict(tw=6, th=5)),
    (T.test_nearest_identity_ct32, dict(tw=3, th=4)),
    (T.test_nearest_identity_ct24_texa, dict(aem=1)),
    (T.test_nearest_identity_ct16_texa, dict(psm=0x02, aem=1)),
    (T.test_nearest_identity_ct16_texa, dict(psm=0x0A, aem=1)),
    (T.test_palette_t8_csm1_ct32, {}),
    (T.test_palette_t4_ct16_csa_banks, {}),
    (T.test_palette_csm2_texclut_offsets, {}),
    (T.test_cld_cached_cbp, {}),
    (T.test_wrap_modes, dict(mode="repeat")),
    (T.test_wrap_modes, dict(mode="clamp")),
    (T.test_wrap_modes, dict(mode="region_clamp")),
    (T.test_wrap_modes, dict(mode="region_clamp_reversed")),
    (T.test_modulate_closed_form, {}),
    (T.test_bilinear_at_texel_centres_is_exact, {}),
    (T.test_bilinear_midpoints_of_linear_ramp, {}),
    (T.test_perspective_constant_q_matches_uv, dict(q=2.0)),
    (T.test_texture_reads_see_pending_primitives, {}),
    (T.test_tex2_reloads_clut, {}),
]


@pytest.mark.parametrize("fn,kw", CLOSED_FORM, ids=[f"{f.__name__}-{i}" for i, (f, _) in enumerate(CLOSED_FORM)])
def test_closed_form_on_gpu(gpu_checked_run, fn, kw):
    fn(**kw)


@pytest.mark.parametrize("seed", list(range(8)))
def test_random_textured(hip_device, seed):
    stream = render_streams.random_textured_stream(seed)
    o, stats = render_both(stream)
    print("passes", stats["render_passes"], "oracle passes", o.stat(2), flush=True)
