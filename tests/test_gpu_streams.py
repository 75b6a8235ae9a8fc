"""GPU: the texture stream (VRAM copies, CLUT loads and texture decodes on a second HIP stream,
beside triangle setup, the sampling heuristic and binning) changes no pixel. Streams with
uploads, CLUT loads, textured passes, local copies and readbacks render to the same VRAM with
one stream and with two, equal to the oracle's, live and in replays (plain and graph)."""
import numpy as np
import pytest

import pgs
import render_streams as R
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


def oracle(stream, ssaa):
    o = Oracle(LOG2[ssaa])
    o.gif_transfer(stream)
    o.flush()
    return o.planes()


def assert_equal(got, ref):
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


CASES = {
    "textured_frames": lambda: pgs.gen_textured(0x51DE, 600, 3, 320, 224, 0.0),
    "random_textured": lambda: R.random_textured_stream(41),
    "transfers": lambda: pgs.gen_transfers_wild(43, 60),
}


@pytest.mark.parametrize("ssaa", [1, 4])
@pytest.mark.parametrize("case", sorted(CASES))
def test_texture_stream_is_bit_exact(hip_device, case, ssaa):
    stream = CASES[case]()
    ref = oracle(stream, ssaa)
    for two in (False, True):
        print(f"[case] {case} ssaa={ssaa} texture_stream={two}", flush=True)
        with pgs.Context(super_sampling=ssaa, texture_stream=two) as ctx:
            ctx.gif_transfer(stream)
            assert_equal(ctx.vram_planes(), ref)


@pytest.mark.parametrize("use_graph", [False, True])
def test_texture_stream_replays(hip_device, use_graph):
    """Replays re-issue the recorded fork and join points: every replay of a textured
    multi-frame recording ends with the VRAM of the live run."""
    stream = pgs.gen_textured(0x51DF, 800, 4, 320, 224, 0.0)
    print(f"[case] replay use_graph={use_graph}", flush=True)
    with pgs.Context(super_sampling=4, texture_stream=True) as ctx:
        ctx.begin_recording()
        ctx.gif_transfer(stream)
        ctx.end_recording()
        live = ctx.vram_planes()
        assert_equal(live, oracle(stream, 4))
        for _ in range(3):
            ctx.replay(use_graph)
            ctx.sync()
            assert_equal(ctx.vram_planes(), live)


REPLAY_CASES = {
    "textured_frames": (lambda: pgs.gen_textured(0x51E0, 700, 3, 320, 224, 0.0), 4, "auto"),
    "random_textured_1x": (lambda: R.random_textured_stream(44), 1, "auto"),
    "random_textured_4x": (lambda: R.random_textured_stream(45), 4, "auto"),
    "transfers": (lambda: pgs.gen_transfers_wild(46, 80), 1, "auto"),
    "feedback_forced": (lambda: R.random_textured_stream(47), 4, "always"),
}


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("case", sorted(REPLAY_CASES))
def test_fused_replays_match_live(hip_device, case, use_graph):
    """On one stream a recording's replay runs its independent launches of each dependency level
    as one fused launch (copy with setup, CLUT load with heuristic, decode with binning): every
    replay ends with the VRAM of the live run, which equals the oracle's."""
    make, ssaa, feedback = REPLAY_CASES[case]
    stream = make()
    print(f"[case] fused replay {case} ssaa={ssaa} feedback={feedback} use_graph={use_graph}", flush=True)
    with pgs.Context(super_sampling=ssaa, sampler_feedback=feedback) as ctx:
        ctx.begin_recording()
        ctx.gif_transfer(stream)
        ctx.end_recording()
        live = ctx.vram_planes()
        assert_equal(live, oracle(stream, ssaa))
        for _ in range(2):
            ctx.replay(use_graph)
            ctx.sync()
            assert_equal(ctx.vram_planes(), live)
