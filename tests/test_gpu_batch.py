"""GPU: batched replay of several contexts (merged multi-context launches, hipGraph) gives each
context exactly the VRAM its own live run and the CPU oracle give, for every tuning."""
import numpy as np
import pytest

import pgs
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


def make_contexts(specs):
    ctxs, live = [], []
    for sp in specs:
        stream = pgs.spec_stream(sp)
        print(f"[case] context n={sp['n_prims']} ssaa={sp['ssaa']} tex={sp['tex_psm']:#x} blend={sp['blend']}",
              flush=True)
        ctx = pgs.Context(super_sampling=sp["ssaa"], arena_bytes=32 << 20)
        ctx.begin_recording()
        ctx.gif_transfer(stream)
        ctx.end_recording()
        live.append(ctx.vram_planes())
        ctxs.append(ctx)
    return ctxs, live


def oracle_planes(sp):
    o = Oracle(LOG2[sp["ssaa"]])
    o.gif_transfer(pgs.spec_stream(sp))
    o.flush()
    return o.planes()


def test_batch_matches_live_and_oracle(hip_device):
    specs = pgs.batched_specs(0x5EED0004, 6, n_min=300, n_max=3000)
    ctxs, live = make_contexts(specs)
    for sp, planes in zip(specs, live):
        assert np.array_equal(planes, oracle_planes(sp))
    batch = pgs.Batch(ctxs)
    print("merged launches", batch.num_launches, flush=True)
    assert batch.num_launches < sum(1 for _ in ctxs) * 6
    for tuning in [(0, 4), (3, 1), (4, 2), (5, 4)]:
        batch.apply_tuning(*tuning)
        batch.replay(True)
        batch.replay(True)
        batch.sync()
        for i, ctx in enumerate(ctxs):
            got = ctx.vram_planes()
            bad = np.flatnonzero(got != live[i])
            assert bad.size == 0, (tuning, i, bad[:8])
    batch.close()
    seq = pgs.Batch(ctxs, sequential=True)
    seq.replay(True)
    seq.sync()
    for i, ctx in enumerate(ctxs):
        assert np.array_equal(ctx.vram_planes(), live[i])
    seq.close()
    for c in ctxs:
        c.close()


def test_autotune_keeps_pixels(hip_device):
    specs = pgs.batched_specs(0x5EED0005, 4, n_min=300, n_max=2000, ssaa_choices=(1, 4))
    ctxs, live = make_contexts(specs)
    batch = pgs.Batch(ctxs)
    table = batch.autotune(2)
    print(table, flush=True)
    assert len(table) >= 6
    batch.replay(True)
    batch.sync()
    for i, ctx in enumerate(ctxs):
        assert np.array_equal(ctx.vram_planes(), live[i])
    batch.close()
    for c in ctxs:
        c.close()


def test_replay_restores_vram(hip_device):
    """A single context's graph replay starts from the recording's initial VRAM every time."""
    stream = pgs.gen_textured(0x5EED0003, 1500, 2, 640, 448, 0.0)
    with pgs.Context(super_sampling=4) as ctx:
        ctx.begin_recording()
        ctx.gif_transfer(stream)
        ctx.end_recording()
        live = ctx.vram_planes()
        for _ in range(3):
            ctx.replay(True)
        ctx.sync()
        assert np.array_equal(ctx.vram_planes(), live)


def test_schedules_are_bit_exact(hip_device):
    """Hardware dispatch, the static and learned work queues and the static and learned
    dispatch orders shade the same tiles: the VRAM of every context matches its live run under
    each schedule, also with 1 and 2 waves per workgroup (chunks of other sizes)."""
    specs = pgs.batched_specs(0x5EED0006, 5, n_min=300, n_max=3000, ssaa_choices=(1, 2, 4, 16))
    ctxs, live = make_contexts(specs)
    batch = pgs.Batch(ctxs)
    model = pgs.CostModel(4)
    runs = [(4, "static"), (4, "learned"), (4, "hardware"), (4, "learned"), (4, "dispatch_static"),
            (4, "dispatch_learned"), (1, "dispatch_learned"), (2, "dispatch_static"), (2, "learned"), (4, "hardware")]
    for waves, name in runs:
        print(f"[case] schedule {name}, {waves} waves per workgroup", flush=True)
        batch.apply_tuning(0, waves)
        batch.set_schedule(name, model if "learned" in name else None)
        batch.replay(True)
        batch.replay(True)
        batch.sync()
        for i, ctx in enumerate(ctxs):
            assert np.array_equal(ctx.vram_planes(), live[i]), (name, waves, i)
    feats, costs = batch.collect_costs()
    print("collected", feats.shape, "cost range", costs.min(), costs.max(), flush=True)
    assert feats.shape[0] > 100 and np.all(np.isfinite(feats)) and np.all(costs > 0)
    for i, ctx in enumerate(ctxs):
        assert np.array_equal(ctx.vram_planes(), live[i])
    batch.close()
    for c in ctxs:
        c.close()


def test_large_contexts_bin_in_several_windows(hip_device):
    """Contexts of 40000 primitives: binning takes a tile's coarse-mask candidates in several
    windows of primitives, live (own launches) and batched (multi-launches). Every primitive is
    binned once: with blending, a primitive listed twice would blend twice."""
    specs = []
    for ssaa, blend in [(1, 0), (4, 1)]:
        sp = pgs.batched_specs(0x5EED0011 + ssaa, 1, n_min=40000, n_max=40000, ssaa_choices=(ssaa,),
                               blend_presets=(blend,))[0]
        specs.append(sp)
    ctxs, live = make_contexts(specs)
    for sp, planes in zip(specs, live):
        ref = oracle_planes(sp)
        bad = np.flatnonzero(planes != ref)
        assert bad.size == 0, (sp, bad[:8])
    batch = pgs.Batch(ctxs)
    batch.replay(True)
    batch.sync()
    for i, ctx in enumerate(ctxs):
        assert np.array_equal(ctx.vram_planes(), live[i]), i
    batch.close()
    for c in ctxs:
        c.close()


def test_sparse_grid_contexts_in_a_batch(hip_device):
    """Contexts on the sparse 1 x 4 and 2 x 8 grids batched beside ordered ones: live runs equal
    the oracle, and every schedule replays them to the same VRAM."""
    specs = pgs.batched_specs(0x5EED0007, 4, n_min=300, n_max=2000, ssaa_choices=(4, 16))
    ctxs, live = [], []
    for k, sp in enumerate(specs):
        sparse = k % 2 == 0
        print(f"[case] context n={sp['n_prims']} ssaa={sp['ssaa']} sparse={sparse}", flush=True)
        ctx = pgs.Context(super_sampling=sp["ssaa"], ordered=not sparse, arena_bytes=32 << 20)
        ctx.begin_recording()
        ctx.gif_transfer(pgs.spec_stream(sp))
        ctx.end_recording()
        live.append(ctx.vram_planes())
        ctxs.append(ctx)
        o = Oracle(LOG2[sp["ssaa"]], sparse)
        o.gif_transfer(pgs.spec_stream(sp))
        o.flush()
        assert np.array_equal(live[-1], o.planes()), k
    batch = pgs.Batch(ctxs)
    model = pgs.CostModel(5)
    for name in ("hardware", "static", "learned", "dispatch_static", "dispatch_learned"):
        print(f"[case] sparse batch schedule {name}", flush=True)
        batch.set_schedule(name, model if "learned" in name else None)
        batch.replay(True)
        batch.replay(True)
        batch.sync()
        for i, ctx in enumerate(ctxs):
            assert np.array_equal(ctx.vram_planes(), live[i]), (name, i)
    batch.close()
    for c in ctxs:
        c.close()
