"""GPU parity of the two batched configurations at full size, against the CPU oracle.

Config 4: 64 contexts (N log-uniform 500..50000, 512/640x448, 1x/2x/4x, CT32/T8/T4 textures with
CLUT, four blend presets) replayed as one merged, autotuned, graph-captured batch.
Config 5: 64 contexts at 16x with Pareto(1.2) primitive areas; the tile-cost MLP is trained on
the GPU from costs measured on a separate training batch, then orders the work queue and the
per-XCD dispatch order. Every context's VRAM (all planes) must equal the oracle's under every
schedule; the trained network's gain over the untrained one and over the static order is
measured and printed (render order per context: gs/gs_renderer.cpp:3126-3285). No config-5 tile
is heavy enough to need splitting, so a batch with one hot tile exercises the learned queue's
split and the learned dispatch order's front-loading. The oracle renders the contexts on
parallel host threads.
"""
import hashlib
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import pgs
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}
THREADS = max(1, min(16, os.cpu_count() or 1))


def digest(a: np.ndarray) -> str:
    return hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=20).hexdigest()


def oracle_digests(specs):
    def run(sp):
        o = Oracle(LOG2[sp["ssaa"]])
        o.gif_transfer(pgs.spec_stream(sp))
        o.flush()
        return digest(o.planes())

    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        return list(ex.map(run, specs))


def record(specs):
    ctxs = []
    for i, sp in enumerate(specs):
        ctx = pgs.Context(super_sampling=sp["ssaa"], arena_bytes=64 << 20)
        ctx.begin_recording()
        ctx.gif_transfer(pgs.spec_stream(sp))
        ctx.end_recording()
        ctxs.append(ctx)
    return ctxs


def gpu_digests(ctxs):
    return [digest(c.vram_planes()) for c in ctxs]


def assert_all_match(ctxs, ref, what):
    got = gpu_digests(ctxs)
    bad = [i for i in range(len(ref)) if got[i] != ref[i]]
    assert not bad, f"{what}: contexts {bad} differ from the oracle"


def time_replays(batch, reps=5):
    import torch

    batch.replay(True)  # planning replay for planned schedules
    batch.replay(True)
    batch.sync()
    s = torch.cuda.ExternalStream(batch.stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        batch.replay(True)
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def test_config4_batched_64_contexts(hip_device):
    specs = pgs.batched_specs(0x5EED0004, 64)
    print(f"[case] config 4: {len(specs)} contexts, {sum(s['n_prims'] for s in specs)} primitives", flush=True)
    ctxs = record(specs)
    ref = oracle_digests(specs)
    assert_all_match(ctxs, ref, "live")
    batch = pgs.Batch(ctxs)
    table = batch.autotune(1)
    print("autotune", table, flush=True)
    batch.replay(True)
    batch.sync()
    assert_all_match(ctxs, ref, "merged batch, autotuned")
    batch.close()
    for c in ctxs:
        c.close()


def spearman(a, b):
    ra = np.argsort(np.argsort(a)).astype(np.float64)
    rb = np.argsort(np.argsort(b)).astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / np.sqrt((ra * ra).sum() * (rb * rb).sum()))


def test_config5_learned_schedule_64_contexts(hip_device):
    common = dict(ssaa_choices=(16,), pareto_alpha=1.2, tex_psms=(0x13,), blend_presets=(0,))
    # Training batch: costs measured by the plain kernels, network trained on the GPU.
    train_ctxs = record(pgs.batched_specs(0x5EED0015, 32, **common))
    tb = pgs.Batch(train_ctxs)
    feats, costs = tb.collect_costs()
    tb.close()
    for c in train_ctxs:
        c.close()
    model = pgs.CostModel(7)
    untrained = pgs.CostModel(7)
    mse = model.train(feats, costs, epochs=8, batch=1024, lr=2e-3, seed=3)
    print(f"[case] config 5: trained on {feats.shape[0]} tile costs, mse {mse:.4f}, var {np.var(costs):.4f}",
          flush=True)

    specs = pgs.batched_specs(0x5EED0005, 64, **common)
    ctxs = record(specs)
    ref = oracle_digests(specs)
    assert_all_match(ctxs, ref, "live")
    batch = pgs.Batch(ctxs)
    ev_feats, ev_costs = batch.collect_costs()
    p_trained, p_untrained = model.predict(ev_feats), untrained.predict(ev_feats)
    rho_t, rho_u = spearman(p_trained, ev_costs), spearman(p_untrained, ev_costs)
    mse_t = float(np.mean((p_trained - ev_costs) ** 2))
    mse_u = float(np.mean((p_untrained - ev_costs) ** 2))
    print(f"evaluation batch: spearman trained {rho_t:.3f} untrained {rho_u:.3f}; "
          f"mse trained {mse_t:.3f} untrained {mse_u:.3f}", flush=True)
    assert rho_t > 0.8 and mse_t < 0.5 * mse_u

    times = {}
    for name, m in (("hardware", None), ("static", None), ("learned", model), ("learned_untrained", untrained),
                    ("dispatch_static", None), ("dispatch_learned", model)):
        batch.set_schedule(name.replace("_untrained", ""), m)
        times[name] = time_replays(batch)
        st = batch.schedule_stats()
        print(f"schedule {name}: {times[name]:.3f} ms per replay, plan {st}", flush=True)
        assert_all_match(ctxs, ref, name)
    print(f"learned gain: vs untrained {times['learned_untrained'] / times['learned']:.4f}, "
          f"vs static queue {times['static'] / times['learned']:.4f}, "
          f"dispatch vs static {times['dispatch_static'] / times['dispatch_learned']:.4f}, "
          f"vs hardware {times['hardware'] / min(times['learned'], times['dispatch_learned']):.4f}", flush=True)
    batch.close()
    for c in ctxs:
        c.close()


def hot_tile_stream(seed, n_hot=4000, n_background=60):
    """A 16x frame whose cost sits in one tile: n_hot blended, depth-tested triangles inside a
    16x16 pixel square, over a few large background triangles."""
    import gif
    import render_streams as rs

    rng = np.random.default_rng(seed)
    ofs = 2048 * 16
    regs = [(0x1A, 1), (0x18, ofs | ofs << 32), (0x4C, rs.frame(0, 10, 0x00)), (0x4E, rs.zbuf(140, 0x1)),
            (0x40, rs.scissor(0, 639, 0, 447)), (0x47, rs.test_reg(zte=1, ztst=2)),
            (0x42, rs.alpha_reg(0, 1, 0, 1, 0)), (0x00, rs.prim_reg(3, iip=1, abe=1))]
    writes = []
    for k in range(n_background + n_hot):
        hot = k >= n_background
        for _ in range(3):
            if hot:
                x, y = 160 * 16 + int(rng.integers(0, 16 * 16)), 96 * 16 + int(rng.integers(0, 16 * 16))
            else:
                x, y = int(rng.integers(0, 640 * 16)), int(rng.integers(0, 448 * 16))
            writes.append((0x01, rs.rgbaq(int(rng.integers(0, 2**32)) & 0x7FFFFFFF)))
            writes.append((0x05, rs.xyz2(ofs + x, ofs + y, int(rng.integers(0, 2**24)))))
    out = [gif.ad(regs)] + [gif.ad(writes[i:i + 600]) for i in range(0, len(writes), 600)]
    return gif.arr(b"".join(out))


def test_learned_schedules_split_and_front_load_a_hot_tile(hip_device):
    """One context whose work sits in one tile, batched with light contexts: the trained network
    predicts the hot tile's cost, the learned queue splits it into more items than its base items
    and the learned dispatch order starts it first on its XCD. Pixels equal the oracle's."""
    common = dict(ssaa_choices=(16,), pareto_alpha=1.2, tex_psms=(0x13,), blend_presets=(0,))
    light = pgs.batched_specs(0x5EED0025, 3, n_min=300, n_max=1500, **common)
    streams = [hot_tile_stream(5)] + [pgs.spec_stream(sp) for sp in light]
    ctxs = []
    for st in streams:
        c = pgs.Context(super_sampling=16, arena_bytes=32 << 20)
        c.begin_recording()
        c.gif_transfer(st)
        c.end_recording()
        ctxs.append(c)

    def run(st):
        o = Oracle(4)
        o.gif_transfer(st)
        o.flush()
        return digest(o.planes())

    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        ref = list(ex.map(run, streams))
    batch = pgs.Batch(ctxs)
    feats, costs = batch.collect_costs()
    model = pgs.CostModel(5)
    model.train(feats, costs, epochs=40, batch=256, lr=3e-3, seed=1)
    hot = int(np.argmax(costs))
    print(f"[case] hot tile: {feats[hot, 0]:.0f} primitives, log2 cost {costs[hot]:.2f} "
          f"(median {np.median(costs):.2f}), predicted {model.predict(feats[hot:hot + 1])[0]:.2f}", flush=True)
    for name in ("learned", "dispatch_learned"):
        batch.set_schedule(name, model)
        batch.replay(True)  # planning replay
        batch.replay(True)
        batch.sync()
        st = batch.schedule_stats()
        print(f"schedule {name}: plan {st}", flush=True)
        if name == "learned":
            assert st["split_tiles"] > 0, "the learned queue must split the hot tile"
        else:
            assert st["heavy_first"] > 0, "the learned dispatch order must start the hot tile first"
        assert_all_match(ctxs, ref, name)
    batch.close()
    for c in ctxs:
        c.close()
