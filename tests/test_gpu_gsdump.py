"""GPU tests of GS dumps: replay (pgs_dump_*) against the oracle, the C ABI writer against the
format, save states through a dump, and the regression replayer.

GSDumpParser / GSDumpGenerator (dump/gs_dump_parser.cpp:6-207, dump/gs_dump_generator.cpp:
39-237) and tools/gs_repro_replayer.cpp:29-330.
"""
import os

import numpy as np
import pytest

import crtc
import gsdump
import pgs
import render_streams as rs
from oracle_lib import Oracle
from pgs import replay
from test_gpu_api import packet_offsets
from test_gsdump_oracle import textured_dump

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


@pytest.mark.parametrize("samples,high_res", [(1, False), (4, True)])
def test_dump_replay_matches_oracle(hip_device, tmp_path, samples, high_res):
    """Every vsync of a dump (two textured frames) replays to the oracle's frame, and VRAM (all
    planes) ends equal."""
    w = textured_dump(0x5EED00E1 + samples, n_prims=400, frames=2)
    path = w.save(str(tmp_path / "frames.gs"))
    o = Oracle(LOG2[samples])
    rc, meta_o, img_o = gsdump.replay_into_oracle(o, bytes(w.data), high_res_scanout=high_res)
    print(f"[case] dump replay {samples}x high_res={high_res}", flush=True)
    with pgs.Context(super_sampling=samples) as ctx:
        r = pgs.DumpReader(ctx, path)
        stops = 0
        while r.iterate_until_vsync(high_res_scanout=high_res):
            stops += 1
        meta, img = r.consume_vsync_result()
        assert stops == 2
        assert meta == meta_o and np.array_equal(img, img_o)
        o.flush()
        assert np.array_equal(ctx.vram_planes(), o.planes())
        # consume hands the result over once.
        assert r.consume_vsync_result()[1] is None
        # restart loads the start state again: the first frame replays to the same pixels.
        assert r.restart()
        assert r.iterate_until_vsync(high_res_scanout=high_res)
        r.close()


def python_dump_of_state(ctx, vram):
    regs = ctx.register_state()
    d = {}
    for n in gsdump.GLOBAL:
        d[n] = int(getattr(regs, n))
    for c in range(2):
        for n in gsdump.PER_CONTEXT:
            d[f"{n}{c}"] = int(getattr(regs.ctx[c], n))
    d["rgbaq"], d["st"], d["uv"], d["fog"] = int(regs.rgbaq), int(regs.st), int(regs.uv), int(regs.fog)
    paths = []
    for i in range(4):
        p = ctx.gif_path(i)
        lo = int(p.tag[0])
        nloop = lo & 0x7FFF
        paths.append(((lo & ~0x7FFF) | ((nloop - int(p.loop)) & 0x7FFF), int(p.tag[1]), int(p.reg)))
    return gsdump.Writer(vram=vram, regs=d, paths=paths, internal_q=float(regs.internal_q),
                         priv=ctx.priv_registers().tobytes())


def test_writer_matches_the_format(hip_device, tmp_path):
    """The C ABI writer's file equals the format built field by field from the context's state,
    with register writes, a PACKED packet, an image upload and a vsync appended."""
    stream = rs.random_render_stream(41, n_batches=3, prims_per_batch=10, width=128, height=96)
    q = stream.view(np.uint64)
    cut = packet_offsets(stream)[2] + 16 * 3
    with pgs.Context() as ctx:
        ctx.gif_transfer(stream[:cut])
        ctx.set_priv_registers(np.frombuffer(crtc.ntsc_interlaced_640x448(), np.uint8))
        path = str(tmp_path / "w.gs")
        w = pgs.DumpWriter(ctx, path)
        expect = python_dump_of_state(ctx, ctx.vram())
        w.write_register(0x4C, 0x0123456789)
        data = np.arange(2 * 3 * 16, dtype=np.uint8)
        w.write_packed(0x56, [1, 5, 0xE], 2, data)
        img = np.random.default_rng(1).integers(0, 256, 64 * 8 * 4, dtype=np.uint8)
        w.write_image_upload(0x2000, 0x00, 64, 8, img)
        w.write_vsync(1)
        w.close()
    import gif

    def ad_eop(addr, value):  # the writer's A+D packets end with EOP set
        return gif.arr(gif.giftag(1, eop=True) + np.array([value, addr], np.uint64).tobytes())

    e = expect
    e.transfer(ad_eop(0x4C, 0x0123456789))
    tag_lo = 2 | 1 << 15 | 1 << 46 | 0x56 << 47 | 3 << 60
    e.transfer(np.concatenate([np.array([tag_lo, 1 | 5 << 4 | 0xE << 8], np.uint64).view(np.uint8), data]))
    bitblt = (0x2000 // 256) << 32 | 1 << 48
    for a, v in ((0x50, bitblt), (0x51, 0), (0x52, 64 | 8 << 32), (0x53, 0)):
        e.transfer(ad_eop(a, v))
    e.transfer(np.concatenate([np.array([img.size // 16 | 1 << 15 | 2 << 58, 0], np.uint64).view(np.uint8), img]))
    e.vsync(1, crtc.ntsc_interlaced_640x448())
    got, want = open(path, "rb").read(), bytes(e.data)
    first = next((i for i in range(min(len(got), len(want))) if got[i] != want[i]), None)
    assert got == want, f"sizes {len(got)} / {len(want)}, first difference at byte {first}"


def test_save_state_through_a_dump(hip_device, tmp_path):
    """Run part of a stream (cut inside a GIF packet), write a dump of the state, open it in a
    fresh context and continue the stream: VRAM equals the oracle's uninterrupted run."""
    stream = rs.random_render_stream(51, n_batches=6, prims_per_batch=20, width=128, height=96)
    q = stream.view(np.uint64)
    regs_packets = [o for o in packet_offsets(stream) if int(q[(o + 16) // 8 + 1]) & 0xFF == 0x4C]
    cut = regs_packets[3] + 16 * 6
    o = Oracle()
    o.gif_transfer(stream)
    ref = o.vram()
    path = str(tmp_path / "state.gs")
    with pgs.Context() as a:
        a.gif_transfer(stream[:cut])
        pgs.DumpWriter(a, path).close()
    print(f"[case] save state through a dump, cut={cut}", flush=True)
    with pgs.Context() as b:
        r = pgs.DumpReader(b, path)
        assert not r.iterate_until_vsync()  # no packets: the end
        b.gif_transfer(stream[cut:])
        got = b.vram()
        r.close()
    assert np.array_equal(got, ref)


def test_replayer_regression_harness(hip_device, tmp_path):
    """The replayer writes references on the first run, passes (infinite PSNR, same statistics)
    on the second, and its reference image equals the oracle's frame; a changed reference
    fails and leaves a fault image."""
    for k in range(2):
        textured_dump(0x5EED00F1 + k, n_prims=300, frames=1).save(str(tmp_path / f"d{k}.gs"))
    assert replay.main(["--dir", str(tmp_path), "--samples", "4"]) == 0
    assert replay.main(["--dir", str(tmp_path), "--samples", "4"]) == 0
    for k in range(2):
        p = str(tmp_path / f"d{k}.gs")
        ok, details = replay.run_dump(p, samples=4)
        assert ok and details["psnr"] == [float("inf")] * 3
        _, _, img = gsdump.replay_into_oracle(Oracle(2), open(p, "rb").read())
        assert np.array_equal(replay.read_png(p + ".reference.png"), replay.frame_rgba(img))
    ref = replay.read_png(str(tmp_path / "d0.gs.reference.png"))
    ref[::2, ::3, :3] ^= 0x40
    replay.write_png(str(tmp_path / "d0.gs.reference.png"), ref)
    assert replay.main(["--dir", str(tmp_path), "--samples", "4"]) == 1
    assert os.path.exists(str(tmp_path / "d0.gs.fault.png"))
