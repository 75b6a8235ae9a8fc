"""GPU parity of the drop-in API beyond gif_transfer (gs/gs_interface.hpp:251-305), through the
C ABI, against the CPU oracle: PRMODE/PRMODECONT, LOCAL->HOST readback of every readable
format with draws still pending, host VRAM writes (texture-cache invalidation and super-sample
invalidation), reset_context_state, save-state round trips (register file, GIF paths, VRAM),
runtime super-sampling rate changes, the signal interface, timestamps, and recordings that
stay replayable after later work."""
import numpy as np
import pytest

import gif
import pgs
import render_streams as rs
from oracle_lib import Oracle, VRAM_SIZE
from test_api_oracle import READABLE, sprite_stream

pytestmark = pytest.mark.gpu

OFS = 2048 * 16


def assert_same(got, ref, what=""):
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{what}: {bad.size} bytes differ, first at {bad[:8]}"


def prmode_stream(seed):
    """The random render stream with PRMODECONT / PRMODE toggled between batches."""
    rng = np.random.default_rng(seed)
    base = rs.random_render_stream(seed, n_batches=10, prims_per_batch=24, width=128, height=96)
    # Split at A+D packets: each random batch starts with a register packet. Interleave toggles.
    out = []
    raw = base.tobytes()
    step = len(raw) // 16 // 6 * 16
    for i in range(0, len(raw), step):
        chunk = raw[i:i + step]
        out.append(chunk)
    # Chunks may cut GIF packets, so toggles go through a different GIF path (path 1).
    toggles = []
    for _ in out:
        toggles.append(gif.ad([(0x1A, int(rng.integers(0, 2))),
                               (0x1B, rs.prim_reg(0, iip=int(rng.integers(0, 2)), abe=int(rng.integers(0, 2)),
                                                  fge=int(rng.integers(0, 2)), ctxt=0))]))
    return [(3, gif.arr(c)) for c in out], [(1, gif.arr(t)) for t in toggles]


@pytest.mark.parametrize("seed", [11, 12])
def test_prmode_toggles_match_oracle(hip_device, seed):
    print(f"[case] prmode seed={seed}", flush=True)
    chunks, toggles = prmode_stream(seed)
    o = Oracle()
    with pgs.Context() as ctx:
        for (p, c), (tp, t) in zip(chunks, toggles):
            for path, data in ((tp, t), (p, c)):
                ctx.gif_transfer(data, path)
                o.gif_transfer(data, path)
        got = ctx.vram()
    assert_same(got, o.vram(), "prmode")


def test_prmode_context_select(hip_device):
    for prmodecont, prmode, prim in ((0, rs.prim_reg(3, ctxt=1), rs.prim_reg(6, ctxt=0)),
                                     (1, rs.prim_reg(3, ctxt=1), rs.prim_reg(6, ctxt=0)),
                                     (1, None, rs.prim_reg(6, ctxt=1))):
        s = sprite_stream(prmodecont, prmode, prim)
        o = Oracle()
        o.gif_transfer(s)
        with pgs.Context() as ctx:
            ctx.gif_transfer(s)
            assert_same(ctx.vram(), o.vram(), f"ac={prmodecont}")


@pytest.mark.parametrize("psm", READABLE)
def test_readback_after_pending_draw(hip_device, psm):
    """Upload, draw a sprite over part of it, then read back with the draw still pending (no
    flush in between): the FIFO equals the oracle's, which renders first."""
    print(f"[case] readback psm={psm:#x}", flush=True)
    rng = np.random.default_rng(psm + 100)
    w, h = 48, 40
    payload = rng.integers(0, 256, w * h * gif.BPP[psm] // 8, dtype=np.uint8).tobytes()
    fb_psm = psm if psm in (0x00, 0x01, 0x02, 0x0A, 0x30, 0x31, 0x32, 0x3A) else 0x00
    draw = gif.ad([(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 1, fb_psm)),
                   (0x40, rs.scissor(0, 63, 0, 63)), (0x47, rs.test_reg(zte=1, ztst=1)), (0x4E, rs.zbuf(32, 0, 1)),
                   (0x00, rs.prim_reg(6)), (0x01, rs.rgbaq(0x8F6F4F2F)),
                   (0x05, rs.xyz2(OFS + 6 * 16, OFS + 3 * 16, 0)), (0x05, rs.xyz2(OFS + 30 * 16, OFS + 21 * 16, 0))])
    read = gif.ad([(0x50, gif.bitbltbuf(sbp=0, sbw=1, spsm=psm)), (0x51, gif.trxpos(2, 1)),
                   (0x52, gif.trxreg(w, h)), (0x53, 1)])
    stream = gif.arr(gif.upload(0, 1, psm, 2, 1, w, h, payload) + draw + read)
    n128 = len(payload) // 16
    o = Oracle()
    o.gif_transfer(stream)
    ref = o.read_transfer_fifo(n128 + 2)
    with pgs.Context() as ctx:
        ctx.gif_transfer(stream)
        got = ctx.read_transfer_fifo(n128 + 2)
    assert_same(got, ref, f"fifo psm {psm:#x}")
    assert got[:len(payload)].tobytes() != payload, "the draw must be visible in the readback"


def textured_sprite(tbp, color=0x80808080):
    return gif.ad([(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 1, 0x00)),
                   (0x40, rs.scissor(0, 63, 0, 63)), (0x47, rs.test_reg(zte=1, ztst=1)), (0x4E, rs.zbuf(32, 0, 1)),
                   (0x14, rs.tex1(0, 0)), (0x08, rs.clamp_reg(1, 1)),
                   (0x06, rs.tex0(tbp, 1, 0x00, 4, 4, tcc=1, tfx=1)),
                   (0x00, rs.prim_reg(6, tme=1, fst=1)), (0x01, rs.rgbaq(color)),
                   (0x03, rs.uv(0, 0)), (0x05, rs.xyz2(OFS, OFS, 0)),
                   (0x03, rs.uv(16 * 16, 16 * 16)), (0x05, rs.xyz2(OFS + 16 * 16, OFS + 16 * 16, 0))])


def test_host_write_invalidates_cached_texture(hip_device):
    """A texture decoded by one pass is cached; a host write over its pages must make the next
    pass decode again (gs/page_tracker.cpp:842-933 invalidation)."""
    rng = np.random.default_rng(3)
    tex_a = rng.integers(0, 256, 16 * 16 * 4, dtype=np.uint8).tobytes()
    tbp = 64 * 32
    o = Oracle()
    with pgs.Context() as ctx:
        first = gif.arr(gif.upload(tbp, 1, 0x00, 0, 0, 16, 16, tex_a) + textured_sprite(tbp))
        ctx.gif_transfer(first)
        o.gif_transfer(first)
        ctx.flush()
        # Overwrite the whole texture page from the host.
        new_page = rng.integers(0, 256, 8192, dtype=np.uint8)
        ctx.write_vram(new_page, tbp * 256)
        o.host_write(new_page, tbp * 256)
        second = gif.arr(textured_sprite(tbp, 0x80406080))
        ctx.gif_transfer(second)
        o.gif_transfer(second)
        got = ctx.vram()
        stats = ctx.consume_flush_stats()
    assert_same(got, o.vram(), "host write")
    assert stats["textures_decoded"] == 2, stats


def test_host_write_ssaa_invalidates_samples(hip_device):
    """At 4x a host write clears the reference plane of its pages, so the next pass reads the
    written single-sampled value (every plane compared)."""
    tri = [(0x1A, 1), (0x18, OFS | OFS << 32), (0x4C, rs.frame(0, 1, 0x00)), (0x40, rs.scissor(0, 63, 0, 63)),
           (0x47, rs.test_reg(zte=0)), (0x00, rs.prim_reg(3, iip=1)),
           (0x01, rs.rgbaq(0x80FF0000)), (0x05, rs.xyz2(OFS + 3, OFS + 5, 100)),
           (0x01, rs.rgbaq(0x8000FF00)), (0x05, rs.xyz2(OFS + 60 * 16 + 7, OFS + 2 * 16, 5000)),
           (0x01, rs.rgbaq(0x800000FF)), (0x05, rs.xyz2(OFS + 9 * 16, OFS + 50 * 16 + 3, 90000))]
    stream = gif.arr(gif.ad(tri))
    data = np.full(8192, 0x55, np.uint8)
    o = Oracle(2)
    with pgs.Context(super_sampling=4) as ctx:
        ctx.gif_transfer(stream)
        o.gif_transfer(stream)
        ctx.write_vram(data[:5000], 100)
        o.host_write(data[:5000], 100)
        ctx.gif_transfer(stream)
        o.gif_transfer(stream)
        got = ctx.vram_planes()
    o.flush()
    assert_same(got, o.planes(), "ssaa host write")


def test_reset_context_state(hip_device):
    """reset_context_state zeroes the register file: AC = 0 afterwards, so the second sprite's
    PRIM context bit is ignored."""
    a = sprite_stream(1, None, rs.prim_reg(6, ctxt=1))
    b = gif.arr(gif.ad([(0x18, OFS | OFS << 32), (0x19, OFS | OFS << 32), (0x4C, rs.frame(0, 1, 0x00)),
                        (0x40, rs.scissor(0, 63, 0, 63)), (0x47, rs.test_reg(zte=1, ztst=1)),
                        (0x4E, rs.zbuf(32, 0, 1)), (0x00, rs.prim_reg(6, ctxt=1)), (0x01, rs.rgbaq(0x11223344)),
                        (0x05, rs.xyz2(OFS + 20 * 16, OFS + 20 * 16, 0)), (0x05, rs.xyz2(OFS + 30 * 16, OFS + 30 * 16, 0))]))
    o = Oracle()
    with pgs.Context() as ctx:
        ctx.gif_transfer(a)
        o.gif_transfer(a)
        ctx.reset_context_state()
        o.reset_context_state()
        assert ctx.register_state().prmodecont == 0
        ctx.gif_transfer(b)
        o.gif_transfer(b)
        got = ctx.vram()
    assert_same(got, o.vram(), "reset")


def packet_offsets(stream):
    """Byte offsets of the GIFtags of a stream of PACKED / IMAGE packets."""
    q = stream.view(np.uint64)
    offs, i = [], 0
    while i < q.size // 2:
        lo = int(q[2 * i])
        nloop, flg, nreg = lo & 0x7FFF, (lo >> 58) & 3, (lo >> 60) & 0xF
        offs.append(i * 16)
        i += 1 + (nloop * (nreg or 16) if flg == 0 else nloop)
    return offs


def test_save_state_round_trip(hip_device):
    """The PCSX2 freeze/defrost path (misc/arch-pkgbuild-pcsx2/0001*.patch:1663-1839): run part
    of a stream (cut inside a GIF packet), save registers, GIF paths and VRAM, restore them into
    a fresh context, clobber, continue the stream; VRAM equals the oracle's uninterrupted run."""
    stream = rs.random_render_stream(21, n_batches=6, prims_per_batch=20, width=128, height=96)
    # Cut 5 registers into the register packet of the fourth batch: the packet ends with a PRIM
    # write, so no half-assembled primitive (which a save state does not keep) spans the cut.
    q = stream.view(np.uint64)
    regs_packets = [o for o in packet_offsets(stream) if int(q[(o + 16) // 8 + 1]) & 0xFF == 0x4C]
    cut = regs_packets[3] + 16 * 6
    o = Oracle()
    o.gif_transfer(stream)
    ref = o.vram()
    with pgs.Context() as a:
        a.gif_transfer(stream[:cut])
        regs = a.register_state()
        paths = [a.gif_path(i) for i in range(4)]
        priv = np.arange(8192, dtype=np.uint64).astype(np.uint8)
        a.set_priv_registers(priv)
        priv_back = a.priv_registers()
        vram = a.vram()
    assert (priv_back == priv).all()
    print(f"[case] save-state cut={cut} path3 loop={paths[3].loop} reg={paths[3].reg}", flush=True)
    assert paths[3].loop != 0 or paths[3].reg != 0, "the cut should fall inside a packet"
    with pgs.Context() as b:
        b.write_vram(vram)
        b.set_register_state(regs)
        for i, p in enumerate(paths):
            b.set_gif_path(i, p)
        b.clobber_register_state()
        b.gif_transfer(stream[cut:])
        got = b.vram()
    assert_same(got, ref, "save state")


def test_dynamic_super_sampling_rate_change(hip_device):
    stream1 = rs.random_render_stream(31, n_batches=3, prims_per_batch=20, width=128, height=96)
    stream2 = rs.random_render_stream(32, n_batches=3, prims_per_batch=20, width=128, height=96)
    o = Oracle(0)
    with pgs.Context(super_sampling=1, dynamic_super_sampling=True) as ctx:
        ctx.gif_transfer(stream1)
        o.gif_transfer(stream1)
        assert ctx.set_super_sampling_rate(4) == 0
        o.set_sample_log2(2)
        assert ctx.super_sampling_rate == 4 and ctx.num_planes == 6
        ctx.gif_transfer(stream2)
        o.gif_transfer(stream2)
        got = ctx.vram_planes()
        assert ctx.set_super_sampling_rate(4, textures=True) == 0
        assert ctx.set_super_sampling_rate(3) == pgs.ERROR_INVALID_ARGUMENT
        assert ctx.super_sampling_rate == 4
    o.flush()
    assert_same(got, o.planes(), "rate change")
    with pgs.Context(super_sampling=2) as ctx:
        assert ctx.set_super_sampling_rate(16) == pgs.ERROR_INVALID_ARGUMENT  # beyond the allocated planes
        assert ctx.set_super_sampling_rate(1) == 0 and ctx.num_planes == 1


def test_create_rejects_unbuilt_options(hip_device):
    with pytest.raises(pgs.PgsError):
        pgs.Context(super_sampling=3)
    # Super-sampled textures are built; at 1x the option is accepted and has no effect.
    for kw in (dict(super_sampling=4, super_sampled_textures=True), dict(super_sampling=1, super_sampled_textures=True)):
        pgs.Context(**kw).close()


def test_signal_interface_and_options(hip_device):
    seen = []
    with pgs.Context() as ctx:
        ctx.set_signal_interface(on_signal=lambda v: seen.append(("signal", v)) or True,
                                 on_finish=lambda v: seen.append(("finish", v)) or False,
                                 on_label=lambda v: seen.append(("label", v)) or False)
        ctx.gif_transfer(gif.arr(gif.ad([(0x60, 0x1234), (0x61, 7), (0x62, 0xABCD00000001)])))
        assert seen == [("signal", 0x1234), ("finish", 7), ("label", 0xABCD00000001)]
        assert ctx.set_hacks(unsynced_readbacks=1) == 0
        assert ctx.set_hacks(backbuffer_promotion=1) == pgs.ERROR_UNSUPPORTED
        assert ctx.set_debug_mode(draw_mode=1) == pgs.ERROR_UNSUPPORTED
        assert ctx.set_debug_mode(timestamps=1, disable_sampler_feedback=1) == 0
        ctx.gif_transfer(rs.random_textured_stream(5, n_batches=3, prims_per_batch=20, width=128, height=96))
        ctx.flush()
        ts = ctx.accumulated_timestamps()
        print(ts, flush=True)
        assert ts["shading"] > 0 and ts["triangle_setup"] > 0 and ts["binning"] > 0
        assert ts["copy_vram"] > 0 and ts["texture_upload"] > 0
        assert all(v >= 0 for v in ts.values())


def test_recording_replays_after_later_work(hip_device):
    """Work issued after a recording (uploads, draws) must not disturb the recording: replays
    still produce the recorded result."""
    rec = rs.random_textured_stream(41, n_batches=4, prims_per_batch=20, width=128, height=96, feedback=False)
    later = gif.arr(gif.upload(300 * 32, 2, 0x00, 0, 0, 128, 64,
                               np.random.default_rng(2).integers(0, 256, 128 * 64 * 4, np.uint8).tobytes()))
    with pgs.Context() as ctx:
        ctx.begin_recording()
        ctx.gif_transfer(rec)
        ctx.end_recording()
        after_rec = ctx.vram()
        for _ in range(3):
            ctx.gif_transfer(later)
            ctx.flush()
        ctx.replay(use_graph=False)
        ctx.replay(use_graph=True)
        got = ctx.vram()
    o = Oracle()
    o.gif_transfer(rec)
    ref = o.vram()
    assert_same(after_rec, ref, "live")
    o.gif_transfer(later)
    assert_same(got, o.vram(), "replayed after later work")
