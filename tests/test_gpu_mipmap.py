"""GPU parity of mipmapped sampling (ubershader.comp:1066-1258, gs_renderer.cpp:3458-3598):

* the shading kernel's Q-derived level of detail is bit-identical to the oracle's over every
  mantissa of several exponent ranges (zero / subnormal, [1, 2), inf / NaN, negative q) for every
  L (pgs_k_lod_eval against orc_lod_sweep, the same hash on both sides);
* random streams with mipmapped filters, LCM / L / K, MIPTBP levels and spread Q match the oracle
  byte for byte at 1x and super-sampled, with sampler feedback forced on, and in a batch replay;
* the closed-form level tests of test_mipmap_oracle.py hold on the GPU (VRAM equal to the
  oracle's);
* Hacks::disable_mipmaps reads level 0 at LOD 0.
"""
import ctypes
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import pgs
import render_streams as rs
import test_mipmap_oracle as T
from oracle_lib import Oracle, oracle_lib, swizzle

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


def device_lod_hash(l, k, first, count):
    import torch

    L = pgs.lib()
    L.pgs_k_lod_eval.argtypes = [ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p,
                                 ctypes.c_void_p]
    h = torch.zeros(1, dtype=torch.int64, device="cuda")
    rc = L.pgs_k_lod_eval(l, k, first, count, h.data_ptr(), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    return int(h.item()) % 2**64


def oracle_lod_hash(l, k, first, count, threads=16):
    L = oracle_lib()
    f = L.orc_lod_sweep
    f.argtypes = [ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    f.restype = None
    step = (count + threads - 1) // threads

    def part(i):
        a = i * step
        n = max(0, min(step, count - a))
        h = ctypes.c_uint64()
        f(l, k, (first + a) & 0xFFFFFFFF, n, ctypes.byref(h))
        return h.value

    with ThreadPoolExecutor(threads) as ex:
        return sum(ex.map(part, range(threads))) % 2**64


@pytest.mark.parametrize("first", [0x00000000, 0x3F800000, 0x3C000000, 0x47000000, 0x7F800000, 0xBF800000])
def test_lod_bit_identical_over_mantissas(hip_device, first):
    for l in range(4):
        for k in (0, -100, 1000):
            print(f"[case] lod first={first:#x} L={l} K={k}", flush=True)
            got = device_lod_hash(l, k, first, 1 << 23)
            ref = oracle_lod_hash(l, k, first, 1 << 23)
            assert got == ref, (hex(first), l, k)


def render_both(stream, ssaa=1, feedback="auto"):
    print(f"[case] mip stream {len(stream)} bytes ssaa={ssaa} feedback={feedback}", flush=True)
    with pgs.Context(super_sampling=ssaa, sampler_feedback=feedback) as ctx:
        ctx.gif_transfer(stream)
        got = ctx.vram_planes()
    o = Oracle(LOG2[ssaa])
    o.gif_transfer(stream)
    o.flush()
    ref = o.planes()
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


@pytest.mark.parametrize("seed,ssaa", [(0, 1), (1, 1), (2, 1), (3, 4), (4, 16), (5, 2)])
def test_random_mipmapped_streams(hip_device, seed, ssaa):
    render_both(rs.random_textured_stream(700 + seed, n_batches=12, prims_per_batch=30, mipmap=True), ssaa)


def test_mipmapped_streams_with_sampler_feedback(hip_device):
    render_both(rs.random_textured_stream(710, n_batches=10, prims_per_batch=30, mipmap=True), 1, "always")


def test_mipmapped_batch_replay(hip_device):
    ctxs, live = [], []
    for i, ssaa in enumerate((1, 4, 1, 16)):
        stream = rs.random_textured_stream(720 + i, n_batches=8, prims_per_batch=30, mipmap=True)
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


def test_closed_form_on_gpu(hip_device, monkeypatch):
    captured = []

    class Mirrored(Oracle):
        def gif_transfer(self, data, path=3):
            captured.append(np.array(data, np.uint8))
            super().gif_transfer(data, path)

        def vram(self):
            ref = super().vram()
            with pgs.Context() as ctx:
                for c in captured:
                    ctx.gif_transfer(c)
                got = ctx.vram()
            captured.clear()
            assert np.array_equal(got, ref)
            return ref

    monkeypatch.setattr(T, "Oracle", Mirrored)
    for fn in (T.test_fixed_lod_selects_level, T.test_without_a_mipmapped_filter_level_zero_is_read,
               T.test_q_lod_selects_level, T.test_nearest_level_rounds_half_up, T.test_trilinear_blends_two_levels,
               T.test_lod_past_the_last_image_level_reads_the_last, T.test_min_and_mag_filters_follow_the_lod):
        print(f"[case] {fn.__name__}", flush=True)
        fn()


def test_disable_mipmaps_hack_reads_level_zero(hip_device):
    import gif

    out = [gif.ad([(0x1A, 1), (0x18, T.OFS | T.OFS << 32)])] + T.mip_setup()
    regs = [(0x4C, rs.frame(0, 1, 0x00)), (0x40, rs.scissor(0, 63, 0, 63)), (0x47, rs.test_reg(zte=1, ztst=1)),
            (0x14, rs.tex1(0, 2, lcm=1, mxl=4, k=48)), (0x08, 0),
            (0x34, rs.miptbp([(T.LEVEL_TBP[i], 1) for i in (1, 2, 3)])),
            (0x06, rs.tex0(T.LEVEL_TBP[0], 1, 0x00, 4, 4, tcc=1, tfx=1)),
            (0x00, rs.prim_reg(6, iip=0, tme=1, fst=1)),
            (0x03, rs.uv(8, 8)), (0x05, rs.xyz2(T.OFS, T.OFS, 0)),
            (0x03, rs.uv(16 * 16, 16 * 16)), (0x05, rs.xyz2(T.OFS + 16 * 16, T.OFS + 16 * 16, 0))]
    out.append(gif.ad(regs))
    stream = gif.arr(b"".join(out))
    for disable, colour in ((False, T.LEVEL_COLOURS[3]), (True, T.LEVEL_COLOURS[0])):
        with pgs.Context() as ctx:
            assert ctx.set_hacks(disable_mipmaps=disable) == 0
            ctx.gif_transfer(stream)
            v = ctx.vram()
        a = swizzle(5, 3, 0, 1, 0x00)
        got = int(v[4 * a:4 * a + 4].view(np.uint32)[0])
        assert got == colour, (disable, hex(got))
