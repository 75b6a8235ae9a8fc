"""Python bindings (ctypes) for the MI355X GS backend.

The product is the C-ABI shared library ``libpgs_amd.so`` (include/pgs_amd.h and
include/pgs_kernels.h). This module only loads it and wraps the context-level calls, the
same way a ctypes integration of the reference's GSInterface would. There is no fallback:
if the library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libpgs_amd.so")
STREAMS_PATH = os.path.join(PKG_DIR, "libpgs_streams.so")

VRAM_SIZE = 4 * 1024 * 1024

PSM = {
    "CT32": 0x00, "CT24": 0x01, "CT16": 0x02, "CT16S": 0x0A,
    "T8": 0x13, "T4": 0x14, "T8H": 0x1B, "T4HL": 0x24, "T4HH": 0x2C,
    "Z32": 0x30, "Z24": 0x31, "Z16": 0x32, "Z16S": 0x3A,
}


class PgsError(RuntimeError):
    pass


class _Options(ctypes.Structure):
    _fields_ = [
        ("super_sampling", ctypes.c_uint32),
        ("ordered_super_sampling", ctypes.c_uint32),
        ("super_sampled_textures", ctypes.c_uint32),
        ("device_index", ctypes.c_uint32),
        ("upload_arena_bytes", ctypes.c_uint64),
        ("sampler_feedback", ctypes.c_uint32),
        ("texture_stream", ctypes.c_uint32),
        ("dynamic_super_sampling", ctypes.c_uint32),
    ]


ERROR_INVALID_ARGUMENT = -1
ERROR_UNSUPPORTED = -1000


class ContextRegisters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("tex0", "tex1", "clamp", "xyoffset", "miptbp1", "miptbp2", "scissor",
                                               "alpha", "test", "fba", "frame", "zbuf")]


class RegisterState(ctypes.Structure):
    _fields_ = ([("ctx", ContextRegisters * 2)] +
                [(n, ctypes.c_uint64) for n in ("prim", "rgbaq", "st", "uv", "fog", "prmodecont", "texclut", "texa",
                                                "fogcol", "dimx", "dthe", "colclamp", "pabe", "bitbltbuf", "trxpos",
                                                "trxreg", "trxdir", "scanmsk")] +
                [("internal_q", ctypes.c_float), ("cached_cbp", ctypes.c_uint32 * 2), ("pad", ctypes.c_uint32)])


class GifPath(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64 * 2), ("reg", ctypes.c_uint32), ("loop", ctypes.c_uint32)]


class Hacks(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("disable_mipmaps", "unsynced_readbacks", "backbuffer_promotion",
                                               "allow_blend_demote")]


class DebugMode(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("feedback_render_target", "timestamps", "deterministic_timeline_query",
                                               "disable_sampler_feedback", "draw_mode")]


SIGNAL_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64)


class SignalInterface(ctypes.Structure):
    _fields_ = [("on_signal", SIGNAL_CB), ("on_finish", SIGNAL_CB), ("on_label", SIGNAL_CB), ("user", ctypes.c_void_p)]


PRIV_REGISTER_BYTES = 8192
VSYNC_FIELDS = ("phase", "force_progressive", "overscan", "anti_blur", "crtc_offsets",
                "adapt_to_internal_horizontal_resolution", "raw_circuit_scanout", "high_resolution_scanout",
                "skip_deinterlace")
SCANOUT_FIELDS = ("width", "height", "internal_width", "internal_height", "mode_width", "mode_height",
                  "high_resolution_scanout", "interlaced", "double_strike", "interlace_phase")


class VSyncInfo(ctypes.Structure):
    """pgs_vsync_info."""
    _fields_ = [(n, ctypes.c_uint32) for n in VSYNC_FIELDS]


class Scanout(ctypes.Structure):
    """pgs_scanout."""
    _fields_ = [("image", ctypes.c_void_p)] + [(n, ctypes.c_uint32) for n in SCANOUT_FIELDS] + \
        [("sampling_rate_mhz", ctypes.c_float)]
TIMESTAMPS = ("sync_host_to_vram", "copy_vram", "palette_update", "texture_upload", "triangle_setup", "binning",
              "shading", "readback", "vsync")


class FlushStats(ctypes.Structure):
    _fields_ = [
        ("transfer_batches", ctypes.c_uint64),
        ("copies", ctypes.c_uint64),
        ("copy_blocks", ctypes.c_uint64),
        ("copy_records", ctypes.c_uint64),
        ("snapshot_blocks", ctypes.c_uint64),
        ("copy_bytes", ctypes.c_uint64),
        ("render_passes", ctypes.c_uint64),
        ("primitives", ctypes.c_uint64),
        ("syncs", ctypes.c_uint64),
        ("textures_decoded", ctypes.c_uint64),
        ("texture_cache_hits", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


_lib: Optional[ctypes.CDLL] = None
_streams: Optional[ctypes.CDLL] = None

# Every symbol declared in include/pgs_amd.h and include/pgs_kernels.h.
EXPORTED_SYMBOLS = [
    "pgs_create", "pgs_destroy", "pgs_reset_context_state", "pgs_gif_transfer",
    "pgs_write_register", "pgs_flush", "pgs_map_vram_read", "pgs_map_vram_write",
    "pgs_end_vram_write", "pgs_read_transfer_fifo", "pgs_clobber_register_state",
    "pgs_consume_flush_stats", "pgs_num_planes", "pgs_read_vram_planes",
    "pgs_get_register_state", "pgs_set_register_state", "pgs_get_gif_path", "pgs_set_gif_path",
    "pgs_get_priv_register_state", "pgs_set_priv_register_state", "pgs_set_super_sampling_rate",
    "pgs_get_super_sampling_rate", "pgs_set_hacks", "pgs_set_debug_mode", "pgs_set_signal_interface",
    "pgs_get_accumulated_timestamps", "pgs_vsync", "pgs_vsync_can_skip", "pgs_read_scanout",
    "pgs_dump_open", "pgs_dump_restart", "pgs_dump_iterate_until_vsync", "pgs_dump_consume_vsync_result",
    "pgs_dump_close", "pgs_dump_writer_create", "pgs_dump_write_register", "pgs_dump_write_packed",
    "pgs_dump_write_vsync", "pgs_dump_write_image_upload", "pgs_dump_writer_close",
    "pgs_begin_recording", "pgs_end_recording", "pgs_replay", "pgs_sync", "pgs_stream",
    "pgs_profile_enable", "pgs_profile_collect",
    "pgs_host_swizzle_addr", "pgs_host_block_lane_pixel",
    "pgs_k_vram_copy", "pgs_k_vram_snapshot", "pgs_k_init_tables", "pgs_k_triangle_setup",
    "pgs_k_binning", "pgs_k_shade", "pgs_phase_lut", "pgs_k_clut_write", "pgs_k_texture_decode", "pgs_k_ssaa_heuristic",
    "pgs_batch_create", "pgs_batch_destroy", "pgs_batch_apply_tuning", "pgs_batch_autotune", "pgs_batch_replay",
    "pgs_batch_sync", "pgs_batch_stream", "pgs_batch_num_launches", "pgs_batch_profile",
    "pgs_cost_model_create", "pgs_cost_model_destroy", "pgs_cost_model_set_params", "pgs_cost_model_get_params",
    "pgs_cost_model_predict", "pgs_cost_model_gradients", "pgs_cost_model_train",
    "pgs_k_mlp_forward", "pgs_k_mlp_gradients", "pgs_k_mlp_gradients_noalloc", "pgs_k_mlp_adam", "pgs_k_mlp_quantize",
    "pgs_batch_set_schedule", "pgs_batch_collect_costs", "pgs_batch_schedule_stats", "pgs_k_shade_queue", "pgs_k_tile_features",
    "pgs_k_queue_order", "pgs_k_fill_u32", "pgs_k_shade_ordered", "pgs_k_sampler_feedback", "pgs_k_restore", "pgs_k_fused", "pgs_k_rcp_eval", "pgs_k_lod_eval",
    "pgs_k_scanout_circuit", "pgs_k_scanout_merge", "pgs_k_scanout_extwrite", "pgs_k_scanout_weave",
]


def lib() -> ctypes.CDLL:
    """Loads libpgs_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PgsError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, u32, u64, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.pgs_create.restype = vp
        L.pgs_create.argtypes = [ctypes.POINTER(_Options)]
        L.pgs_destroy.argtypes = [vp]
        L.pgs_destroy.restype = None
        L.pgs_reset_context_state.argtypes = [vp]
        L.pgs_gif_transfer.argtypes = [vp, u32, vp, sz]
        L.pgs_write_register.argtypes = [vp, u32, u64]
        L.pgs_flush.argtypes = [vp]
        L.pgs_map_vram_read.argtypes = [vp, sz, sz]
        L.pgs_map_vram_read.restype = vp
        L.pgs_map_vram_write.argtypes = [vp, sz, sz]
        L.pgs_map_vram_write.restype = vp
        L.pgs_end_vram_write.argtypes = [vp, sz, sz]
        L.pgs_read_transfer_fifo.argtypes = [vp, vp, u32]
        L.pgs_clobber_register_state.argtypes = [vp]
        L.pgs_consume_flush_stats.argtypes = [vp, ctypes.POINTER(FlushStats)]
        L.pgs_num_planes.argtypes = [vp]
        L.pgs_num_planes.restype = u32
        L.pgs_read_vram_planes.argtypes = [vp, vp]
        L.pgs_begin_recording.argtypes = [vp]
        L.pgs_end_recording.argtypes = [vp]
        L.pgs_replay.argtypes = [vp, i32]
        L.pgs_sync.argtypes = [vp]
        L.pgs_profile_enable.argtypes = [vp, i32]
        L.pgs_profile_collect.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64), u32]
        L.pgs_stream.argtypes = [vp]
        L.pgs_stream.restype = vp
        L.pgs_host_swizzle_addr.argtypes = [u32, u32, u32, u32, u32]
        L.pgs_host_swizzle_addr.restype = u32
        L.pgs_host_block_lane_pixel.argtypes = [u32, u32, u32, ctypes.POINTER(u32), ctypes.POINTER(u32)]
        L.pgs_host_block_lane_pixel.restype = None
        L.pgs_get_register_state.argtypes = [vp, ctypes.POINTER(RegisterState)]
        L.pgs_set_register_state.argtypes = [vp, ctypes.POINTER(RegisterState)]
        L.pgs_get_gif_path.argtypes = [vp, u32, ctypes.POINTER(GifPath)]
        L.pgs_set_gif_path.argtypes = [vp, u32, ctypes.POINTER(GifPath)]
        L.pgs_get_priv_register_state.argtypes = [vp, vp, sz]
        L.pgs_set_priv_register_state.argtypes = [vp, vp, sz]
        L.pgs_set_super_sampling_rate.argtypes = [vp, u32, u32, u32]
        L.pgs_get_super_sampling_rate.argtypes = [vp]
        L.pgs_get_super_sampling_rate.restype = u32
        L.pgs_set_hacks.argtypes = [vp, ctypes.POINTER(Hacks)]
        L.pgs_set_debug_mode.argtypes = [vp, ctypes.POINTER(DebugMode)]
        L.pgs_set_signal_interface.argtypes = [vp, ctypes.POINTER(SignalInterface)]
        L.pgs_get_accumulated_timestamps.argtypes = [vp, u32]
        L.pgs_get_accumulated_timestamps.restype = ctypes.c_double
        L.pgs_vsync.argtypes = [vp, ctypes.POINTER(VSyncInfo), ctypes.POINTER(Scanout)]
        L.pgs_vsync_can_skip.argtypes = [vp, ctypes.POINTER(VSyncInfo)]
        L.pgs_read_scanout.argtypes = [vp, vp, sz]
        cp = ctypes.c_char_p
        L.pgs_dump_open.argtypes = [vp, cp, u32]
        L.pgs_dump_open.restype = vp
        L.pgs_dump_restart.argtypes = [vp]
        L.pgs_dump_iterate_until_vsync.argtypes = [vp, u32, u32]
        L.pgs_dump_consume_vsync_result.argtypes = [vp, ctypes.POINTER(Scanout)]
        L.pgs_dump_close.argtypes = [vp]
        L.pgs_dump_close.restype = None
        L.pgs_dump_writer_create.argtypes = [vp, cp]
        L.pgs_dump_writer_create.restype = vp
        L.pgs_dump_write_register.argtypes = [vp, u32, u64]
        L.pgs_dump_write_packed.argtypes = [vp, u64, ctypes.POINTER(u32), u32, u32, vp]
        L.pgs_dump_write_vsync.argtypes = [vp, u32]
        L.pgs_dump_write_image_upload.argtypes = [vp, u32, u32, u32, u32, vp, sz]
        L.pgs_dump_writer_close.argtypes = [vp]
        _lib = L
    return _lib


def phase_lut(rate_x: int, rate_y: int) -> np.ndarray:
    """pgs_phase_lut: the phase filter table of super-sampled textures for a 2^rate_x x 2^rate_y
    sampling grid (a host function: no device needed)."""
    out = np.zeros(2 * (64 * 64 + 32 * 32), np.uint32)
    L = lib()
    L.pgs_phase_lut.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    L.pgs_phase_lut.restype = None
    L.pgs_phase_lut(rate_x, rate_y, out.ctypes.data)
    return out


def streams() -> ctypes.CDLL:
    """Loads the synthetic stream generator library."""
    global _streams
    if _streams is None:
        if not os.path.exists(STREAMS_PATH):
            raise PgsError(f"{STREAMS_PATH} is missing: run __graft_entry__.build() first")
        S = ctypes.CDLL(STREAMS_PATH)
        S.pgs_gen_transfers.restype = ctypes.c_void_p
        S.pgs_gen_transfers.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_size_t)]
        S.pgs_gen_plain.restype = ctypes.c_void_p
        S.pgs_gen_plain.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.POINTER(ctypes.c_size_t)]
        S.pgs_gen_textured.restype = ctypes.c_void_p
        S.pgs_gen_textured.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_size_t)]
        S.pgs_gen_transfers_wild.restype = ctypes.c_void_p
        S.pgs_gen_transfers_wild.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_size_t)]
        _streams = S
    return _streams


def _copy_stream(ptr, size) -> np.ndarray:
    buf = (ctypes.c_uint8 * size.value).from_address(ptr)
    return np.frombuffer(buf, dtype=np.uint8).copy()


def gen_transfers(seed: int, n_h2l: int, n_l2l: int, max_dim: int = 512) -> np.ndarray:
    """GIF stream of the VRAM transfer microbenchmark (host->local + local->local)."""
    size = ctypes.c_size_t()
    ptr = streams().pgs_gen_transfers(seed, n_h2l, n_l2l, max_dim, ctypes.byref(size))
    return _copy_stream(ptr, size)


def gen_plain(seed: int, n_tris: int = 2000, width: int = 640, height: int = 448) -> np.ndarray:
    """GIF stream of the plain 1x rendering workload (untextured Gouraud triangles)."""
    size = ctypes.c_size_t()
    ptr = streams().pgs_gen_plain(seed, n_tris, width, height, ctypes.byref(size))
    return _copy_stream(ptr, size)


def gen_textured(seed: int, n_prims: int, frames: int, width: int = 640, height: int = 448,
                 pareto_alpha: float = 0.0, tex_psm: int = 0x13, blend_preset: int = 0) -> np.ndarray:
    """Textured + blended stream: per frame a 256x256 texture (PSMCT32, PSMT8 or PSMT4; the
    palette formats with a PSMCT32 CLUT loaded by CLD=1), n_prims primitives (70% triangles,
    30% sprites), bilinear MODULATE, a blend preset (0: (Cs-Cd)*As+Cd, 1: Cs*As+Cd,
    2: Cd-Cs*As, 3: (Cs-Cd)*64+Cd), GEQUAL depth, 20% ST/Q. pareto_alpha > 0 draws primitive
    areas from Pareto(alpha)."""
    size = ctypes.c_size_t(0)
    ptr = streams().pgs_gen_textured(seed, n_prims, frames, width, height, pareto_alpha, tex_psm, blend_preset,
                                     ctypes.byref(size))
    return _copy_stream(ptr, size)


def gen_transfers_wild(seed: int, n: int) -> np.ndarray:
    """GIF stream of small transfers with arbitrary geometry (edge cases)."""
    size = ctypes.c_size_t()
    ptr = streams().pgs_gen_transfers_wild(seed, n, ctypes.byref(size))
    return _copy_stream(ptr, size)


class Context:
    """One GS context (GSInterface) on a HIP device."""

    SAMPLER_FEEDBACK = {"auto": 0, "always": 1, "never": 2}

    def __init__(self, super_sampling: int = 1, device_index: int = 0, arena_bytes: int = 0,
                 ordered: bool = True, super_sampled_textures: bool = False, sampler_feedback: str = "auto",
                 texture_stream: bool = False, dynamic_super_sampling: bool = False):
        """texture_stream=True runs VRAM copies, CLUT loads and texture decodes on a second HIP
        stream beside setup and binning; by default every launch goes to one stream."""
        opts = _Options(super_sampling, int(ordered), int(super_sampled_textures), device_index, arena_bytes,
                        self.SAMPLER_FEEDBACK[sampler_feedback], int(texture_stream), int(dynamic_super_sampling))
        self._L = lib()
        self._h = self._L.pgs_create(ctypes.byref(opts))
        if not self._h:
            raise PgsError("pgs_create failed (no HIP device or out of memory)")

    def close(self):
        if self._h:
            self._L.pgs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise PgsError(f"{what} failed with {rc}")

    def gif_transfer(self, data: np.ndarray, path: int = 3):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        self._check(self._L.pgs_gif_transfer(self._h, path, data.ctypes.data, data.nbytes), "gif_transfer")

    def write_register(self, addr: int, value: int):
        self._check(self._L.pgs_write_register(self._h, addr, value & ((1 << 64) - 1)), "write_register")

    def flush(self):
        self._check(self._L.pgs_flush(self._h), "flush")

    def reset_context_state(self):
        self._check(self._L.pgs_reset_context_state(self._h), "reset_context_state")

    def clobber_register_state(self):
        self._check(self._L.pgs_clobber_register_state(self._h), "clobber_register_state")

    def sync(self):
        self._check(self._L.pgs_sync(self._h), "sync")

    @property
    def num_planes(self) -> int:
        return int(self._L.pgs_num_planes(self._h))

    def vram(self) -> np.ndarray:
        """Plane 0 (the 4 MiB GS local memory) as uint8."""
        p = self._L.pgs_map_vram_read(self._h, 0, VRAM_SIZE)
        if not p:
            raise PgsError("map_vram_read failed")
        return np.frombuffer((ctypes.c_uint8 * VRAM_SIZE).from_address(p), dtype=np.uint8).copy()

    def vram_planes(self) -> np.ndarray:
        out = np.empty(self.num_planes * VRAM_SIZE, dtype=np.uint8)
        self._check(self._L.pgs_read_vram_planes(self._h, out.ctypes.data), "read_vram_planes")
        return out

    def write_vram(self, data: np.ndarray, offset: int = 0):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        p = self._L.pgs_map_vram_write(self._h, offset, data.nbytes)
        if not p:
            raise PgsError("map_vram_write failed")
        ctypes.memmove(p, data.ctypes.data, data.nbytes)
        self._check(self._L.pgs_end_vram_write(self._h, offset, data.nbytes), "end_vram_write")

    def read_transfer_fifo(self, n128: int) -> np.ndarray:
        out = np.zeros(n128 * 16, dtype=np.uint8)
        self._check(self._L.pgs_read_transfer_fifo(self._h, out.ctypes.data, n128), "read_transfer_fifo")
        return out

    def consume_flush_stats(self) -> dict:
        s = FlushStats()
        self._check(self._L.pgs_consume_flush_stats(self._h, ctypes.byref(s)), "consume_flush_stats")
        return s.as_dict()

    def begin_recording(self):
        self._check(self._L.pgs_begin_recording(self._h), "begin_recording")

    def end_recording(self):
        self._check(self._L.pgs_end_recording(self._h), "end_recording")

    def replay(self, use_graph: bool = True):
        self._check(self._L.pgs_replay(self._h, int(use_graph)), "replay")

    def profile_enable(self, enable: bool = True):
        self._check(self._L.pgs_profile_enable(self._h, int(enable)), "profile_enable")

    def profile_collect(self, kinds: int = 16):
        ms = (ctypes.c_double * kinds)()
        n = (ctypes.c_uint64 * kinds)()
        self._check(self._L.pgs_profile_collect(self._h, ms, n, kinds), "profile_collect")
        return list(ms), list(n)

    @property
    def stream(self) -> int:
        return int(self._L.pgs_stream(self._h) or 0)

    def register_state(self) -> RegisterState:
        r = RegisterState()
        self._check(self._L.pgs_get_register_state(self._h, ctypes.byref(r)), "get_register_state")
        return r

    def set_register_state(self, r: RegisterState):
        self._check(self._L.pgs_set_register_state(self._h, ctypes.byref(r)), "set_register_state")

    def gif_path(self, path: int) -> GifPath:
        g = GifPath()
        self._check(self._L.pgs_get_gif_path(self._h, path, ctypes.byref(g)), "get_gif_path")
        return g

    def set_gif_path(self, path: int, g: GifPath):
        self._check(self._L.pgs_set_gif_path(self._h, path, ctypes.byref(g)), "set_gif_path")

    def priv_registers(self) -> np.ndarray:
        out = np.zeros(PRIV_REGISTER_BYTES, np.uint8)
        self._check(self._L.pgs_get_priv_register_state(self._h, out.ctypes.data, out.nbytes), "get_priv")
        return out

    def set_priv_registers(self, data: np.ndarray):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        self._check(self._L.pgs_set_priv_register_state(self._h, data.ctypes.data, data.nbytes), "set_priv")

    def set_super_sampling_rate(self, samples: int, ordered: bool = True, textures: bool = False) -> int:
        """Returns the C return code (0, ERROR_UNSUPPORTED or ERROR_INVALID_ARGUMENT)."""
        return int(self._L.pgs_set_super_sampling_rate(self._h, samples, int(ordered), int(textures)))

    @property
    def super_sampling_rate(self) -> int:
        return int(self._L.pgs_get_super_sampling_rate(self._h))

    def set_hacks(self, **kw) -> int:
        h = Hacks(**{k: int(v) for k, v in kw.items()})
        return int(self._L.pgs_set_hacks(self._h, ctypes.byref(h)))

    def set_debug_mode(self, **kw) -> int:
        d = DebugMode(**{k: int(v) for k, v in kw.items()})
        return int(self._L.pgs_set_debug_mode(self._h, ctypes.byref(d)))

    def set_signal_interface(self, on_signal=None, on_finish=None, on_label=None):
        """Each callback takes the 64-bit payload and returns True to flush."""
        def wrap(f):
            return SIGNAL_CB(lambda _user, v: int(bool(f(int(v))))) if f else SIGNAL_CB()
        self._signal = SignalInterface(wrap(on_signal), wrap(on_finish), wrap(on_label), None)
        self._check(self._L.pgs_set_signal_interface(self._h, ctypes.byref(self._signal)), "set_signal_interface")

    def accumulated_timestamps(self) -> dict:
        return {name: float(self._L.pgs_get_accumulated_timestamps(self._h, i)) for i, name in enumerate(TIMESTAMPS)}

    @staticmethod
    def _vsync_info(info: dict) -> VSyncInfo:
        bad = set(info) - set(VSYNC_FIELDS)
        if bad:
            raise TypeError(f"unknown vsync fields {bad}")
        return VSyncInfo(**{k: int(v) for k, v in info.items()})

    def vsync(self, read: bool = True, **info):
        """pgs_vsync. Returns (status, metadata dict, image): the image is the frame as a
        (height, width) uint32 RGBA8 array copied to the host when `read`, else the device
        pointer; None when there is no frame."""
        out = Scanout()
        rc = int(self._L.pgs_vsync(self._h, ctypes.byref(self._vsync_info(info)), ctypes.byref(out)))
        return self._scanout_result(rc, out, read)

    def _scanout_result(self, rc: int, out: "Scanout", read: bool):
        meta = {n: int(getattr(out, n)) for n in SCANOUT_FIELDS}
        meta["sampling_rate_mhz"] = float(out.sampling_rate_mhz)
        if rc != 0 or not out.image:
            return rc, meta, None
        if not read:
            return rc, meta, int(out.image)
        img = np.zeros((out.height, out.width), np.uint32)
        self._check(self._L.pgs_read_scanout(self._h, img.ctypes.data, img.nbytes), "read_scanout")
        return rc, meta, img

    def vsync_can_skip(self, **info) -> bool:
        rc = int(self._L.pgs_vsync_can_skip(self._h, ctypes.byref(self._vsync_info(info))))
        if rc < 0:
            raise PgsError(f"vsync_can_skip failed: {rc}")
        return rc == 1


class DumpReader:
    """GS dump reader bound to a context (pgs_dump_open)."""

    def __init__(self, ctx: "Context", path: str, raw: bool = False):
        self._L, self._ctx = ctx._L, ctx
        self._h = self._L.pgs_dump_open(ctx._h, os.fsencode(path), int(raw))
        if not self._h:
            raise PgsError(f"cannot open GS dump {path}")

    def restart(self) -> bool:
        return self._L.pgs_dump_restart(self._h) == 0

    def iterate_until_vsync(self, high_res_scanout: bool = False, conservative_crtc: bool = False) -> bool:
        rc = int(self._L.pgs_dump_iterate_until_vsync(self._h, int(high_res_scanout), int(conservative_crtc)))
        if rc < 0:
            raise PgsError(f"dump replay failed: {rc}")
        return rc == 1

    def consume_vsync_result(self, read: bool = True):
        """(metadata, image) of the last vsync; image None when there was none."""
        out = Scanout()
        rc = int(self._L.pgs_dump_consume_vsync_result(self._h, ctypes.byref(out)))
        _, meta, img = self._ctx._scanout_result(rc, out, read)
        return meta, img

    def close(self):
        if getattr(self, "_h", None):
            self._L.pgs_dump_close(self._h)
            self._h = None

    def __del__(self):
        self.close()


class DumpWriter:
    """GS dump writer bound to a context (pgs_dump_writer_create): the header holds the
    context's state at creation."""

    def __init__(self, ctx: "Context", path: str):
        self._L = ctx._L
        self._h = self._L.pgs_dump_writer_create(ctx._h, os.fsencode(path))
        if not self._h:
            raise PgsError(f"cannot create GS dump {path}")

    def _ok(self, rc, what):
        if rc != 0:
            raise PgsError(f"dump {what} failed: {rc}")

    def write_register(self, addr: int, payload: int):
        self._ok(self._L.pgs_dump_write_register(self._h, addr, payload & ((1 << 64) - 1)), "write_register")

    def write_packed(self, prim: int, registers, num_loops: int, data: np.ndarray):
        regs = (ctypes.c_uint32 * len(registers))(*registers)
        data = np.ascontiguousarray(data, dtype=np.uint8)
        self._ok(self._L.pgs_dump_write_packed(self._h, prim, regs, len(registers), num_loops, data.ctypes.data),
                 "write_packed")

    def write_vsync(self, field: int):
        self._ok(self._L.pgs_dump_write_vsync(self._h, field), "write_vsync")

    def write_image_upload(self, addr: int, psm: int, width: int, height: int, data: np.ndarray):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        self._ok(self._L.pgs_dump_write_image_upload(self._h, addr, psm, width, height, data.ctypes.data,
                                                     data.nbytes), "write_image_upload")

    def close(self):
        if getattr(self, "_h", None):
            rc = self._L.pgs_dump_writer_close(self._h)
            self._h = None
            self._ok(rc, "close")

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.pgs_dump_writer_close(self._h)
            self._h = None


def host_swizzle_addr(x: int, y: int, base: int, stride: int, psm: int) -> int:
    return int(lib().pgs_host_swizzle_addr(x, y, base, stride, psm))


def host_block_lane_pixel(psm: int, lane: int, k: int):
    dx, dy = ctypes.c_uint32(), ctypes.c_uint32()
    lib().pgs_host_block_lane_pixel(psm, lane, k, ctypes.byref(dx), ctypes.byref(dy))
    return int(dx.value), int(dy.value)


class _Tuning(ctypes.Structure):
    _fields_ = [("tile_log2_1x", ctypes.c_uint32), ("waves_per_group", ctypes.c_uint32)]


class _TuningResult(ctypes.Structure):
    _fields_ = [("tuning", _Tuning), ("ms_per_replay", ctypes.c_float)]


class Batch:
    """Batched replay of several contexts' recordings (merged multi-context launches)."""

    def __init__(self, contexts, sequential: bool = False):
        L = lib()
        L.pgs_batch_create.restype = ctypes.c_void_p
        L.pgs_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint32, ctypes.c_uint32]
        L.pgs_batch_destroy.argtypes = [ctypes.c_void_p]
        L.pgs_batch_apply_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Tuning)]
        L.pgs_batch_autotune.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(_TuningResult), ctypes.c_uint32]
        L.pgs_batch_replay.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.pgs_batch_sync.argtypes = [ctypes.c_void_p]
        L.pgs_batch_stream.restype = ctypes.c_void_p
        L.pgs_batch_stream.argtypes = [ctypes.c_void_p]
        L.pgs_batch_num_launches.restype = ctypes.c_uint32
        L.pgs_batch_num_launches.argtypes = [ctypes.c_void_p]
        self._L = L
        self._contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self._contexts))(*[c._h for c in self._contexts])
        self._h = L.pgs_batch_create(arr, len(self._contexts), 1 if sequential else 0)
        if not self._h:
            raise PgsError("pgs_batch_create failed")

    def close(self):
        if self._h:
            self._L.pgs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def replay(self, use_graph: bool = True):
        rc = self._L.pgs_batch_replay(self._h, int(use_graph))
        if rc != 0:
            raise PgsError(f"pgs_batch_replay failed with {rc}")

    def sync(self):
        if self._L.pgs_batch_sync(self._h) != 0:
            raise PgsError("pgs_batch_sync failed")

    def apply_tuning(self, tile_log2_1x: int, waves_per_group: int):
        t = _Tuning(tile_log2_1x, waves_per_group)
        if self._L.pgs_batch_apply_tuning(self._h, ctypes.byref(t)) != 0:
            raise PgsError("pgs_batch_apply_tuning failed")

    def autotune(self, reps: int = 3):
        res = (_TuningResult * 64)()
        n = self._L.pgs_batch_autotune(self._h, reps, res, 64)
        if n < 0:
            raise PgsError("pgs_batch_autotune failed")
        return [(r.tuning.tile_log2_1x, r.tuning.waves_per_group, r.ms_per_replay) for r in res[:n]]

    @property
    def stream(self) -> int:
        return int(self._L.pgs_batch_stream(self._h) or 0)

    @property
    def num_launches(self) -> int:
        return int(self._L.pgs_batch_num_launches(self._h))

    # "static" / "learned": the persistent work queue; "dispatch_static" / "dispatch_learned":
    # the hardware dispatcher started in tile order or in predicted-cost order.
    SCHEDULES = {"hardware": 0, "static": 1, "learned": 2, "dispatch_static": 3, "dispatch_learned": 4}

    def set_schedule(self, name: str, model=None):
        self._L.pgs_batch_set_schedule.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        self._model = model  # kept alive with the batch
        if self._L.pgs_batch_set_schedule(self._h, self.SCHEDULES[name], model._h if model else None) != 0:
            raise PgsError("pgs_batch_set_schedule failed")

    def schedule_stats(self):
        """Contents of the current schedule's order after its planning replay (dict), or None
        under the hardware schedule / before planning."""
        class _S(ctypes.Structure):
            _fields_ = [(n, ctypes.c_uint32) for n in ("shading_launches", "tiles", "entries", "split_tiles",
                                                        "heavy_first")]
        self._L.pgs_batch_schedule_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(_S)]
        s = _S()
        if self._L.pgs_batch_schedule_stats(self._h, ctypes.byref(s)) != 0:
            return None
        return {n: int(getattr(s, n)) for n, _ in _S._fields_}

    def collect_costs(self, max_items: int = 1 << 22):
        fp = ctypes.POINTER(ctypes.c_float)
        self._L.pgs_batch_collect_costs.argtypes = [ctypes.c_void_p, fp, fp, ctypes.c_uint32]
        feats = np.zeros((max_items, MLP_INPUTS), np.float32)
        costs = np.zeros(max_items, np.float32)
        n = self._L.pgs_batch_collect_costs(self._h, feats.ctypes.data_as(fp), costs.ctypes.data_as(fp), max_items)
        if n < 0:
            raise PgsError("pgs_batch_collect_costs failed")
        return feats[:n].copy(), costs[:n].copy()

    def profile(self, kinds: int = 16):
        """Per-kind (ms, launches) of one event-bracketed replay."""
        self._L.pgs_batch_profile.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32]
        ms = (ctypes.c_double * kinds)()
        n = (ctypes.c_uint64 * kinds)()
        if self._L.pgs_batch_profile(self._h, ms, n, kinds) != 0:
            raise PgsError("pgs_batch_profile failed")
        return list(ms), list(n)


def batched_specs(seed: int, n_contexts: int, n_min: int = 500, n_max: int = 50000, ssaa_choices=(1, 2, 4),
                  pareto_alpha: float = 0.0, tex_psms=(0x00, 0x13, 0x14), blend_presets=(0, 1, 2, 3)):
    """Per-context workload choices of the batched configurations: primitive count log-uniform
    in [n_min, n_max], frame 512x448 or 640x448, a sampling rate, a texture format and a blend
    preset, each drawn from a seeded numpy generator. Returns a list of dicts."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_contexts):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        specs.append(dict(seed=int(seed) * 1000 + i, n_prims=n, width=int(rng.choice([512, 640])), height=448,
                          ssaa=int(rng.choice(ssaa_choices)), tex_psm=int(rng.choice(tex_psms)),
                          blend=int(rng.choice(blend_presets)), pareto_alpha=float(pareto_alpha)))
    return specs


def spec_stream(spec) -> np.ndarray:
    """The GIF stream of one batched-context spec (one frame)."""
    return gen_textured(spec["seed"], spec["n_prims"], 1, spec["width"], spec["height"], spec["pareto_alpha"],
                        spec["tex_psm"], spec["blend"])


MLP_INPUTS, MLP_PARAMS = 24, 5825


class CostModel:
    """The learned tile-cost MLP (parameters on the GPU; training runs on the GPU)."""

    def __init__(self, seed: int = 1):
        L = lib()
        fp = ctypes.POINTER(ctypes.c_float)
        L.pgs_cost_model_create.restype = ctypes.c_void_p
        L.pgs_cost_model_create.argtypes = [ctypes.c_uint64]
        L.pgs_cost_model_destroy.argtypes = [ctypes.c_void_p]
        L.pgs_cost_model_set_params.argtypes = [ctypes.c_void_p, fp]
        L.pgs_cost_model_get_params.argtypes = [ctypes.c_void_p, fp]
        L.pgs_cost_model_predict.argtypes = [ctypes.c_void_p, fp, ctypes.c_uint32, fp]
        L.pgs_cost_model_gradients.argtypes = [ctypes.c_void_p, fp, fp, ctypes.c_uint32, fp]
        L.pgs_cost_model_train.argtypes = [ctypes.c_void_p, fp, fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_float, ctypes.c_uint64, fp]
        self._L = L
        self._h = L.pgs_cost_model_create(seed)
        if not self._h:
            raise PgsError("pgs_cost_model_create failed")

    @staticmethod
    def _f(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def close(self):
        if self._h:
            self._L.pgs_cost_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        a, ptr = self._f(p)
        assert a.size == MLP_PARAMS
        if self._L.pgs_cost_model_set_params(self._h, ptr) != 0:
            raise PgsError("set_params failed")

    def get_params(self):
        out = np.zeros(MLP_PARAMS, np.float32)
        if self._L.pgs_cost_model_get_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) != 0:
            raise PgsError("get_params failed")
        return out

    def predict(self, X):
        x, xp = self._f(X)
        n = x.shape[0]
        out = np.zeros(n, np.float32)
        if self._L.pgs_cost_model_predict(self._h, xp, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) != 0:
            raise PgsError("predict failed")
        return out

    def gradients(self, X, t):
        x, xp = self._f(X)
        tt, tp = self._f(t)
        g = np.zeros(MLP_PARAMS, np.float32)
        if self._L.pgs_cost_model_gradients(self._h, xp, tp, x.shape[0], g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) != 0:
            raise PgsError("gradients failed")
        return g

    def train(self, X, t, epochs=10, batch=1024, lr=1e-3, seed=7):
        x, xp = self._f(X)
        tt, tp = self._f(t)
        mse = ctypes.c_float(0.0)
        if self._L.pgs_cost_model_train(self._h, xp, tp, x.shape[0], epochs, batch, lr, seed, ctypes.byref(mse)) != 0:
            raise PgsError("train failed")
        return float(mse.value)
