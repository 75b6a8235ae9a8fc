"""ctypes loader for the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker the HIP path is compared against.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
VRAM_SIZE = 4 * 1024 * 1024

_L = None

VSYNC_FIELDS = ("phase", "force_progressive", "overscan", "anti_blur", "crtc_offsets",
                "adapt_to_internal_horizontal_resolution", "raw_circuit_scanout", "high_resolution_scanout",
                "skip_deinterlace")
SCANOUT_FIELDS = ("width", "height", "internal_width", "internal_height", "mode_width", "mode_height",
                  "high_resolution_scanout", "interlaced", "double_strike", "interlace_phase")


class VsyncInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in VSYNC_FIELDS]


class ScanoutOut(ctypes.Structure):
    _fields_ = [("has_image", ctypes.c_uint32)] + [(n, ctypes.c_uint32) for n in SCANOUT_FIELDS] + \
        [("sampling_rate_mhz", ctypes.c_float)]


def vsync_info(**kw) -> VsyncInfo:
    bad = set(kw) - set(VSYNC_FIELDS)
    if bad:
        raise TypeError(f"unknown vsync fields {bad}")
    return VsyncInfo(**{k: int(v) for k, v in kw.items()})


def oracle_lib() -> ctypes.CDLL:
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run __graft_entry__.build()")
        L = ctypes.CDLL(ORACLE_PATH)
        vp, sz, u32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint64
        L.orc_create.restype = vp
        L.orc_create.argtypes = [u32]
        L.orc_destroy.argtypes = [vp]
        L.orc_destroy.restype = None
        L.orc_gif_transfer.argtypes = [vp, u32, vp, sz]
        L.orc_gif_transfer.restype = None
        L.orc_write_register.argtypes = [vp, u32, u64]
        L.orc_write_register.restype = None
        L.orc_flush.argtypes = [vp]
        L.orc_flush.restype = None
        L.orc_vram.argtypes = [vp]
        L.orc_vram.restype = vp
        L.orc_vram_size.argtypes = [vp]
        L.orc_vram_size.restype = sz
        L.orc_set_vram.argtypes = [vp, vp, sz]
        L.orc_set_vram.restype = None
        L.orc_swizzle.argtypes = [u32, u32, u32, u32, u32]
        L.orc_swizzle.restype = u32
        L.orc_stat.argtypes = [vp, u32]
        L.orc_stat.restype = u64
        L.orc_read_transfer_fifo.argtypes = [vp, vp, u32]
        L.orc_read_transfer_fifo.restype = None
        L.orc_host_write.argtypes = [vp, sz, vp, sz]
        L.orc_host_write.restype = None
        L.orc_set_sample_log2.argtypes = [vp, u32]
        L.orc_set_sample_log2.restype = None
        L.orc_reset_context_state.argtypes = [vp]
        L.orc_reset_context_state.restype = None
        L.orc_set_sparse_grid.argtypes = [vp, u32]
        L.orc_set_sparse_grid.restype = None
        L.orc_set_super_sampled_textures.argtypes = [vp, u32]
        L.orc_set_super_sampled_textures.restype = None
        L.orc_phase_lut.argtypes = [u32, u32, vp]
        L.orc_phase_lut.restype = None
        L.orc_set_priv.argtypes = [vp, vp, sz]
        L.orc_set_priv.restype = None
        L.orc_complex_scanmsk_timeout.argtypes = [vp]
        L.orc_complex_scanmsk_timeout.restype = u32
        L.orc_vsync_can_skip.argtypes = [vp, ctypes.POINTER(VsyncInfo)]
        L.orc_vsync_can_skip.restype = ctypes.c_int
        L.orc_vsync.argtypes = [vp, ctypes.POINTER(VsyncInfo), ctypes.POINTER(ScanoutOut), vp, sz]
        L.orc_vsync.restype = ctypes.c_int
        _L = L
    return _L


class Oracle:
    """Immediate-mode CPU GS (one context)."""

    def __init__(self, sample_log2: int = 0, sparse: bool = False, super_sampled_textures: bool = False):
        """sparse: the 1x4 (4x) / 2x8 (16x) sampling grids instead of the ordered ones;
        super_sampled_textures: sprites may sample rendered pages per sample (off at 1x)."""
        self._L = oracle_lib()
        self._h = self._L.orc_create(sample_log2)
        if sparse:
            self._L.orc_set_sparse_grid(self._h, 1)
        if super_sampled_textures and sample_log2:
            self._L.orc_set_super_sampled_textures(self._h, 1)

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.orc_destroy(self._h)
            self._h = None

    def gif_transfer(self, data: np.ndarray, path: int = 3):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        self._L.orc_gif_transfer(self._h, path, data.ctypes.data, data.nbytes)

    def write_register(self, addr: int, value: int):
        self._L.orc_write_register(self._h, addr, value & ((1 << 64) - 1))

    def flush(self):
        self._L.orc_flush(self._h)

    def planes(self) -> np.ndarray:
        n = self._L.orc_vram_size(self._h)
        p = self._L.orc_vram(self._h)
        return np.frombuffer((ctypes.c_uint8 * n).from_address(p), dtype=np.uint8).copy()

    def vram(self) -> np.ndarray:
        self.flush()
        return self.planes()[:VRAM_SIZE]

    def set_vram(self, data: np.ndarray):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        self._L.orc_set_vram(self._h, data.ctypes.data, data.nbytes)

    def stat(self, which: int) -> int:
        return int(self._L.orc_stat(self._h, which))

    def read_transfer_fifo(self, n128: int) -> np.ndarray:
        out = np.zeros(n128 * 16, dtype=np.uint8)
        self._L.orc_read_transfer_fifo(self._h, out.ctypes.data, n128)
        return out

    def host_write(self, data: np.ndarray, offset: int = 0):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        self._L.orc_host_write(self._h, offset, data.ctypes.data, data.nbytes)

    def reset_context_state(self):
        self._L.orc_reset_context_state(self._h)

    def set_sample_log2(self, sample_log2: int, sparse: bool = False):
        self.flush()  # pending primitives render on the old grid
        self._L.orc_set_sparse_grid(self._h, int(sparse))
        self._L.orc_set_sample_log2(self._h, sample_log2)

    def set_priv(self, block: bytes):
        buf = np.frombuffer(bytes(block), dtype=np.uint8)
        self._L.orc_set_priv(self._h, buf.ctypes.data, buf.nbytes)

    def complex_scanmsk_timeout(self) -> int:
        return int(self._L.orc_complex_scanmsk_timeout(self._h))

    def vsync_can_skip(self, **info) -> bool:
        i = vsync_info(**info)
        return bool(self._L.orc_vsync_can_skip(self._h, ctypes.byref(i)))

    def vsync(self, **info):
        """Returns (status, metadata dict, image as (height, width) uint32 RGBA8 or None)."""
        i = vsync_info(**info)
        out = ScanoutOut()
        cap = 8192 * 4096
        pixels = np.zeros(cap, dtype=np.uint32)
        rc = self._L.orc_vsync(self._h, ctypes.byref(i), ctypes.byref(out), pixels.ctypes.data, cap)
        meta = {n: int(getattr(out, n)) for n in SCANOUT_FIELDS}
        meta["sampling_rate_mhz"] = float(out.sampling_rate_mhz)
        img = pixels[:out.width * out.height].reshape(out.height, out.width).copy() if out.has_image else None
        return int(rc), meta, img


def swizzle(x, y, base, stride, psm) -> int:
    return int(oracle_lib().orc_swizzle(x, y, base, stride, psm))


def oracle_phase_lut(rate_x: int, rate_y: int) -> np.ndarray:
    """The oracle's phase filter table of super-sampled textures (2 * (64 * 64 + 32 * 32) words)."""
    out = np.zeros(2 * (64 * 64 + 32 * 32), np.uint32)
    oracle_lib().orc_phase_lut(rate_x, rate_y, out.ctypes.data)
    return out
