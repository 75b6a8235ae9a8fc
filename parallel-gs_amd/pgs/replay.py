"""Regression replayer for GS dumps (the reference's tools/gs_repro_replayer.cpp:29-330).

Every `*.gs` file of a folder is replayed on a fresh context (8x super-sampling by default,
as the reference's tool does) to its last vsync. The frame (alpha forced to 255) is compared
with `<dump>.reference.png` by per-channel PSNR (at least 40 dB passes) and the context's
flush statistics with `<dump>.reference.json`; a missing reference, or `--update`, writes it.

    python -m pgs.replay --dir <folder> [--update] [--samples N] [--high-res]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import struct
import sys
import zlib

import numpy as np

from . import Context, DumpReader

STAT_KEYS = ("render_passes", "primitives", "copies", "transfer_batches", "textures_decoded")


# ---- PNG (8-bit RGBA, as the reference's stb_image_write output) ----------------------------

def _chunk(kind: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def write_png(path: str, rgba: np.ndarray):
    """rgba: (h, w, 4) uint8."""
    h, w, _ = rgba.shape
    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(h))
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) + \
        _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def read_png(path: str) -> np.ndarray:
    """Reads an 8-bit RGB or RGBA, non-interlaced PNG as (h, w, 4) uint8."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG")
    pos, idat, w = 8, b"", 0
    while pos < len(data):
        (n,) = struct.unpack_from(">I", data, pos)
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if kind == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            if depth != 8 or ctype not in (2, 6) or interlace:
                raise ValueError(f"{path}: only 8-bit RGB/RGBA non-interlaced PNGs are read")
            c = 4 if ctype == 6 else 3
        elif kind == b"IDAT":
            idat += body
        elif kind == b"IEND":
            break
    raw = np.frombuffer(zlib.decompress(idat), np.uint8)
    stride = w * c
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        ft = raw[y * (stride + 1)]
        line = raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)].astype(np.int32)
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 255
        else:
            cur = np.zeros(stride, np.int32)
            for x in range(stride):
                a = cur[x - c] if x >= c else 0
                b = prev[x]
                cc = prev[x - c] if x >= c else 0
                if ft == 1:
                    p = a
                elif ft == 3:
                    p = (a + b) >> 1
                else:
                    pa, pb, pc = abs(b - cc), abs(a - cc), abs(a + b - 2 * cc)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else cc)
                cur[x] = (line[x] + p) & 255
        out[y] = cur
        prev = cur
    img = out.reshape(h, w, c)
    if c == 3:
        img = np.concatenate([img, np.full((h, w, 1), 255, np.uint8)], -1)
    return img


def frame_rgba(frame: np.ndarray) -> np.ndarray:
    """uint32 RGBA8 words -> (h, w, 4) bytes with alpha 255 (the reference's readback does so)."""
    rgba = frame.view(np.uint8).reshape(frame.shape[0], frame.shape[1], 4).copy()
    rgba[..., 3] = 255
    return rgba


def psnr(new: np.ndarray, ref: np.ndarray):
    """Per-channel PSNR of R, G, B in dB (inf when equal)."""
    out = []
    for ch in range(3):
        err = np.sum((new[..., ch].astype(np.int64) - ref[..., ch].astype(np.int64)) ** 2)
        peak = 255 * 255 * new.shape[0] * new.shape[1]
        out.append(math.inf if err == 0 else 10.0 * math.log10(peak / err))
    return out


def run_dump(path: str, samples: int = 8, update: bool = False, high_res: bool = False, log=print):
    """Replays one dump; returns (passed, details)."""
    log(f"=== Testing {path} ===")
    with Context(super_sampling=samples) as ctx:
        reader = DumpReader(ctx, path)
        while reader.iterate_until_vsync(high_res_scanout=high_res):
            pass
        meta, frame = reader.consume_vsync_result()
        reader.close()
        stats = ctx.consume_flush_stats()
    if frame is None:
        log("No image retrieved from vsync result.")
        return False, {}
    rgba = frame_rgba(frame)
    details = {"width": int(rgba.shape[1]), "height": int(rgba.shape[0]), **meta}
    ok = True
    ref_png, ref_json = path + ".reference.png", path + ".reference.json"
    if os.path.exists(ref_png):
        ref = read_png(ref_png)
        if ref.shape != rgba.shape:
            log("Mismatch in vsync image dimensions.")
            ok = False
        else:
            p = psnr(rgba, ref)
            details["psnr"] = p
            log("PSNR is [%.3f, %.3f, %.3f] dB." % tuple(p))
            if min(p) < 40.0:
                log("PSNR is too low for reference.")
                ok = False
        if not ok and not update:
            write_png(path + ".fault.png", rgba)
            log(f"Writing faulting image to {path}.fault.png.")
            return False, details
    if update or not os.path.exists(ref_png):
        write_png(ref_png, rgba)
    kept = {k: int(stats[k]) for k in STAT_KEYS}
    details["stats"] = kept
    if os.path.exists(ref_json):
        with open(ref_json) as f:
            old = json.load(f)
        for k in STAT_KEYS:
            log(f"{k}: {old.get(k)} -> {kept[k]}")
            if old.get(k) != kept[k]:
                log(f"{k} mismatch.")
                ok = False
    if update or not os.path.exists(ref_json):
        with open(ref_json, "w") as f:
            json.dump(kept, f, indent=2)
    return ok or update, details


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--dir", default=".")
    ap.add_argument("--update", action="store_true")
    ap.add_argument("--samples", type=int, default=8)
    ap.add_argument("--high-res", action="store_true")
    a = ap.parse_args(argv)
    for name in sorted(os.listdir(a.dir)):
        if name.endswith(".gs"):
            ok, _ = run_dump(os.path.join(a.dir, name), a.samples, a.update, a.high_res)
            if not ok:
                return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
