"""GS dump files for tests: a writer of the PCSX2 GS-dump format (state versions 8 and 9) and
a replay of a dump into the CPU oracle.

Format (dump/gs_dump_parser.cpp:6-207, dump/gs_dump_generator.cpp:39-237): u32 0xffffffff,
u32 header size, a 36-byte header (version, state size, serial / CRC / screenshot fields), the
register state (u32 version; 15 u64: PRIM, PRMODECONT, TEXCLUT, SCANMSK, TEXA, FOGCOL, DIMX,
DTHE, COLCLAMP, PABE, BITBLTBUF, TRXDIR, TRXPOS, TRXREG, a dummy; per context 12 u64:
XYOFFSET, TEX0, TEX1, CLAMP, MIPTBP1, MIPTBP2, SCISSOR, ALPHA, TEST, FBA, FRAME, ZBUF; RGBAQ, ST;
u32 UV and u32 FOG low words; u64 XYZ; u32 x 4 dummies; version 9 adds 61 bytes), 4 MiB of VRAM,
four GIF paths (16-byte tag + u32 register index), f32 internal Q, 8 KiB of privileged
registers; then packets: 0 Transfer (u8 path, u32 size, data), 1 Vsync (u8 field),
2 ReadFIFO (u32 size), 3 PrivRegisters (8 KiB).
"""
from __future__ import annotations

import struct

import numpy as np

VRAM_SIZE = 4 * 1024 * 1024
GLOBAL = ("prim", "prmodecont", "texclut", "scanmsk", "texa", "fogcol", "dimx", "dthe", "colclamp", "pabe",
          "bitbltbuf", "trxdir", "trxpos", "trxreg")
PER_CONTEXT = ("xyoffset", "tex0", "tex1", "clamp", "miptbp1", "miptbp2", "scissor", "alpha", "test", "fba", "frame",
               "zbuf")
TRANSFER, VSYNC, READ_FIFO, PRIV_REGISTERS = 0, 1, 2, 3
STATE_SIZE = VRAM_SIZE + (18 + 12 * 2) * 8 + 7 * 4 + 4 * 5 * 4 + 4


class Writer:
    """Builds a dump in memory."""

    def __init__(self, vram=None, regs=None, paths=None, internal_q=1.0, priv=None, state_version=8,
                 header_extra=b""):
        regs = regs or {}
        out = bytearray()
        out += struct.pack("<II", 0xFFFFFFFF, 36 + len(header_extra))
        out += struct.pack("<9I", 9, STATE_SIZE, 0, 0, 0, 0, 0, 0, 0) + header_extra
        out += struct.pack("<I", state_version)
        out += struct.pack("<14Q", *[regs.get(n, 0) for n in GLOBAL])
        out += struct.pack("<Q", regs.get("trxreg", 0))
        for c in range(2):
            out += struct.pack("<12Q", *[regs.get(f"{n}{c}", 0) for n in PER_CONTEXT])
        out += struct.pack("<QQII", regs.get("rgbaq", 0), regs.get("st", 0), regs.get("uv", 0) & 0xFFFFFFFF,
                           regs.get("fog", 0) & 0xFFFFFFFF)
        out += struct.pack("<QIIII", 0, 0xFFFFFFFF, 0xFFFFFFFF, 0, 0)
        if state_version >= 9:
            out += bytes(61)
        v = np.zeros(VRAM_SIZE, np.uint8) if vram is None else np.asarray(vram, np.uint8)
        assert v.size == VRAM_SIZE
        out += v.tobytes()
        for i in range(4):
            lo, hi, reg = (paths or [(0, 0, 0)] * 4)[i]
            out += struct.pack("<QQI", lo, hi, reg)
        out += struct.pack("<f", internal_q)
        out += bytes(priv) if priv is not None else bytes(8192)
        self.data = out

    def transfer(self, stream, path=1):
        b = np.ascontiguousarray(stream, np.uint8).tobytes()
        self.data += struct.pack("<BBI", TRANSFER, path, len(b)) + b

    def priv(self, block: bytes):
        assert len(block) == 8192
        self.data += struct.pack("<B", PRIV_REGISTERS) + bytes(block)

    def vsync(self, field: int, priv: bytes | None = None):
        if priv is not None:
            self.priv(priv)
        self.data += struct.pack("<BB", VSYNC, field)

    def read_fifo(self, size: int):
        self.data += struct.pack("<BI", READ_FIFO, size)

    def save(self, path):
        with open(path, "wb") as f:
            f.write(bytes(self.data))
        return path


def parse(data: bytes):
    """Splits a dump into its start state and packet list (for checks and the oracle replay)."""
    pos = 0

    def take(fmt):
        nonlocal pos
        vals = struct.unpack_from(fmt, data, pos)
        pos += struct.calcsize(fmt)
        return vals

    crc, header_size = take("<II")
    assert crc == 0xFFFFFFFF and header_size >= 36
    header = take("<9I")
    pos += header_size - 36
    (version,) = take("<I")
    g = take("<14Q")
    take("<Q")
    regs = dict(zip(GLOBAL, g))
    for c in range(2):
        regs.update({f"{n}{c}": v for n, v in zip(PER_CONTEXT, take("<12Q"))})
    regs["rgbaq"], regs["st"], regs["uv"], regs["fog"] = take("<QQII")
    take("<QIIII")
    if version >= 9:
        pos += 61
    vram = np.frombuffer(data, np.uint8, VRAM_SIZE, pos)
    pos += VRAM_SIZE
    paths = [take("<QQI") for _ in range(4)]
    (q,) = take("<f")
    priv = data[pos:pos + 8192]
    pos += 8192
    packets = []
    while pos < len(data):
        t = data[pos]
        pos += 1
        if t == TRANSFER:
            p, n = take("<BI")
            packets.append(("transfer", p, np.frombuffer(data, np.uint8, n, pos)))
            pos += n
        elif t == VSYNC:
            packets.append(("vsync", data[pos]))
            pos += 1
        elif t == READ_FIFO:
            packets.append(("read_fifo", take("<I")[0]))
        elif t == PRIV_REGISTERS:
            packets.append(("priv", data[pos:pos + 8192]))
            pos += 8192
    return dict(header=header, version=version, regs=regs, vram=vram, paths=paths, internal_q=q, priv=priv,
                packets=packets)


def replay_into_oracle(oracle, data: bytes, high_res_scanout=False):
    """GSDumpParser::restart + iterate_until_vsync to the end for a dump whose start state is
    the reset register state with idle GIF paths: returns the last vsync's (status, metadata,
    image)."""
    d = parse(data)
    assert all(v == 0 for v in d["regs"].values()), "the oracle replay starts from the reset register state"
    assert all(lo & 0x7FFF == 0 for lo, _, _ in d["paths"]), "the oracle replay starts with idle GIF paths"
    oracle.host_write(np.array(d["vram"]), 0)
    oracle.set_priv(d["priv"])
    last = (0, None, None)
    info = dict(adapt_to_internal_horizontal_resolution=1, high_resolution_scanout=int(high_res_scanout),
                force_progressive=1, anti_blur=1)
    for pkt in d["packets"]:
        if pkt[0] == "transfer":
            n = len(pkt[2]) & ~15
            oracle.gif_transfer(pkt[2][:n], pkt[1] & 3)
        elif pkt[0] == "priv":
            oracle.set_priv(pkt[1])
        elif pkt[0] == "vsync":
            last = oracle.vsync(phase=pkt[1], **info)
        elif pkt[0] == "read_fifo":
            break
    return last
