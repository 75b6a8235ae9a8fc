"""Tiny GIF packet builder for hand-made test streams."""
from __future__ import annotations

import struct

import numpy as np


def giftag(nloop, flg=0, nreg=1, regs=0xE, pre=False, prim=0, eop=False) -> bytes:
    lo = (nloop & 0x7FFF) | (int(eop) << 15) | (int(pre) << 46) | ((prim & 0x7FF) << 47) | \
         ((flg & 3) << 58) | ((nreg & 0xF) << 60)
    return struct.pack("<QQ", lo, regs)


def ad(writes) -> bytes:
    """PACKED A+D writes: list of (addr, value)."""
    out = [giftag(len(writes))]
    for addr, value in writes:
        out.append(struct.pack("<QQ", value & (2**64 - 1), addr))
    return b"".join(out)


def image(payload: bytes) -> bytes:
    assert len(payload) % 16 == 0
    out = []
    q = len(payload) // 16
    off = 0
    while off < q:
        n = min(q - off, 32767)
        out.append(giftag(n, flg=2, nreg=0, regs=0))
        out.append(payload[off * 16:(off + n) * 16])
        off += n
    return b"".join(out)


def bitbltbuf(sbp=0, sbw=0, spsm=0, dbp=0, dbw=0, dpsm=0):
    return sbp | sbw << 16 | spsm << 24 | dbp << 32 | dbw << 48 | dpsm << 56


def trxpos(ssax=0, ssay=0, dsax=0, dsay=0):
    return ssax | ssay << 16 | dsax << 32 | dsay << 48


def trxreg(w, h):
    return w | h << 32


def upload(dbp, dbw, dpsm, dsax, dsay, w, h, payload: bytes) -> bytes:
    return ad([(0x50, bitbltbuf(dbp=dbp, dbw=dbw, dpsm=dpsm)), (0x51, trxpos(dsax=dsax, dsay=dsay)),
               (0x52, trxreg(w, h)), (0x53, 0)]) + image(payload)


def local_copy(sbp, sbw, spsm, ssax, ssay, dbp, dbw, dpsm, dsax, dsay, w, h) -> bytes:
    return ad([(0x50, bitbltbuf(sbp, sbw, spsm, dbp, dbw, dpsm)), (0x51, trxpos(ssax, ssay, dsax, dsay)),
               (0x52, trxreg(w, h)), (0x53, 2)])


def arr(b: bytes) -> np.ndarray:
    return np.frombuffer(b, dtype=np.uint8)


BPP = {0x00: 32, 0x01: 24, 0x02: 16, 0x0A: 16, 0x13: 8, 0x14: 4, 0x1B: 8, 0x24: 4, 0x2C: 4,
       0x30: 32, 0x31: 24, 0x32: 16, 0x3A: 16}


def unpack_pixels(payload: bytes, bpp: int, n: int) -> np.ndarray:
    b = np.frombuffer(payload, dtype=np.uint8)
    if bpp == 32:
        return b[:n * 4].view(np.uint32).astype(np.uint64)
    if bpp == 24:
        t = b[:n * 3].reshape(-1, 3).astype(np.uint64)
        return t[:, 0] | (t[:, 1] << 8) | (t[:, 2] << 16)
    if bpp == 16:
        return b[:n * 2].view(np.uint16).astype(np.uint64)
    if bpp == 8:
        return b[:n].astype(np.uint64)
    lo = (b & 0xF).astype(np.uint64)
    hi = (b >> 4).astype(np.uint64)
    return np.stack([lo, hi], axis=1).reshape(-1)[:n]
