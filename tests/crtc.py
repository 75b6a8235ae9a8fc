"""Privileged CRTC register encoders for scanout tests (bit layouts: gs/gs_registers.hpp:562-736;
register block layout: gs/gs_interface.hpp:70-108, one 16-byte slot per register)."""
from __future__ import annotations

import struct

SLOTS = {"PMODE": 0, "SMODE1": 1, "SMODE2": 2, "DISPFB1": 7, "DISPLAY1": 8, "DISPFB2": 9, "DISPLAY2": 10,
         "EXTBUF": 11, "EXTDATA": 12, "EXTWRITE": 13, "BGCOLOR": 14}

CMOD_PROGRESSIVE, CMOD_NTSC, CMOD_PAL = 0, 2, 3
LC_ANALOG, LC_HDTV, LC_VGA = 32, 22, 15


def pmode(en1=1, en2=0, mmod=1, slbg=0, alp=0xFF, amod=0, crtmd=1) -> int:
    return en1 | en2 << 1 | crtmd << 2 | mmod << 5 | amod << 6 | slbg << 7 | alp << 8


def smode1(cmod=CMOD_NTSC, lc=LC_ANALOG) -> int:
    return lc << 3 | cmod << 13


def smode2(interlace=1, ffmd=0) -> int:
    return interlace | ffmd << 1


def dispfb(fbp=0, fbw=10, psm=0, dbx=0, dby=0) -> int:
    return fbp | fbw << 9 | psm << 15 | dbx << 32 | dby << 43


def display(dx=636, dy=50, magh=3, magv=0, dw=2559, dh=447) -> int:
    return dx | dy << 12 | magh << 23 | magv << 27 | dw << 32 | dh << 44


def extbuf(exbp=0, exbw=10, fbin=0, wffmd=1, emoda=0, emodc=0, wdx=0, wdy=0) -> int:
    return exbp | exbw << 14 | fbin << 20 | wffmd << 22 | emoda << 23 | emodc << 25 | wdx << 32 | wdy << 43


def extdata(sx=0, sy=0, smph=0, smpv=0, ww=639, wh=223) -> int:
    return sx | sy << 12 | smph << 23 | smpv << 27 | ww << 32 | wh << 44


def bgcolor(r=0, g=0, b=0) -> int:
    return r | g << 8 | b << 16


def priv_block(**regs) -> bytes:
    """8 KiB block with the named registers set (others 0)."""
    buf = bytearray(8192)
    for name, value in regs.items():
        struct.pack_into("<Q", buf, 16 * SLOTS[name], value & (2**64 - 1))
    return bytes(buf)


def ntsc_interlaced_640x448(fbp=0, fbw=10, psm=0, ffmd=0, **extra) -> bytes:
    """A common NTSC set-up: 640x448 frame buffer shown by circuit 1, field or frame mode."""
    regs = dict(PMODE=pmode(), SMODE1=smode1(), SMODE2=smode2(1, ffmd), DISPFB1=dispfb(fbp, fbw, psm),
                DISPLAY1=display())
    regs.update(extra)
    return priv_block(**regs)
