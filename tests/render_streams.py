"""Random GS drawing streams for parity tests (any stream is valid input for both the oracle
and the GPU path; these exercise register state the benchmark workloads do not)."""
from __future__ import annotations

import struct

import numpy as np

import gif

FB_PSMS = [0x00, 0x01, 0x02, 0x0A, 0x30, 0x31, 0x32, 0x3A]
Z_PSMS = [0x0, 0x1, 0x2, 0xA]  # ZBUF.PSM low bits: Z32, Z24, Z16, Z16S


def prim_reg(prim_type, iip=1, tme=0, fge=0, abe=0, aa1=0, fst=1, ctxt=0, fix=0):
    return prim_type | iip << 3 | tme << 4 | fge << 5 | abe << 6 | aa1 << 7 | fst << 8 | ctxt << 9 | fix << 10


def frame(fbp, fbw, psm, fbmsk=0):
    return fbp | fbw << 16 | psm << 24 | fbmsk << 32


def zbuf(zbp, psm, zmsk=0):
    return zbp | (psm & 0xF) << 24 | zmsk << 32


def test_reg(ate=0, atst=1, aref=0, afail=0, date=0, datm=0, zte=1, ztst=1):
    return ate | atst << 1 | aref << 4 | afail << 12 | date << 14 | datm << 15 | zte << 16 | ztst << 17


def alpha_reg(a, b, c, d, fix):
    return a | b << 2 | c << 4 | d << 6 | fix << 32


def scissor(x0, x1, y0, y1):
    return x0 | x1 << 16 | y0 << 32 | y1 << 48


def xyz2(x, y, z):
    return (x & 0xFFFF) | (y & 0xFFFF) << 16 | (z & 0xFFFFFFFF) << 32


def xyzf2(x, y, z, f):
    return (x & 0xFFFF) | (y & 0xFFFF) << 16 | (z & 0xFFFFFF) << 32 | (f & 0xFF) << 56


def rgbaq(rgba, q=1.0):
    return (rgba & 0xFFFFFFFF) | struct.unpack("<I", struct.pack("<f", q))[0] << 32


def random_render_stream(seed: int, n_batches: int = 12, prims_per_batch: int = 40, width: int = 256,
                         height: int = 192, textured: bool = False) -> np.ndarray:
    """A sequence of batches; each batch randomises FB/Z format, test, blend, dither, masks
    and primitive type, then draws primitives with A+D vertex writes."""
    rng = np.random.default_rng(seed)
    ofs = 2048 * 16
    out = [gif.ad([(0x1A, 1), (0x18, ofs | ofs << 32)])]
    fbw = (width + 63) // 64
    for b in range(n_batches):
        fb_psm = int(rng.choice(FB_PSMS))
        z_low = int(rng.choice(Z_PSMS))
        zbp = 64 + int(rng.integers(0, 8)) * 32
        regs = [
            (0x4C, frame(0, fbw, fb_psm, int(rng.choice([0, 0, 0, 0xFF000000, 0x00FF00FF, int(rng.integers(0, 2**32))])))),
            (0x4E, zbuf(zbp, z_low, int(rng.integers(0, 4) == 0))),
            (0x40, scissor(int(rng.integers(0, 16)), width - 1 - int(rng.integers(0, 16)),
                           int(rng.integers(0, 16)), height - 1 - int(rng.integers(0, 16)))),
            (0x47, test_reg(ate=int(rng.integers(0, 2)), atst=int(rng.integers(0, 8)), aref=int(rng.integers(0, 256)),
                            afail=int(rng.integers(0, 4)), date=int(rng.integers(0, 4) == 0), datm=int(rng.integers(0, 2)),
                            zte=int(rng.integers(0, 5) != 0), ztst=int(rng.integers(1, 4)))),
            (0x42, alpha_reg(int(rng.integers(0, 3)), int(rng.integers(0, 3)), int(rng.integers(0, 3)),
                             int(rng.integers(0, 3)), int(rng.integers(0, 256)))),
            (0x45, int(rng.integers(0, 3) == 0)),
            (0x44, int(rng.integers(0, 2**64, dtype=np.uint64))),
            (0x46, int(rng.integers(0, 2))),
            (0x49, int(rng.integers(0, 2))),
            (0x4A, int(rng.integers(0, 2))),
            (0x3D, int(rng.integers(0, 2**24))),
        ]
        ptype = int(rng.choice([0, 1, 2, 3, 3, 3, 4, 5, 6, 6]))
        prim = prim_reg(ptype, iip=int(rng.integers(0, 2)), abe=int(rng.integers(0, 2)),
                        fge=int(rng.integers(0, 2)), fst=1, fix=int(rng.integers(0, 6) == 0))
        regs.append((0x00, prim))
        out.append(gif.ad(regs))
        writes = []
        quads = rng.integers(0, 3) == 0 and ptype == 3
        for p in range(prims_per_batch):
            if quads:
                # two triangles forming an axis-aligned rectangle (parallelogram fusion)
                x0, y0 = int(rng.integers(0, width * 16)), int(rng.integers(0, height * 16))
                w, h = int(rng.integers(16, 64 * 16)), int(rng.integers(16, 64 * 16))
                z = int(rng.integers(0, 2**24))
                col = int(rng.integers(0, 2**32))
                corners = [(x0, y0), (x0 + w, y0), (x0, y0 + h), (x0 + w, y0), (x0 + w, y0 + h), (x0, y0 + h)]
                for (x, y) in corners:
                    writes.append((0x01, rgbaq(col)))
                    writes.append((0x05, xyz2(ofs + x, ofs + y, z)))
                continue
            nv = {0: 1, 1: 2, 2: 2, 3: 3, 4: 3, 5: 3, 6: 2}[ptype]
            for v in range(nv):
                x = ofs + int(rng.integers(-32 * 16, (width + 32) * 16))
                y = ofs + int(rng.integers(-32 * 16, (height + 32) * 16))
                z = int(rng.integers(0, 2**32)) if rng.integers(0, 2) else int(rng.integers(0, 2**24))
                writes.append((0x01, rgbaq(int(rng.integers(0, 2**32)))))
                if rng.integers(0, 2):
                    writes.append((0x04, xyzf2(x, y, z, int(rng.integers(0, 256)))))
                else:
                    writes.append((0x0A, int(rng.integers(0, 256)) << 56))
                    writes.append((0x05, xyz2(x, y, z)))
        for i in range(0, len(writes), 200):
            out.append(gif.ad(writes[i:i + 200]))
    return gif.arr(b"".join(out))
