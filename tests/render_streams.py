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
                         height: int = 192, textured: bool = False, aa1: bool = False) -> np.ndarray:
    """A sequence of batches; each batch randomises FB/Z format, test, blend, dither, masks
    and primitive type, then draws primitives with A+D vertex writes. aa1: half of the
    batches set PRIM.AA1 (the other seeds' streams are unchanged)."""
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
        if aa1 and rng.integers(0, 2):
            prim |= 1 << 7
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


# ---------------------------------------------------------------------------------------
# Texture state (GS register layouts: TEX0, CLAMP, TEX1, TEXA, TEXCLUT, UV, ST).

def tex0(tbp, tbw, psm, tw, th, tcc=1, tfx=1, cbp=0, cpsm=0, csm=0, csa=0, cld=0):
    return (tbp | tbw << 14 | psm << 20 | tw << 26 | th << 30 | tcc << 34 | tfx << 35 | cbp << 37 |
            cpsm << 51 | csm << 55 | csa << 56 | cld << 61)


def clamp_reg(wms=0, wmt=0, minu=0, maxu=0, minv=0, maxv=0):
    return wms | wmt << 2 | minu << 4 | maxu << 14 | minv << 24 | maxv << 34


def tex1(mmag=0, mmin=0, lcm=0, mxl=0, l=0, k=0):
    """TEX1: LCM bit 0, MXL bits 2..4, MMAG bit 5, MMIN bits 6..8, L bits 19..20, K (signed
    7.4 fixed point, 12 bits) bits 32..43."""
    return lcm | mxl << 2 | mmag << 5 | mmin << 6 | l << 19 | (k & 0xFFF) << 32


def miptbp(levels):
    """MIPTBP1 / MIPTBP2: three (TBP, TBW) pairs of 14 + 6 bits."""
    v = 0
    for i, (tbp, tbw) in enumerate(levels):
        v |= (tbp | tbw << 14) << (20 * i)
    return v


def texa(ta0=0, aem=0, ta1=0x80):
    return ta0 | aem << 15 | ta1 << 32


def texclut(cbw, cou, cov):
    return cbw | cou << 6 | cov << 12


def uv(u, v):
    return (u & 0x3FFF) | (v & 0x3FFF) << 16


def st(s, t):
    f = lambda x: struct.unpack("<I", struct.pack("<f", x))[0]
    return f(s) | f(t) << 32


def csm1_position(i):
    """CLUT entry i of a CSM1 palette: 8x2 column blocks, the two halves of each 32-entry group
    side by side (the GS CLUT layout for 256-colour palettes)."""
    group, j = divmod(i, 32)
    return (j // 16) * 8 + j % 8, 2 * group + (j // 8) % 2


TEX_PSMS = [0x00, 0x01, 0x02, 0x0A, 0x13, 0x14, 0x1B, 0x24, 0x2C]


def _random_payload(rng, nbytes):
    return rng.integers(0, 256, size=(nbytes + 15) // 16 * 16, dtype=np.uint8).tobytes()


def random_textured_stream(seed: int, n_batches: int = 10, prims_per_batch: int = 30, width: int = 256,
                           height: int = 192, feedback: bool = True, mipmap: bool = False) -> np.ndarray:
    """Textured drawing: random texture formats / sizes / wrap modes / filters / combiners,
    CLUT loads (CSM1, CSM2, CSA, CLD 0..5), UV and ST/Q coordinates, and (optionally) textures
    read from the frame buffer being drawn to (render-pass hazard flushes). With `mipmap`, TEX1
    takes random mipmapped filters (MMIN 2..5 mostly), MXL, LCM, L and K, MIPTBP1 / MIPTBP2 point
    the levels at random texture pages, and Q spreads over 2^-3..2^3 so the LOD varies."""
    rng = np.random.default_rng(seed)
    ofs = 2048 * 16
    out = [gif.ad([(0x1A, 1), (0x18, ofs | ofs << 32)])]
    fbw = (width + 63) // 64
    # Texture and CLUT area: pages 400..511.
    tex_pages = [400 + 8 * i for i in range(12)]
    # Fill the texture area with random bytes (as 32-bit pixels, 64 px wide rows).
    for page in tex_pages:
        out.append(gif.upload(page * 32, 1, 0x00, 0, 0, 64, 64, _random_payload(rng, 64 * 64 * 4)))
    clut_bp = 505 * 32
    out.append(gif.upload(clut_bp, 1, 0x00, 0, 0, 64, 32, _random_payload(rng, 64 * 32 * 4)))
    for b in range(n_batches):
        fb_psm = int(rng.choice([0x00, 0x00, 0x01, 0x02]))
        zbp = 64 + int(rng.integers(0, 8)) * 32
        psm = int(rng.choice(TEX_PSMS))
        tw, th = int(rng.integers(2, 8)), int(rng.integers(2, 8))
        from_fb = feedback and rng.integers(0, 6) == 0
        if from_fb:
            psm, tbp, tbw = fb_psm, 0, fbw
        else:
            tbp = int(rng.choice(tex_pages)) * 32 + int(rng.integers(0, 4)) * 32
            tbw = int(rng.integers(1, 5)) * 2
        cpsm = int(rng.choice([0x00, 0x02, 0x0A]))
        csm = int(rng.integers(0, 4) == 0)
        csa = int(rng.integers(0, 32)) if rng.integers(0, 2) else 0
        cld = int(rng.choice([0, 1, 1, 2, 3, 4, 5]))
        cbp = clut_bp + int(rng.integers(0, 4)) * 4
        wms, wmt = int(rng.integers(0, 4)), int(rng.integers(0, 4))
        lim = lambda n: int(rng.integers(0, 1 << n))
        cl = clamp_reg(wms, wmt, lim(tw), lim(tw), lim(th), lim(th))
        if wms == 3:
            cl = (cl & ~(0xFFFFF << 4)) | (int(rng.integers(0, 1 << tw)) << 4) | (int(rng.integers(0, 64)) << 14)
        if wmt == 3:
            cl = (cl & ~(0xFFFFF << 24)) | (int(rng.integers(0, 1 << th)) << 24) | (int(rng.integers(0, 64)) << 34)
        regs = [
            (0x4C, frame(0, fbw, fb_psm)),
            (0x4E, zbuf(zbp, 1, int(rng.integers(0, 3) == 0))),
            (0x40, scissor(0, width - 1, 0, height - 1)),
            (0x47, test_reg(ate=int(rng.integers(0, 2)), atst=int(rng.integers(0, 8)), aref=int(rng.integers(0, 256)),
                            afail=int(rng.integers(0, 4)), zte=1, ztst=int(rng.integers(1, 4)))),
            (0x42, alpha_reg(0, 1, 0, 1, 0)),
            (0x3B, texa(int(rng.integers(0, 256)), int(rng.integers(0, 2)), int(rng.integers(0, 256)))),
            (0x1C, texclut(int(rng.integers(1, 4)), int(rng.integers(0, 4)), int(rng.integers(0, 8)))),
            (0x14, tex1(int(rng.integers(0, 2)), int(rng.integers(0, 2))) if not mipmap else
             tex1(int(rng.integers(0, 2)), int(rng.choice([2, 3, 4, 5, 2, 3, 4, 5, 0, 1, 6])),
                  lcm=int(rng.integers(0, 2)), mxl=int(rng.integers(0, 8)), l=int(rng.integers(0, 4)),
                  k=int(rng.integers(-96, 160)))),
            (0x08, cl),
            (0x06, tex0(tbp, tbw, psm, tw, th, tcc=int(rng.integers(0, 2)), tfx=int(rng.integers(0, 4)), cbp=cbp,
                        cpsm=cpsm, csm=csm, csa=csa, cld=cld)),
            (0x3D, int(rng.integers(0, 2**24))),
        ]
        if mipmap:
            # Now and then a level is read from the frame buffer being drawn (a hazard on a level
            # other than 0).
            lv = lambda: ((0, fbw) if rng.integers(0, 6) == 0 else
                          (int(rng.choice(tex_pages)) * 32 + int(rng.integers(0, 8)) * 4, int(rng.integers(1, 5))))
            regs.insert(0, (0x34, miptbp([lv(), lv(), lv()])))
            regs.insert(0, (0x36, miptbp([lv(), lv(), lv()])))
        ptype = int(rng.choice([3, 3, 4, 5, 6, 6, 1, 2]))
        fst = int(rng.integers(0, 3) != 0)
        prim = prim_reg(ptype, iip=int(rng.integers(0, 2)), tme=1, abe=int(rng.integers(0, 2)),
                        fge=int(rng.integers(0, 2)), fst=fst)
        regs.append((0x00, prim))
        out.append(gif.ad(regs))
        writes = []
        nv = {1: 2, 2: 2, 3: 3, 4: 3, 5: 3, 6: 2}[ptype]
        tsz_u, tsz_v = (1 << tw) * 16, (1 << th) * 16
        for p in range(prims_per_batch):
            q_const = float(rng.choice([1.0, 0.5, 2.0]))
            if mipmap:
                q_const = float(2.0 ** rng.uniform(-3, 3)) if rng.integers(0, 2) else float(2.0 ** rng.integers(-3, 4))
            for v in range(nv):
                x = ofs + int(rng.integers(-16 * 16, (width + 16) * 16))
                y = ofs + int(rng.integers(-16 * 16, (height + 16) * 16))
                z = int(rng.integers(0, 2**24))
                q = q_const if rng.integers(0, 2) else float(rng.uniform(0.5, 2.0))
                if mipmap and rng.integers(0, 2):
                    q = q_const * float(2.0 ** rng.uniform(-1, 1))
                writes.append((0x01, rgbaq(int(rng.integers(0, 2**32)), q)))
                if fst:
                    writes.append((0x03, uv(int(rng.integers(0, 2 * tsz_u)), int(rng.integers(0, 2 * tsz_v)))))
                else:
                    writes.append((0x02, st(float(rng.uniform(-0.5, 2.0)) * q, float(rng.uniform(-0.5, 2.0)) * q)))
                writes.append((0x05, xyz2(x, y, z)))
        for i in range(0, len(writes), 240):
            out.append(gif.ad(writes[i:i + 240]))
    return gif.arr(b"".join(out))


def random_fb_feedback_stream(seed: int, n_batches: int = 10, prims_per_batch: int = 12) -> np.ndarray:
    """Frame-buffer feedback: a 128 x 64 frame (CT32, CT24 or CT16) with random contents and 3D
    triangles (varying depth), then sprites and triangles textured from the frame buffer itself
    (same layout, 24-bit or palette views in the top bits of the word, a 16-bit view of a 32-bit
    frame) or from the depth buffer, with UVs on the pixels' own texels (offsets 0..15 / 16, half
    a texel with LINEAR) or off them, in UV and ST form, with random combiners, TEXA, blending
    and depth tests."""
    rng = np.random.default_rng(seed)
    ofs = 2048 * 16
    fw, fh = 128, 64
    zbp = 64
    out = [gif.ad([(0x1A, 1), (0x18, ofs | ofs << 32)])]
    out.append(gif.upload(0, 2, 0x00, 0, 0, fw, fh, _random_payload(rng, fw * fh * 4)))
    out.append(gif.upload(zbp * 32, 2, 0x00, 0, 0, fw, fh, _random_payload(rng, fw * fh * 4)))
    clut_bp = 2048
    out.append(gif.upload(clut_bp, 1, 0x00, 0, 0, 16, 16, _random_payload(rng, 16 * 16 * 4)))
    for b in range(n_batches):
        fb_psm = int(rng.choice([0x00, 0x00, 0x01, 0x02]))
        z_psm = int(rng.integers(0, 3))
        regs = [(0x4C, frame(0, 2, fb_psm)), (0x4E, zbuf(zbp, z_psm, int(rng.integers(0, 2)))),
                (0x40, scissor(0, fw - 1, 0, fh - 1)),
                (0x47, test_reg(zte=1, ztst=int(rng.integers(1, 4)))),
                (0x42, alpha_reg(0, 1, 0, 1, 0)),
                (0x3B, texa(int(rng.integers(0, 256)), int(rng.integers(0, 2)), int(rng.integers(0, 256))))]
        out.append(gif.ad(regs))
        writes = []
        for p in range(prims_per_batch):
            kind = int(rng.integers(0, 4))
            if kind == 0:
                # A 3D triangle.
                writes.append((0x00, prim_reg(3, iip=1, abe=int(rng.integers(0, 2)))))
                for v in range(3):
                    writes.append((0x01, rgbaq(int(rng.integers(0, 2**32)))))
                    writes.append((0x05, xyz2(ofs + int(rng.integers(-8 * 16, (fw + 8) * 16)),
                                              ofs + int(rng.integers(-8 * 16, (fh + 8) * 16)),
                                              int(rng.integers(0, 2**24)))))
                continue
            depth = rng.integers(0, 5) == 0
            if depth:
                tpsm = [0x30, 0x31, 0x32][z_psm]
                tbp, cpsm = zbp * 32, 0
            else:
                choices = [fb_psm]
                if fb_psm == 0x00:
                    choices += [0x01, 0x1B, 0x24, 0x2C, 0x02]
                tpsm = int(rng.choice(choices))
                tbp, cpsm = 0, int(rng.choice([0x00, 0x02]))
            mmag = int(rng.integers(0, 2))
            d = int(rng.choice([0, 3, 8, 8, 15, 24, 8 if mmag else 5]))
            t1 = tex1(mmag, 0)
            t0 = tex0(tbp, 2, tpsm, 7, 6, tcc=int(rng.integers(0, 2)), tfx=int(rng.integers(0, 4)), cbp=clut_bp,
                      cpsm=cpsm, csa=int(rng.integers(0, 16)), cld=int(rng.choice([0, 1])))
            fst = int(rng.integers(0, 4) != 0)
            writes += [(0x14, t1), (0x08, 0), (0x06, t0)]
            x0, y0 = int(rng.integers(0, fw - 8)), int(rng.integers(0, fh - 8))
            x1, y1 = x0 + int(rng.integers(1, fw - x0)), y0 + int(rng.integers(1, fh - y0))
            if kind == 1 or kind == 3:
                corners = [(x0, y0), (x1, y1)]
                writes.append((0x00, prim_reg(6, iip=0, tme=1, abe=int(rng.integers(0, 2)), fst=fst)))
            else:
                corners = [(x0, y0), (x1, y0), (x0, y1)]
                writes.append((0x00, prim_reg(3, iip=1, tme=1, abe=int(rng.integers(0, 2)), fst=fst)))
            q = float(rng.choice([1.0, 0.5, 2.0]))
            for (x, y) in corners:
                u, v = x * 16 + d, y * 16 + d
                writes.append((0x01, rgbaq(int(rng.integers(0, 2**32)), q)))
                if fst:
                    writes.append((0x03, uv(u, v)))
                else:
                    writes.append((0x02, st(u / (fw * 16.0) * q, v / (fh * 16.0) * q)))
                writes.append((0x05, xyz2(ofs + x * 16, ofs + y * 16, int(rng.integers(0, 2**24)))))
        for i in range(0, len(writes), 240):
            out.append(gif.ad(writes[i:i + 240]))
    return gif.arr(b"".join(out))


def random_fbz_alias_stream(seed: int, n_batches: int = 10, prims_per_batch: int = 16) -> np.ndarray:
    """Depth buffer at the colour buffer's base (FBP == ZBP == 0, FBW 2, 128 x 64 of random
    contents): batches of random colour / depth formats whose primitives (sprites of every size,
    flat and 3D triangles) switch between no depth test, test-only (ZMSK) and depth writes, with
    random colours, depths, blending and alpha tests; some batches draw test-only sprites only
    (runs that end at overlapping sprites)."""
    rng = np.random.default_rng(seed)
    ofs = 2048 * 16
    fw, fh = 128, 64
    out = [gif.ad([(0x1A, 1), (0x18, ofs | ofs << 32)])]
    out.append(gif.upload(0, 2, 0x00, 0, 0, fw, fh, _random_payload(rng, fw * fh * 4)))
    for b in range(n_batches):
        fb_psm = int(rng.choice([0x00, 0x00, 0x01, 0x02, 0x0A]))
        z_psm = int(rng.choice([0, 1, 2, 0xA] if fb_psm in (0x02, 0x0A) else [0, 0, 1, 2]))
        zmax = 0xFFFF if z_psm in (2, 0xA) else 0xFFFFFF if z_psm == 1 else 0xFFFFFFFF
        test_only = rng.integers(0, 4) == 0
        out.append(gif.ad([(0x4C, frame(0, 2, fb_psm)), (0x40, scissor(0, fw - 1, 0, fh - 1)),
                           (0x42, alpha_reg(0, 1, 0, 1, 0))]))
        writes = []
        for p in range(prims_per_batch):
            mode = 1 if test_only else int(rng.choice([0, 1, 1, 2]))  # 0 no depth, 1 test only, 2 writes
            ztst = 1 if mode == 0 else int(rng.integers(1 if mode == 2 else 2, 4))
            ate = int(rng.integers(0, 4) == 0)
            writes += [(0x4E, zbuf(0, z_psm, 0 if mode == 2 else 1)),
                       (0x47, test_reg(ate=ate, atst=int(rng.integers(0, 8)), aref=int(rng.integers(0, 256)),
                                       afail=int(rng.integers(0, 4)), zte=1, ztst=ztst))]
            kind = 0 if test_only else int(rng.integers(0, 3))
            abe = int(rng.integers(0, 3) == 0)
            if kind == 0:
                big = rng.integers(0, 3) == 0
                w = int(rng.integers(8, fw)) if big else int(rng.integers(1, 24))
                h = int(rng.integers(8, fh)) if big else int(rng.integers(1, 24))
                x0, y0 = int(rng.integers(-4, fw - 2)), int(rng.integers(-4, fh - 2))
                z = int(rng.integers(0, zmax + 1))
                writes += [(0x00, prim_reg(6, iip=0, abe=abe)), (0x01, rgbaq(int(rng.integers(0, 2**32)))),
                           (0x05, xyz2(ofs + x0 * 16, ofs + y0 * 16, z)),
                           (0x05, xyz2(ofs + (x0 + w) * 16, ofs + (y0 + h) * 16, z))]
            else:
                flat = kind == 1
                z = int(rng.integers(0, zmax + 1))
                writes.append((0x00, prim_reg(3, iip=int(not flat), abe=abe)))
                for v in range(3):
                    writes.append((0x01, rgbaq(int(rng.integers(0, 2**32)))))
                    writes.append((0x05, xyz2(ofs + int(rng.integers(-8 * 16, (fw + 8) * 16)),
                                              ofs + int(rng.integers(-8 * 16, (fh + 8) * 16)),
                                              z if flat else int(rng.integers(0, zmax + 1)))))
        for i in range(0, len(writes), 240):
            out.append(gif.ad(writes[i:i + 240]))
    return gif.arr(b"".join(out))


def random_ssaa_texture_stream(seed: int, n_batches: int = 6, sprites_per_batch: int = 10) -> np.ndarray:
    """Super-sampled textures: a 128 x 64 source frame (CT32, CT24 or CT16) drawn with 3D
    triangles (whose edges leave super samples), sometimes partly overwritten by an upload
    (those pages stop being super-sampled), then sprites into a second frame textured from the
    source in its own format or a 24-bit / 8H / 8-bit view, with NEAREST or LINEAR, at 1:1,
    near 1:1, 2x up and down, in UV and ST form, with CLAMP / REGION_CLAMP / REGION_REPEAT, and
    blur-like passes (one texture at several phases, Cs * alpha + Cd blending, no depth test)."""
    rng = np.random.default_rng(seed)
    ofs = 2048 * 16
    fw, fh = 128, 64
    dst_fbp, zbp, clut_bp = 8, 16, 2048
    out = [gif.ad([(0x1A, 1), (0x18, ofs | ofs << 32)])]
    out.append(gif.upload(0, 2, 0x00, 0, 0, fw, fh, _random_payload(rng, fw * fh * 4)))
    out.append(gif.upload(dst_fbp * 32, 2, 0x00, 0, 0, fw, fh, _random_payload(rng, fw * fh * 4)))
    out.append(gif.upload(clut_bp, 1, 0x00, 0, 0, 16, 16, _random_payload(rng, 16 * 16 * 4)))
    for b in range(n_batches):
        src_psm = int(rng.choice([0x00, 0x00, 0x01, 0x02]))
        out.append(gif.ad([(0x4C, frame(0, 2, src_psm)), (0x4E, zbuf(zbp, 0, 0)),
                           (0x40, scissor(0, fw - 1, 0, fh - 1)), (0x47, test_reg(zte=1, ztst=1)),
                           (0x42, alpha_reg(0, 1, 0, 1, 0))]))
        writes = []
        for t in range(int(rng.integers(4, 12))):
            writes.append((0x00, prim_reg(3, iip=1, abe=int(rng.integers(0, 4) == 0))))
            for v in range(3):
                writes.append((0x01, rgbaq(int(rng.integers(0, 2**32)))))
                writes.append((0x05, xyz2(ofs + int(rng.integers(-8 * 16, (fw + 8) * 16)),
                                          ofs + int(rng.integers(-8 * 16, (fh + 8) * 16)), 0)))
        out.append(gif.ad(writes))
        if rng.integers(0, 4) == 0:
            x, y = int(rng.integers(0, 2)) * 64, int(rng.integers(0, 2)) * 32
            out.append(gif.upload(0, 2, 0x00, x, y, 64, 32, _random_payload(rng, 64 * 32 * 4)))
        dst_psm = int(rng.choice([0x00, 0x00, 0x02]))
        blur = rng.integers(0, 3) == 0
        ztst = 1 if blur else int(rng.integers(1, 4))
        regs = [(0x4C, frame(dst_fbp, 2, dst_psm)), (0x40, scissor(0, fw - 1, 0, fh - 1)),
                (0x47, test_reg(zte=1, ztst=ztst)),
                (0x42, alpha_reg(0, 2, int(rng.integers(0, 3)), 1, int(rng.integers(0, 256))) if blur
                 else alpha_reg(*[int(rng.integers(0, 3)) for _ in range(4)], int(rng.integers(0, 256))))]
        out.append(gif.ad(regs))
        writes = []
        views = {0x00: [0x00, 0x01, 0x1B, 0x13], 0x01: [0x01, 0x1B], 0x02: [0x02]}[src_psm]
        mmag_pass = int(rng.integers(0, 2))
        for p in range(sprites_per_batch):
            tpsm = int(rng.choice(views))
            mmag = mmag_pass if blur else int(rng.integers(0, 2))
            wm = int(rng.choice([0, 0, 1, 2, 3]))
            if wm == 3:
                cl = clamp_reg(3, 3, int(rng.choice([0x3F, 0x7F, 0x1F])), 0, int(rng.choice([0x3F, 0x1F])), 0)
            elif wm == 2:
                lo, hi = sorted(int(v) for v in rng.integers(0, fw, 2))
                lov, hiv = sorted(int(v) for v in rng.integers(0, fh, 2))
                cl = clamp_reg(2, 2, lo, hi, lov, hiv)
            else:
                cl = clamp_reg(wm, wm)
            t0 = tex0(0, 2, tpsm, 7, 6, tcc=int(rng.integers(0, 2)), tfx=int(rng.choice([0, 1, 1])),
                      cbp=clut_bp, cpsm=0, csa=int(rng.integers(0, 8)), cld=int(rng.choice([0, 1])))
            fst = int(rng.integers(0, 5) != 0)
            writes += [(0x14, tex1(mmag, mmag)), (0x08, cl), (0x06, t0),
                       (0x00, prim_reg(6, iip=0, tme=1, abe=int(blur or rng.integers(0, 2)), fst=fst))]
            w = int(rng.choice([8, 16, 24, 32, 37, 64]))
            h = int(rng.choice([8, 16, 20, 32]))
            x0, y0 = int(rng.integers(0, fw - w + 1)), int(rng.integers(0, fh - h + 1))
            scale = float(rng.choice([1.0, 1.0, 0.5, 2.0, 0.97, 1.03]))
            d = int(rng.choice([0, 0, 3, 8, 15])) if not blur else int(rng.integers(0, 16))
            u0, v0 = int(rng.integers(0, fw - 8)) * 16 + d, int(rng.integers(0, fh - 8)) * 16 + d
            if blur:
                u0, v0 = x0 * 16 + d, y0 * 16 + d
            corners = [(x0, y0, u0, v0), (x0 + w, y0 + h, u0 + int(w * 16 * scale), v0 + int(h * 16 * scale))]
            q = float(rng.choice([1.0, 1.0, 2.0]))
            for (x, y, u, v) in corners:
                writes.append((0x01, rgbaq(int(rng.integers(0, 2**32)), q)))
                if fst:
                    writes.append((0x03, uv(u, v)))
                else:
                    writes.append((0x02, st(u / (fw * 16.0) * q, v / (fh * 16.0) * q)))
                writes.append((0x05, xyz2(ofs + x * 16, ofs + y * 16, int(rng.integers(0, 2**24)))))
        out.append(gif.ad(writes))
    return gif.arr(b"".join(out))
