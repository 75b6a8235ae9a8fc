// CPU oracle for the GS rendering hot path: GIF parsing, register file, VRAM transfers.
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp). Exported through a plain C ABI so the
// Python tests and the benchmark's CPU-baseline leg can drive it with ctypes.
//
// Semantics restated (not copied) from the reference:
//   gs/gs_interface.cpp:4290-4397   gif_transfer (PACKED / REGLIST / IMAGE)
//   gs/gs_interface.cpp:3168-3276   init_transfer
//   gs/gs_interface.cpp:2979-3148   flush_pending_transfer (host->local)
//   gs/shaders/vram_copy.comp       per-pixel read/merge rules for every PSM
//   gs/shaders/swizzle_utils.h      swizzle_PS2
#include "oracle_core.hpp"
#include "oracle_render.hpp"
#include <cstring>
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <limits>

namespace orc
{
// gs/shaders/swizzle_utils.h:18-55
uint32_t bits_per_pixel(uint32_t psm)
{
	switch (psm)
	{
	case CT24: case Z24: return 24;
	case CT16: case CT16S: case Z16: case Z16S: return 16;
	case T8: case T8H: return 8;
	case T4: case T4HL: case T4HH: return 4;
	default: return 32;
	}
}

namespace
{
struct Layout { uint32_t pw_log2, ph_log2, bw_log2, bh_log2, ch_log2; };

// Page / block / column geometry per PSM (gs/shaders/swizzle_utils.h:93-160).
Layout layout_of(uint32_t psm)
{
	switch (psm)
	{
	case CT16: case CT16S: case Z16: case Z16S: return { 6, 6, 4, 3, 1 };
	case T8: return { 7, 6, 4, 4, 2 };
	case T4: return { 7, 7, 5, 4, 2 };
	default: return { 6, 5, 3, 3, 1 };
	}
}

inline uint32_t bit(uint32_t v, uint32_t b) { return (v >> b) & 1u; }
}

// Restates swizzle_PS2 (gs/shaders/swizzle_utils.h:329-522). Block order within a page is
// a bit interleave of the block coordinates; Z formats flip block bits 3 and 4 after the
// base offset is added; 8/4-bit formats interleave pixels inside a 64-byte column.
uint32_t swizzle(uint32_t x, uint32_t y, uint32_t base_block, uint32_t page_stride, uint32_t psm)
{
	const Layout L = layout_of(psm);
	if (psm == T4 || psm == T8)
		page_stride >>= 1;

	uint32_t page = (y >> L.ph_log2) * page_stride + (x >> L.pw_log2);
	uint32_t bx = (x & ((1u << L.pw_log2) - 1u)) >> L.bw_log2;
	uint32_t by = (y & ((1u << L.ph_log2) - 1u)) >> L.bh_log2;
	uint32_t col = (y & ((1u << L.bh_log2) - 1u)) >> L.ch_log2;
	uint32_t px = x & ((1u << L.bw_log2) - 1u);
	uint32_t py = y & ((1u << L.ch_log2) - 1u);

	uint32_t block;
	bool zfmt = psm == Z32 || psm == Z24 || psm == Z16 || psm == Z16S;
	if (psm == CT16 || psm == Z16 || psm == T4)
		block = bit(by, 0) | bit(bx, 0) << 1 | bit(by, 1) << 2 | bit(bx, 1) << 3 | bit(by, 2) << 4;
	else if (psm == CT16S || psm == Z16S)
		block = bit(by, 0) | bit(bx, 0) << 1 | bit(by, 2) << 2 | bit(by, 1) << 3 | bit(bx, 1) << 4;
	else
		block = bit(bx, 0) | bit(by, 0) << 1 | bit(bx, 1) << 2 | bit(by, 1) << 3 | bit(bx, 2) << 4;
	block += base_block;
	if (zfmt)
		block ^= 0x18u;

	uint32_t column_base = page * 8192u + block * 256u + col * 64u;

	switch (psm)
	{
	case CT16: case CT16S: case Z16: case Z16S:
	{
		uint32_t pi = bit(px, 3) | bit(px, 0) << 1 | bit(py, 0) << 2 | bit(px, 1) << 3 | bit(px, 2) << 4;
		return ((column_base + pi * 2u) & VRAM_MASK) >> 1;
	}
	case T8:
	{
		uint32_t hi = bit(px, 2) ^ bit(py, 1) ^ bit(col, 0);
		uint32_t pi = bit(py, 1) | bit(px, 3) << 1 | bit(px, 0) << 2 | bit(py, 0) << 3 | bit(px, 1) << 4 | hi << 5;
		return (column_base + pi) & VRAM_MASK;
	}
	case T4:
	{
		uint32_t hi = bit(px, 2) ^ bit(py, 1) ^ bit(col, 0);
		uint32_t pi = bit(py, 1) | bit(px, 3) << 1 | bit(px, 4) << 2 | bit(px, 0) << 3 | bit(py, 0) << 4 |
		              bit(px, 1) << 5 | hi << 6;
		return ((column_base << 1) + pi) & ((VRAM_MASK << 1) | 1u);
	}
	default:
	{
		uint32_t pi = bit(px, 0) | bit(py, 0) << 1 | bit(px, 1) << 2 | bit(px, 2) << 3;
		return ((column_base + pi * 4u) & VRAM_MASK) >> 2;
	}
	}
}

// ---------------------------------------------------------------------------------------
// VRAM element access helpers on a single 4 MiB plane.
namespace
{
inline uint32_t rd32(const uint8_t *v, uint32_t word) { uint32_t r; memcpy(&r, v + word * 4u, 4); return r; }
inline void wr32(uint8_t *v, uint32_t word, uint32_t d) { memcpy(v + word * 4u, &d, 4); }
inline uint32_t rd16(const uint8_t *v, uint32_t hw) { uint16_t r; memcpy(&r, v + hw * 2u, 2); return r; }
inline void wr16(uint8_t *v, uint32_t hw, uint32_t d) { uint16_t s = uint16_t(d); memcpy(v + hw * 2u, &s, 2); }
}

// Reads one source pixel of a local->local copy with the source PSM's rules
// (gs/shaders/vram_copy.comp:199-231).
uint32_t read_local_pixel(const uint8_t *v, uint32_t psm, uint32_t x, uint32_t y, uint32_t base, uint32_t stride)
{
	uint32_t a = swizzle(x, y, base, stride, psm);
	switch (psm)
	{
	case CT16: case CT16S: case Z16: case Z16S: return rd16(v, a);
	case T8: return v[a];
	case T8H: return v[a * 4u + 3u];
	case T4: return (v[a >> 1] >> ((a & 1u) * 4u)) & 0xfu;
	case T4HL: return (rd32(v, a) >> 24) & 0xfu;
	case T4HH: return (rd32(v, a) >> 28) & 0xfu;
	default: return rd32(v, a);
	}
}

// Writes one destination pixel with the destination PSM's lane mask
// (gs/shaders/vram_copy.comp:328-613; gs/gs_util.hpp:112-205).
void write_local_pixel(uint8_t *v, uint32_t psm, uint32_t x, uint32_t y, uint32_t base, uint32_t stride, uint32_t d)
{
	uint32_t a = swizzle(x, y, base, stride, psm);
	switch (psm)
	{
	case CT24: case Z24:
		wr32(v, a, (rd32(v, a) & 0xff000000u) | (d & 0xffffffu));
		break;
	case CT16: case CT16S: case Z16: case Z16S:
		wr16(v, a, d);
		break;
	case T8:
		v[a] = uint8_t(d);
		break;
	case T8H:
		v[a * 4u + 3u] = uint8_t(d);
		break;
	case T4:
	{
		uint8_t &b = v[a >> 1];
		if (a & 1u)
			b = uint8_t((b & 0x0fu) | ((d & 0xfu) << 4));
		else
			b = uint8_t((b & 0xf0u) | (d & 0xfu));
		break;
	}
	case T4HL:
		wr32(v, a, (rd32(v, a) & 0xf0ffffffu) | ((d & 0xfu) << 24));
		break;
	case T4HH:
		wr32(v, a, (rd32(v, a) & 0x0fffffffu) | ((d & 0xfu) << 28));
		break;
	default:
		wr32(v, a, d);
		break;
	}
}

// Host payload element for destination PSM (gs/shaders/vram_copy.comp:232-256):
// pixels are packed at the destination's bits-per-pixel, nibbles low-first.
static bool read_host_pixel(const uint8_t *p, size_t bytes, uint32_t psm, uint32_t idx, uint32_t &out)
{
	switch (bits_per_pixel(psm))
	{
	case 24:
		if (idx >= bytes / 3) return false;
		out = p[idx * 3] | (p[idx * 3 + 1] << 8) | (p[idx * 3 + 2] << 16);
		return true;
	case 16:
		if (idx >= bytes / 2) return false;
		out = p[idx * 2] | (p[idx * 2 + 1] << 8);
		return true;
	case 8:
		if (idx >= bytes) return false;
		out = p[idx];
		return true;
	case 4:
		if (idx >= bytes * 2) return false;
		out = (p[idx >> 1] >> ((idx & 1u) * 4u)) & 0xfu;
		return true;
	default:
		if (idx >= bytes / 4) return false;
		memcpy(&out, p + idx * 4, 4);
		return true;
	}
}

// The GPU copy skips pixels whose X lies beyond the destination stride, except on the last
// page row of the transfer (gs/shaders/vram_copy.comp:282-300).
static bool stride_skip(uint32_t psm, uint32_t dbw, uint32_t dsay, uint32_t h, uint32_t cx, uint32_t cy)
{
	uint32_t eff = dbw;
	if (psm == T4 || psm == T8)
		eff &= ~1u;
	if (cx < eff * 64u)
		return false;
	uint32_t ph_log2 = (psm == CT16 || psm == CT16S || psm == Z16 || psm == Z16S || psm == T8) ? 6u :
	                   (psm == T4 ? 7u : 5u);
	uint32_t last_page = (dsay + h - 1u) >> ph_log2;
	uint32_t cur_page = (dsay + cy) >> ph_log2;
	return last_page != cur_page;
}

// ---------------------------------------------------------------------------------------
// GS state machine.

GS::GS(uint32_t ssaa_log2)
{
	sample_log2 = ssaa_log2;
	uint32_t planes = ssaa_log2 ? (2u + (1u << ssaa_log2)) : 1u;
	vram.assign(size_t(planes) * VRAM_SIZE, 0);
	clut.assign(512u, 0);
	reset_registers();
}

void GS::reset_registers()
{
	memset(&regs, 0, sizeof(regs));
	regs.rgbaq_q = 0.0f; // RGBAQ is zero after reset (RegisterState{}, gs/gs_interface.hpp:36-61)
	internal_q = 1.0f;
	for (auto &p : paths)
		p = {};
	vq_count = 0;
	xfer = {};
}

void GS::reset_context_state()
{
	flush();
	memset(&regs, 0, sizeof(regs));
	internal_q = 1.0f;
	for (auto &p : paths)
		p = {};
	vq_count = 0;
	cached_cbp[0] = cached_cbp[1] = 0;
	state_dirty = true;
}

void GS::set_reg(uint64_t &reg, uint64_t v, bool marks_dirty)
{
	if (reg != v)
	{
		reg = v;
		if (marks_dirty)
			state_dirty = true;
	}
}

// A+D register writes (gs/gs_interface.cpp:3507-3986). Registers that feed primitive state
// mark the state dirty when their value changes (update_internal_register,
// gs_interface.cpp:3877-3885); a dirty state blocks parallelogram fusion of the next triangle.
void GS::write_register(uint32_t addr, uint64_t v)
{
	switch (addr)
	{
	case 0x00:
		// gs/gs_interface.cpp:3507-3556: PRMODECONT.AC = 0 keeps every PRIM bit but the type.
		if (regs.prmodecont & 1u)
			set_reg(regs.prim, v & 0x7ffu, true);
		else
		{
			if ((regs.prim & 7u) != (v & 7u) && ((v >> 7) & 1u))
				state_dirty = true;
			regs.prim = (regs.prim & ~7ull) | (v & 7u);
		}
		vq_count = 0;
		break;
	case 0x01: case 0x11:
		regs.rgbaq = uint32_t(v);
		{ uint32_t q = uint32_t(v >> 32); memcpy(&regs.rgbaq_q, &q, 4); }
		break;
	case 0x02: { uint32_t s = uint32_t(v), t = uint32_t(v >> 32); memcpy(&regs.s, &s, 4); memcpy(&regs.t, &t, 4); } break;
	case 0x03: regs.uv = v; break;
	case 0x04: kick_xyzf(v, false); break;
	case 0x05: kick_xyz(v, false); break;
	case 0x0c: kick_xyzf(v, true); break;
	case 0x0d: kick_xyz(v, true); break;
	case 0x06: case 0x07: write_tex0(addr - 0x06, v); break;
	case 0x08: case 0x09: set_reg(regs.clamp[addr - 0x08], v, true); break;
	case 0x0a: regs.fog = uint32_t(v >> 56); break;
	case 0x14: case 0x15: set_reg(regs.tex1[addr - 0x14], v, true); break;
	case 0x16: case 0x17: write_tex2(addr - 0x16, v); break;
	case 0x18: case 0x19: regs.xyoffset[addr - 0x18] = v; break;
	case 0x1a: regs.prmodecont = v; break;
	case 0x1b:
		// gs/gs_interface.cpp:3692-3722: PRMODE supplies PRIM's other bits while AC = 0.
		regs.prmode = v;
		if (!(regs.prmodecont & 1u))
			set_reg(regs.prim, (v & 0x7f8u) | (regs.prim & 7u), true);
		break;
	case 0x1c: regs.texclut = v; break;
	case 0x22: set_reg(regs.scanmsk, v, true); break;
	case 0x34: case 0x35: set_reg(regs.miptbp1[addr - 0x34], v, true); break;
	case 0x36: case 0x37: set_reg(regs.miptbp2[addr - 0x36], v, true); break;
	case 0x3b: set_reg(regs.texa, v, true); break;
	case 0x3d: regs.fogcol = v; break;
	case 0x40: case 0x41: set_reg(regs.scissor[addr - 0x40], v, true); break; // STATE_DIRTY_SCISSOR_BIT
	case 0x42: case 0x43: set_reg(regs.alpha[addr - 0x42], v, true); break;
	case 0x44: set_reg(regs.dimx, v, true); break;
	case 0x45: set_reg(regs.dthe, v, true); break;
	case 0x46: set_reg(regs.colclamp, v, true); break;
	case 0x47: case 0x48: set_reg(regs.test[addr - 0x47], v, true); break;
	case 0x49: set_reg(regs.pabe, v, true); break;
	case 0x4a: case 0x4b: set_reg(regs.fba[addr - 0x4a], v, true); break;
	case 0x4c: case 0x4d: set_reg(regs.frame[addr - 0x4c], v, true); break;
	case 0x4e: case 0x4f: set_reg(regs.zbuf[addr - 0x4e], v, true); break;
	case 0x50: regs.bitbltbuf = v; break;
	case 0x51: regs.trxpos = v; break;
	case 0x52: regs.trxreg = v; break;
	case 0x53: regs.trxdir = v; init_transfer(); break;
	case 0x54: hwreg(&v, 1); break;
	default: break;
	}
}

// gs/gs_interface.cpp:3168-3276
void GS::init_transfer()
{
	finish_host_transfer();
	uint64_t b = regs.bitbltbuf, p = regs.trxpos, r = regs.trxreg;
	uint32_t xdir = uint32_t(regs.trxdir & 3u);
	Copy c;
	c.sbp = uint32_t(b & 0x3fff);
	c.sbw = uint32_t((b >> 16) & 0x3f);
	c.spsm = uint32_t((b >> 24) & 0x3f);
	c.dbp = uint32_t((b >> 32) & 0x3fff);
	c.dbw = uint32_t((b >> 48) & 0x3f);
	c.dpsm = uint32_t((b >> 56) & 0x3f);
	c.ssax = uint32_t(p & 0x7ff);
	c.ssay = uint32_t((p >> 16) & 0x7ff);
	c.dsax = uint32_t((p >> 32) & 0x7ff);
	c.dsay = uint32_t((p >> 48) & 0x7ff);
	c.w = uint32_t(r & 0xfff);
	c.h = uint32_t((r >> 32) & 0xfff);

	// Transfers leave their destination pages single-sampled (gs/page_tracker.cpp:255-265,
	// 650-660); the draws before them marked their pages as they were drawn.
	if ((xdir == 0 || xdir == 2) && c.w && c.h)
		mark_pages(c.dbp, c.dbw, c.dpsm, c.dsax, c.dsay, c.dsax + c.w - 1u, c.dsay + c.h - 1u, false);
	if (xdir == 2)
	{
		flush_render_pass();
		local_copy(c);
	}
	else if (xdir == 1)
	{
		flush_render_pass();
		local_to_host(c);
	}
	else if (xdir == 0)
	{
		xfer.copy = c;
		xfer.required_qwords = uint32_t((uint64_t(c.w) * c.h * bits_per_pixel(c.dpsm)) / 64u);
		xfer.active = xfer.required_qwords != 0;
		xfer.payload.clear();
		xfer.written_pixels = 0;
	}
}

// Local->host readback into the FIFO, pixels in raster order packed at the source PSM's
// bits per pixel; 24-bit formats store 3 bytes per pixel; 4-bit formats are not read back
// (their FIFO stays zero) (gs/gs_util.hpp:49-110 vram_readback; gs/gs_interface.cpp:3213-3275).
void GS::local_to_host(const Copy &c)
{
	const uint32_t bytes = uint32_t((uint64_t(c.w) * c.h * bits_per_pixel(c.spsm)) / 8u);
	fifo.assign(bytes, 0);
	fifo_offset_128 = 0;
	const uint8_t *v = vram.data();
	for (uint32_t y = 0; y < c.h; y++)
		for (uint32_t x = 0; x < c.w; x++)
		{
			const size_t i = size_t(y) * c.w + x;
			const uint32_t px = (x + c.ssax) & 2047u, py = (y + c.ssay) & 2047u;
			switch (c.spsm)
			{
			case CT32: case Z32:
			{
				uint32_t d = read_local_pixel(v, c.spsm, px, py, c.sbp, c.sbw);
				memcpy(&fifo[i * 4], &d, 4);
				break;
			}
			case CT24: case Z24:
			{
				const uint32_t d = read_local_pixel(v, c.spsm, px, py, c.sbp, c.sbw); // whole word
				if (i * 3 + 3 <= bytes)
					for (uint32_t k = 0; k < 3; k++)
						fifo[i * 3 + k] = uint8_t(d >> (8 * k));
				break;
			}
			case CT16: case CT16S: case Z16: case Z16S:
			{
				uint32_t d = read_local_pixel(v, c.spsm, px, py, c.sbp, c.sbw);
				fifo[i * 2] = uint8_t(d);
				fifo[i * 2 + 1] = uint8_t(d >> 8);
				break;
			}
			case T8: case T8H:
				fifo[i] = uint8_t(read_local_pixel(v, c.spsm, px, py, c.sbp, c.sbw));
				break;
			default:
				break;
			}
		}
}

void GS::read_transfer_fifo(void *data, uint32_t n)
{
	const uint32_t size_128 = uint32_t(fifo.size() / 16u);
	const uint32_t avail = size_128 - std::min(size_128, fifo_offset_128);
	const uint32_t copy = std::min(n, avail);
	if (copy)
		memcpy(data, fifo.data() + 16u * fifo_offset_128, copy * 16u);
	if (copy < n)
		memset(static_cast<uint8_t *>(data) + 16u * copy, 0, (n - copy) * 16u);
	fifo_offset_128 += copy;
}

// Local->local: the whole source rectangle is read before any destination pixel is
// written (gs/gs_renderer.cpp:3921-3976 shadow-page semantics + page_tracker.cpp:221-325
// flush-on-hazard ordering).
void GS::local_copy(const Copy &c)
{
	if (xfer_hook)
		xfer_hook(xfer_hook_user, 2, c.w, c.h, c.dpsm);
	std::vector<uint32_t> src(size_t(c.w) * c.h);
	for (uint32_t y = 0; y < c.h; y++)
		for (uint32_t x = 0; x < c.w; x++)
			src[size_t(y) * c.w + x] = read_local_pixel(vram.data(), c.spsm, (x + c.ssax) & 2047u,
			                                            (y + c.ssay) & 2047u, c.sbp, c.sbw);
	for (uint32_t y = 0; y < c.h; y++)
		for (uint32_t x = 0; x < c.w; x++)
		{
			if (stride_skip(c.dpsm, c.dbw, c.dsay, c.h, x, y))
				continue;
			write_local_pixel(vram.data(), c.dpsm, (x + c.dsax) & 2047u, (y + c.dsay) & 2047u,
			                  c.dbp, c.dbw, src[size_t(y) * c.w + x]);
			invalidate_ssaa(c.dpsm, (x + c.dsax) & 2047u, (y + c.dsay) & 2047u, c.dbp, c.dbw);
		}
}

// Host->local: pixels are written in raster order as the payload arrives. A partial flush
// writes the pixels whose data is complete; the remainder follows when more data arrives
// (gs/gs_interface.cpp:2970-3148; vram_copy.comp host_offset / source_size guards).
void GS::hwreg(const uint64_t *data, size_t count)
{
	if (!xfer.active)
		return;
	xfer.payload.insert(xfer.payload.end(), data, data + count);
	if (xfer.payload.size() >= xfer.required_qwords)
		finish_host_transfer();
}

void GS::write_host_pixels()
{
	flush_render_pass();
	const Copy &c = xfer.copy;
	const uint8_t *p = reinterpret_cast<const uint8_t *>(xfer.payload.data());
	size_t bytes = xfer.payload.size() * 8u;
	uint32_t total = c.w * c.h;
	uint32_t idx = xfer.written_pixels;
	for (; idx < total; idx++)
	{
		uint32_t d;
		if (!read_host_pixel(p, bytes, c.dpsm, idx, d))
			break;
		uint32_t x = idx % c.w, y = idx / c.w;
		if (stride_skip(c.dpsm, c.dbw, c.dsay, c.h, x, y))
			continue;
		uint32_t vx = (x + c.dsax) & 2047u, vy = (y + c.dsay) & 2047u;
		write_local_pixel(vram.data(), c.dpsm, vx, vy, c.dbp, c.dbw, d);
		invalidate_ssaa(c.dpsm, vx, vy, c.dbp, c.dbw);
	}
	xfer.written_pixels = idx;
}

void GS::finish_host_transfer()
{
	if (!xfer.active)
		return;
	if (xfer_hook)
		xfer_hook(xfer_hook_user, 0, xfer.copy.w, xfer.copy.h, xfer.copy.dpsm);
	write_host_pixels();
	xfer.active = false;
	xfer.payload.clear();
	xfer.written_pixels = 0;
}

void GS::partial_flush_transfer()
{
	if (xfer.active)
		write_host_pixels();
}

// Transfers clear the "super-sampled valid" reference plane of what they overwrite, so a
// later draw reads the single-sampled value (vram_copy.comp INVALIDATE_SUPER_SAMPLE path).
void GS::invalidate_ssaa(uint32_t psm, uint32_t x, uint32_t y, uint32_t base, uint32_t stride)
{
	if (!sample_log2)
		return;
	uint8_t *ref = vram.data() + VRAM_SIZE;
	write_local_pixel(ref, psm, x, y, base, stride, 0);
}

// gs/gs_interface.cpp:4290-4397
void GS::gif_transfer(uint32_t path_index, const void *data, size_t size)
{
	auto &path = paths[path_index & 3u];
	size /= 16;
	const uint64_t *q = static_cast<const uint64_t *>(data);

	for (size_t i = 0; i < size;)
	{
		if (path.loop == path.nloop)
		{
			uint64_t lo = q[2 * i], hi = q[2 * i + 1];
			path.nloop = uint32_t(lo & 0x7fff);
			path.flg = uint32_t((lo >> 58) & 3u);
			path.nreg = uint32_t((lo >> 60) & 0xf);
			path.regs = hi;
			bool pre = ((lo >> 46) & 1u) != 0;
			if (path.flg == 0 && pre && path.nloop)
				write_register(0x00, (lo >> 47) & 0x7ffu);
			path.loop = 0;
			path.reg = 0;
			if (path.nloop)
				internal_q = 1.0f;
			i++;
			continue;
		}

		uint32_t nreg = path.nreg ? path.nreg : 16u;
		if (path.flg == 0)
		{
			uint32_t addr = uint32_t(path.regs >> (4u * path.reg)) & 0xfu;
			packed_write(addr, q + 2 * i);
			i++;
			if (++path.reg == nreg)
			{
				path.reg = 0;
				path.loop++;
			}
		}
		else if (path.flg == 1)
		{
			for (uint32_t j = 0; j < 2; j++)
			{
				uint32_t addr = uint32_t(path.regs >> (4u * path.reg)) & 0xfu;
				if (addr != 0xe && addr != 0xf)
					write_register(addr, q[2 * i + j]);
				if (++path.reg == nreg)
				{
					path.reg = 0;
					path.loop++;
					if (path.loop == path.nloop)
						break;
				}
			}
			i++;
		}
		else
		{
			size_t n = std::min<size_t>(size - i, path.nloop - path.loop);
			hwreg(q + 2 * i, n * 2);
			i += n;
			path.loop += uint32_t(n);
		}
	}
}

// PACKED register formats (gs/gs_interface.cpp:3992-4138).
void GS::packed_write(uint32_t addr, const uint64_t *w)
{
	uint64_t lo = w[0], hi = w[1];
	switch (addr)
	{
	case 0x0: write_register(0x00, lo & 0x7ffu); break;
	case 0x1:
	{
		uint32_t r = uint32_t(lo & 0xff), g = uint32_t((lo >> 32) & 0xff);
		uint32_t b = uint32_t(hi & 0xff), a = uint32_t((hi >> 32) & 0xff);
		regs.rgbaq = r | g << 8 | b << 16 | a << 24;
		regs.rgbaq_q = internal_q;
		break;
	}
	case 0x2:
	{
		uint32_t s = uint32_t(lo), t = uint32_t(lo >> 32), qq = uint32_t(hi);
		memcpy(&regs.s, &s, 4);
		memcpy(&regs.t, &t, 4);
		memcpy(&internal_q, &qq, 4);
		break;
	}
	case 0x3:
		regs.uv = (lo & 0x3fffu) | (((lo >> 32) & 0x3fffu) << 16);
		break;
	case 0x4: case 0xc:
	{
		uint64_t x = lo & 0xffff, y = (lo >> 32) & 0xffff;
		uint64_t z = (hi >> 4) & 0xffffff, f = (hi >> 36) & 0xff;
		bool adc = addr == 0xc || ((hi >> 47) & 1u);
		kick_xyzf(x | y << 16 | z << 32 | f << 56, adc);
		break;
	}
	case 0x5: case 0xd:
	{
		uint64_t x = lo & 0xffff, y = (lo >> 32) & 0xffff;
		uint64_t z = hi & 0xffffffffu;
		bool adc = addr == 0xd || ((hi >> 47) & 1u);
		kick_xyz(x | y << 16 | z << 32, adc);
		break;
	}
	case 0x6: case 0x7: case 0x8: case 0x9:
		write_register(addr, lo);
		break;
	case 0xa:
		regs.fog = uint32_t((hi >> 36) & 0xff);
		break;
	case 0xe:
		write_register(uint32_t(hi & 0x7f), lo);
		break;
	default:
		break;
	}
}
}

// ---------------------------------------------------------------------------------------
// C ABI for tests / benchmark CPU leg.
extern "C"
{
void *orc_create(uint32_t ssaa_log2) { return new orc::GS(ssaa_log2); }
void orc_destroy(void *g) { delete static_cast<orc::GS *>(g); }
void orc_gif_transfer(void *g, uint32_t path, const void *data, size_t size)
{
	static_cast<orc::GS *>(g)->gif_transfer(path, data, size);
}
void orc_write_register(void *g, uint32_t addr, uint64_t v) { static_cast<orc::GS *>(g)->write_register(addr, v); }
void orc_flush(void *g) { static_cast<orc::GS *>(g)->flush(); }
const uint8_t *orc_vram(void *g) { return static_cast<orc::GS *>(g)->vram.data(); }
size_t orc_vram_size(void *g) { return static_cast<orc::GS *>(g)->vram.size(); }
void orc_set_vram(void *g, const void *data, size_t size)
{
	auto *gs = static_cast<orc::GS *>(g);
	memcpy(gs->vram.data(), data, std::min(size, gs->vram.size()));
}
uint32_t orc_swizzle(uint32_t x, uint32_t y, uint32_t base, uint32_t stride, uint32_t psm)
{
	return orc::swizzle(x, y, base, stride, psm);
}
uint64_t orc_stat(void *g, uint32_t which) { return static_cast<orc::GS *>(g)->stat(which); }
// GSInterface::reset_context_state (gs/gs_interface.cpp:26-39): pending work is flushed (a
// host->local transfer keeps what arrived and stays active), then the register file, GIF paths,
// vertex queue and cached CLUT base pointers are reset.
void orc_reset_context_state(void *g) { static_cast<orc::GS *>(g)->reset_context_state(); }
void orc_read_transfer_fifo(void *g, void *data, uint32_t n) { static_cast<orc::GS *>(g)->read_transfer_fifo(data, n); }
// Host VRAM write (gs/gs_interface.cpp:4184-4228): the bytes land in plane 0 after all earlier
// work; every 8 KiB page the range touches loses its super-samples (reference plane cleared,
// gs/gs_renderer.cpp:1478-1492).
void orc_host_write(void *g, size_t offset, const void *data, size_t size)
{
	auto *gs = static_cast<orc::GS *>(g);
	if (!size || offset + size > orc::VRAM_SIZE)
		return;
	gs->flush();
	memcpy(gs->vram.data() + offset, data, size);
	for (size_t page = offset / 8192u; page <= (offset + size - 1u) / 8192u; page++)
		gs->may_super_sample[page % 512u] = 0; // gs/page_tracker.cpp:700-711
	if (gs->sample_log2)
	{
		const size_t first = offset / 8192u * 8192u, last = ((offset + size - 1u) / 8192u + 1u) * 8192u;
		memset(gs->vram.data() + orc::VRAM_SIZE + first, 0, last - first);
	}
}
// GSInterface::set_super_sampling_rate (gs/gs_interface.cpp:65-118): pending work is done at the
// old rate, then every plane above plane 0 is cleared (gs/gs_renderer.cpp:315-342).
void orc_set_sparse_grid(void *g, uint32_t sparse) { static_cast<orc::GS *>(g)->sparse_grid = sparse != 0; }
void orc_phase_lut(uint32_t rate_x, uint32_t rate_y, uint32_t *out)
{
	std::vector<uint32_t> t;
	orc::phase_lut(rate_x, rate_y, t);
	std::copy(t.begin(), t.end(), out);
}
void orc_set_super_sampled_textures(void *g, uint32_t on)
{
	auto *gs = static_cast<orc::GS *>(g);
	gs->flush();
	gs->super_sampled_textures = on != 0;
}
void orc_set_sample_log2(void *g, uint32_t sample_log2)
{
	auto *gs = static_cast<orc::GS *>(g);
	gs->flush();
	gs->sample_log2 = sample_log2;
	const size_t planes = sample_log2 ? 2u + (1u << sample_log2) : 1u;
	gs->vram.resize(planes * orc::VRAM_SIZE);
	std::fill(gs->vram.begin() + orc::VRAM_SIZE, gs->vram.end(), uint8_t(0));
}
}

// Exhaustive sweeps of the reciprocals for the tests: the hash is the wrapping sum over inputs of
// ((output bits << 32) | input bits) * 0x9E3779B97F4A7C15, which the device evaluation kernel
// (pgs_k_rcp_eval) computes the same way; max_ulp is the largest |result - 1/x| in units of the
// last place of the correctly rounded float 1/x.
static double ulps_off(float r, double exact)
{
	const float f = float(exact);
	const double ulp = double(std::nextafter(f, std::numeric_limits<float>::infinity())) - double(f);
	return std::fabs(double(r) - exact) / ulp;
}

extern "C"
{
void orc_rcp_fixed_sweep(uint32_t first, uint64_t count, uint64_t *hash, double *max_ulp)
{
	uint64_t h = 0;
	double worst = 0.0;
	for (uint64_t k = 0; k < count; k++)
	{
		const uint32_t b = uint32_t(first + k);
		const float r = orc::rcp_fixed(b);
		uint32_t bits;
		memcpy(&bits, &r, 4);
		h += ((uint64_t(bits) << 32) | b) * 0x9E3779B97F4A7C15ull;
		worst = std::max(worst, ulps_off(r, 1.0 / double(b)));
	}
	*hash = h;
	*max_ulp = worst;
}

void orc_rcp_float_sweep(uint32_t first_bits, uint64_t count, uint64_t *hash, double *max_ulp)
{
	uint64_t h = 0;
	double worst = 0.0;
	for (uint64_t k = 0; k < count; k++)
	{
		const uint32_t u = uint32_t(first_bits + k);
		float v;
		memcpy(&v, &u, 4);
		const float r = orc::rcp_float(v);
		uint32_t bits;
		memcpy(&bits, &r, 4);
		h += ((uint64_t(bits) << 32) | u) * 0x9E3779B97F4A7C15ull;
		worst = std::max(worst, ulps_off(r, 1.0 / double(v)));
	}
	*hash = h;
	*max_ulp = worst;
}

// The oracle's Q-derived mipmap LOD over consecutive float bit patterns, hashed like the device
// evaluation kernel (pgs_k_lod_eval); one evaluation through orc_lod_from_q.
void orc_lod_sweep(uint32_t l, int32_t k, uint32_t first_bits, uint64_t count, uint64_t *hash)
{
	uint64_t h = 0;
	for (uint64_t i = 0; i < count; i++)
	{
		const uint32_t u = uint32_t(first_bits + i);
		float q;
		memcpy(&q, &u, 4);
		int32_t n;
		orc::lod_from_q(q, k, l, n);
		h += ((uint64_t(uint32_t(n)) << 32) | u) * 0x9E3779B97F4A7C15ull;
	}
	*hash = h;
}

int32_t orc_lod_from_q(float q, int32_t k, uint32_t l, int32_t *positive)
{
	int32_t n;
	*positive = orc::lod_from_q(q, k, l, n) ? 1 : 0;
	return n;
}

const uint8_t *orc_rcp_table() { return orc::rcp_table(); }
float orc_rcp_fixed(uint32_t v) { return orc::rcp_fixed(v); }
float orc_rcp_float(float v) { return orc::rcp_float(v); }
}
