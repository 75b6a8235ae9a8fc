// Texture rectangle, texel addressing and page footprints (host side).
//
// The effective rectangle follows compute_effective_texture_extent (gs/gs_util.cpp:31-79) for
// a single mip level; texel addressing follows the upload addressing
// (gs/shaders/upload.comp:121-123): ((x & MSK) + base) | FIX, with MSK / FIX only for
// REGION_REPEAT. Page spans give the render pass a conservative way to decide whether a
// texture read touches pages the pass writes (the ordering rule of gs/page_tracker.cpp).
#pragma once
#include "gs_defs.hpp"
#include "gs_raster.hpp"
#include <algorithm>

namespace pgs
{
inline uint32_t bits(uint64_t v, uint32_t off, uint32_t n) { return uint32_t((v >> off) & ((1ull << n) - 1ull)); }

struct TexLayout
{
	uint32_t tbp, tbw, psm, tw, th;
	uint32_t width, height, bx, by;
	uint32_t umsk, ufix, vmsk, vfix;
	uint32_t tex_x(uint32_t x) const { return ((x & umsk) + bx) | ufix; }
	uint32_t tex_y(uint32_t y) const { return ((y & vmsk) + by) | vfix; }
};

inline uint32_t effective_extent(uint32_t ext, uint32_t mode, uint32_t lo, uint32_t hi, uint32_t &base)
{
	base = 0;
	if (mode == 2) // REGION_CLAMP
	{
		lo = std::min(lo, hi);
		ext = std::max(hi, lo) - lo + 1u;
		base = lo;
	}
	else if (mode == 3 && lo == 0) // REGION_REPEAT with an empty mask: one texel column
	{
		ext = 1;
		base = hi;
	}
	else if (mode == 3)
	{
		// (x & MSK) | FIX equals (x & MSK) + FIX when FIX has no bit at or below MSK's top bit.
		const uint32_t msb = 31u - uint32_t(__builtin_clz(lo));
		const uint32_t lsb = hi ? uint32_t(__builtin_ctz(hi)) : 32u;
		if (lsb > msb)
		{
			ext = std::min(1u << (msb + 1u), ext);
			base = hi;
		}
	}
	return ext;
}

inline TexLayout tex_layout(uint64_t tex0, uint64_t clamp)
{
	TexLayout l{};
	l.tbp = bits(tex0, 0, 14);
	l.tbw = bits(tex0, 14, 6);
	l.psm = bits(tex0, 20, 6);
	l.tw = std::min(bits(tex0, 26, 4), 10u);
	l.th = std::min(bits(tex0, 30, 4), 10u);
	const uint32_t wms = bits(clamp, 0, 2), wmt = bits(clamp, 2, 2);
	const uint32_t minu = bits(clamp, 4, 10), maxu = bits(clamp, 14, 10);
	const uint32_t minv = bits(clamp, 24, 10), maxv = bits(clamp, 34, 10);
	l.width = effective_extent(1u << l.tw, wms, minu, maxu, l.bx);
	l.height = effective_extent(1u << l.th, wmt, minv, maxv, l.by);
	l.umsk = wms == 3 ? minu : 0xffffffffu;
	l.ufix = wms == 3 ? maxu : 0u;
	l.vmsk = wmt == 3 ? minv : 0xffffffffu;
	l.vfix = wmt == 3 ? maxv : 0u;
	return l;
}

// Inclusive interval of 8 KiB pages a rectangle of a buffer touches; `last` may run past the
// end of VRAM (addresses wrap). The page index (y >> page_h) * stride + (x >> page_w) grows
// with x and y, so the corners bound it; Z formats' block XOR stays inside a 32-block group
// and an unaligned base can spill into the next page.
struct PageSpan { uint32_t first, last; };

inline PageSpan page_span(uint32_t base_block, uint32_t stride, uint32_t psm, uint32_t x0, uint32_t y0, uint32_t x1,
                          uint32_t y1)
{
	const Geometry g = psm_geometry(psm);
	const uint32_t sc = psm_storage_class(psm);
	if (sc == SC_8 || sc == SC_4)
		stride >>= 1;
	const uint32_t lo = (y0 >> g.ph) * stride + (x0 >> g.pw);
	const uint32_t hi = (y1 >> g.ph) * stride + (x1 >> g.pw);
	return { (base_block >> 5) + lo, ((base_block + 31u) >> 5) + hi };
}

inline bool spans_overlap(PageSpan a, PageSpan b)
{
	const uint32_t pages = kNumPages;
	if (a.last - a.first + 1u >= pages || b.last - b.first + 1u >= pages)
		return true;
	const uint32_t a0 = a.first % pages, a1 = a0 + (a.last - a.first);
	const uint32_t b0 = b.first % pages, b1 = b0 + (b.last - b.first);
	auto hit = [](uint32_t p0, uint32_t p1, uint32_t q0, uint32_t q1) { return p0 <= q1 && q0 <= p1; };
	return hit(a0, a1, b0, b1) || hit(a0 + pages, a1 + pages, b0, b1) || hit(a0, a1, b0 + pages, b1 + pages);
}

inline PageSpan texture_span(const TexLayout &l)
{
	uint32_t x0 = ~0u, x1 = 0, y0 = ~0u, y1 = 0;
	for (uint32_t x = 0; x < l.width; x++)
	{
		x0 = std::min(x0, l.tex_x(x));
		x1 = std::max(x1, l.tex_x(x));
	}
	for (uint32_t y = 0; y < l.height; y++)
	{
		y0 = std::min(y0, l.tex_y(y));
		y1 = std::max(y1, l.tex_y(y));
	}
	return page_span(l.tbp, l.tbw, l.psm, x0, y0, x1, y1);
}

// Decode work for one texture of a render pass (the upload.comp push constants).
struct TexDecode
{
	uint32_t image_offset;   // dwords into the pass's image arena
	uint32_t width, height;
	uint32_t tbp, tbw, psm;
	uint32_t bx, by, umsk, ufix, vmsk, vfix;
	uint32_t clut16, csa;    // CLUT format class and bank offset (palette formats)
	uint32_t aem, ta0, ta1;  // TEXA expansion
	uint32_t group_begin;    // first 256-texel workgroup of this texture in the launch
};
static_assert(sizeof(TexDecode) == 72, "TexDecode layout");
}
