// Texture rectangle, texel addressing and page footprints (host side).
//
// The effective rectangle follows compute_effective_texture_extent (gs/gs_util.cpp:31-79) and
// compute_effective_texture_rect (gs/gs_renderer.cpp:1235-1265), mip levels included; texel
// addressing follows the upload addressing (gs/shaders/upload.comp:121-123): ((x & MSK) + base)
// | FIX, with MSK / FIX only for REGION_REPEAT, and level l of a mipmapped texture reads from
// its own TBP / TBW with base, MSK and FIX shifted right by l (gs/gs_renderer.cpp:3458-3598).
// Page spans give the render pass a conservative way to decide whether a texture read touches
// pages the pass writes (the ordering rule of gs/page_tracker.cpp).
#pragma once
#include "gs_defs.hpp"
#include "gs_raster.hpp"
#include <algorithm>

namespace pgs
{
inline uint32_t bits(uint64_t v, uint32_t off, uint32_t n) { return uint32_t((v >> off) & ((1ull << n) - 1ull)); }

constexpr uint32_t kMaxTexLevels = 7;

struct TexLayout
{
	uint32_t tbp, tbw, psm, tw, th;
	uint32_t width, height, bx, by;
	uint32_t umsk, ufix, vmsk, vfix;
	uint32_t levels;                     // mip levels of the image (1 without mipmapping)
	uint32_t level_tbp[kMaxTexLevels], level_tbw[kMaxTexLevels];
	uint32_t level_width(uint32_t l) const { return std::max(width >> l, 1u); }
	uint32_t level_height(uint32_t l) const { return std::max(height >> l, 1u); }
	uint32_t texels() const
	{
		uint32_t n = 0;
		for (uint32_t l = 0; l < levels; l++)
			n += level_width(l) * level_height(l);
		return n;
	}
	uint32_t tex_x(uint32_t x, uint32_t l = 0) const { return ((x & (umsk >> l)) + (bx >> l)) | (ufix >> l); }
	uint32_t tex_y(uint32_t y, uint32_t l = 0) const { return ((y & (vmsk >> l)) + (by >> l)) | (vfix >> l); }
};

inline uint32_t effective_extent(uint32_t ext, uint32_t mode, uint32_t lo, uint32_t hi, uint32_t levels,
                                 uint32_t &base)
{
	base = 0;
	if (mode == 2) // REGION_CLAMP
	{
		lo = std::min(lo, hi);
		ext = std::max(hi, lo) - lo + 1u;
		if (levels > 1)
		{
			// Every level's rectangle starts and ends on a whole texel: the start is rounded
			// down and the extent up to a multiple of 2^(levels - 1).
			const uint32_t mask = (1u << (levels - 1u)) - 1u;
			ext += lo & mask;
			lo &= ~mask;
			ext = (ext + mask) & ~mask;
		}
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

// TEX1.MMIN selects a mipmapped minification filter (gs/gs_registers.hpp:94).
inline bool mmin_has_mipmap(uint64_t tex1)
{
	const uint32_t mmin = bits(tex1, 6, 3);
	return mmin >= 2u && mmin <= 5u;
}

// Mip levels the texture image holds (gs/gs_renderer.cpp:1240-1243, gs/gs_interface.cpp:1539-1540):
// MXL + 1, at most one more than the smaller size's log2 and at most 7; one level when MMIN
// has no mipmap or mipmaps are disabled.
inline uint32_t texture_levels(uint64_t tex0, uint64_t tex1, bool disable_mipmaps)
{
	if (!mmin_has_mipmap(tex1) || disable_mipmaps)
		return 1;
	const uint32_t tw = std::min(bits(tex0, 26, 4), 10u), th = std::min(bits(tex0, 30, 4), 10u);
	return std::min(std::min(bits(tex1, 2, 3) + 1u, std::min(tw, th) + 1u), kMaxTexLevels);
}

// MIPTBP1 holds levels 1..3, MIPTBP2 levels 4..6, each as TBP (14 bits) and TBW (6 bits).
inline TexLayout tex_layout(uint64_t tex0, uint64_t clamp, uint32_t levels = 1, uint64_t mip1 = 0, uint64_t mip2 = 0)
{
	TexLayout l{};
	l.tbp = bits(tex0, 0, 14);
	l.tbw = bits(tex0, 14, 6);
	l.psm = bits(tex0, 20, 6);
	l.tw = std::min(bits(tex0, 26, 4), 10u);
	l.th = std::min(bits(tex0, 30, 4), 10u);
	l.levels = levels;
	l.level_tbp[0] = l.tbp;
	l.level_tbw[0] = l.tbw;
	for (uint32_t i = 1; i < kMaxTexLevels; i++)
	{
		const uint64_t m = i <= 3 ? mip1 : mip2;
		const uint32_t off = ((i - 1u) % 3u) * 20u;
		l.level_tbp[i] = i < levels ? bits(m, off, 14) : 0u;
		l.level_tbw[i] = i < levels ? bits(m, off + 14u, 6) : 0u;
	}
	const uint32_t wms = bits(clamp, 0, 2), wmt = bits(clamp, 2, 2);
	const uint32_t minu = bits(clamp, 4, 10), maxu = bits(clamp, 14, 10);
	const uint32_t minv = bits(clamp, 24, 10), maxv = bits(clamp, 34, 10);
	l.width = effective_extent(1u << l.tw, wms, minu, maxu, levels, l.bx);
	l.height = effective_extent(1u << l.th, wmt, minv, maxv, levels, l.by);
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

// Rectangle of 8 KiB pages a pixel rectangle of a buffer touches, as the page tracker walks it
// (gs/gs_util.cpp:141-171 compute_page_rect): the page of the first pixel (block base rounded
// down to its page), `width` x `height` pages, rows `stride` pages apart; page numbers wrap.
struct PageRect { uint32_t base, stride, width, height; };
inline PageRect page_rect(uint32_t base_block, uint32_t stride, uint32_t psm, uint32_t x0, uint32_t y0, uint32_t x1,
                          uint32_t y1)
{
	const Geometry g = psm_geometry(psm);
	if (psm == PSMT8 || psm == PSMT4)
		stride >>= 1;
	PageRect r;
	r.base = (base_block >> 5) + (x0 >> g.pw) + stride * (y0 >> g.ph);
	r.stride = stride;
	r.width = (x1 >> g.pw) - (x0 >> g.pw) + 1u;
	r.height = (y1 >> g.ph) - (y0 >> g.ph) + 1u;
	return r;
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

// Page spans of every mip level of a texture.
struct TexSpans
{
	PageSpan s[kMaxTexLevels];
	uint32_t n;
	bool overlaps(PageSpan o) const
	{
		for (uint32_t i = 0; i < n; i++)
			if (spans_overlap(s[i], o))
				return true;
		return false;
	}
};

inline TexSpans texture_span(const TexLayout &l)
{
	TexSpans t{};
	t.n = l.levels;
	for (uint32_t lv = 0; lv < l.levels; lv++)
	{
		uint32_t x0 = ~0u, x1 = 0, y0 = ~0u, y1 = 0;
		for (uint32_t x = 0; x < l.level_width(lv); x++)
		{
			x0 = std::min(x0, l.tex_x(x, lv));
			x1 = std::max(x1, l.tex_x(x, lv));
		}
		for (uint32_t y = 0; y < l.level_height(lv); y++)
		{
			y0 = std::min(y0, l.tex_y(y, lv));
			y1 = std::max(y1, l.tex_y(y, lv));
		}
		t.s[lv] = page_span(l.level_tbp[lv], l.level_tbw[lv], l.psm, x0, y0, x1, y1);
	}
	return t;
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
	uint32_t levels;         // mip levels, stored one after another from image_offset
	uint32_t texels;         // texels of all levels (0: nothing to decode)
	uint32_t level_buf[kMaxTexLevels - 1]; // TBP | TBW << 14 of levels 1..6
	// Super-sampled texture: samples per pixel S (0 otherwise). The image then holds S + 1
	// layers of width x height texels: layer 0 from VRAM, layer k from the plane of sample k - 1
	// where the reference plane marks the word super-sampled.
	uint32_t samples;
};
static_assert(sizeof(TexDecode) == 108, "TexDecode layout");
}
