// CPU oracle: CLUT loads, texture-cache decode and texture sampling (TEST INFRASTRUCTURE ONLY).
//
// Restated (not copied) from the reference:
//   gs/gs_interface.cpp:550-797          CLUT load conditions (CLD, cached CBP), CSA banks
//   gs/shaders/clut_write.comp:46-155    CSM1 / CSM2 palette reads into the 512-entry CLUT
//   gs/gs_renderer.cpp:1235-1265,        effective texture rectangle
//   gs/gs_util.cpp:31-79
//   gs/gs_interface.cpp:1754-1781        TexInfo sizes / region / bias
//   gs/shaders/upload.comp:108-228       texel decode (CLUT lookup, TEXA expansion)
//   gs/shaders/ubershader.comp:947-953,  texture_clamp and the non-array sampling path
//   1042-1258
//
// Filtering follows the Vulkan sampler model with REPEAT addressing on the decoded RGBA8
// image. The reference relies on the driver's sampler for NEAREST/LINEAR; the LINEAR weight
// arithmetic is not fixed by the reference (it lives in the Vulkan implementation), so this
// oracle fixes one: 8-bit fractional weights (floor(frac * 256)) and an integer lerp with
// rounding, first along U then along V. The HIP kernels implement the same definition.
#include "oracle_render.hpp"
#include <cstring>
#include <cmath>
#include <algorithm>

namespace orc
{
namespace
{
inline uint32_t field(uint64_t v, uint32_t off, uint32_t bits) { return uint32_t((v >> off) & ((1ull << bits) - 1ull)); }
bool is_palette(uint32_t psm) { return psm == T8 || psm == T4 || psm == T8H || psm == T4HL || psm == T4HH; }

uint32_t rgba16_to_32(uint32_t c16, uint32_t aem, uint32_t ta0, uint32_t ta1)
{
	uint32_t r = (c16 & 0x1fu) << 3, g = ((c16 >> 5) & 0x1fu) << 3, b = ((c16 >> 10) & 0x1fu) << 3;
	bool zero_alpha = aem && c16 == 0;
	uint32_t a = zero_alpha ? 0u : (((c16 >> 15) & 1u) ? ta1 : ta0);
	return r | g << 8 | b << 16 | std::min(a, 255u) << 24;
}

uint32_t rgb24_to_32(uint32_t c, uint32_t aem, uint32_t ta0)
{
	bool zero_alpha = aem && (c & 0xffffffu) == 0;
	return (c & 0xffffffu) | ((zero_alpha ? 0u : ta0) << 24);
}
}

// Palette loads happen when TEX0 is written (gs/gs_interface.cpp:550-797). A load reads VRAM,
// so any primitives already queued in the render pass are rendered first.
void GS::load_clut(uint32_t ctx)
{
	const uint64_t t0 = regs.tex0[ctx];
	const uint32_t psm = field(t0, 20, 6);
	if (!is_palette(psm))
		return;
	uint32_t cpsm = field(t0, 51, 4);
	if (cpsm == CT24)
		cpsm = CT32;
	uint32_t csa = field(t0, 56, 5);
	const uint32_t cld = field(t0, 61, 3);
	const uint32_t cbp = field(t0, 37, 14);
	bool load = false;
	switch (cld)
	{
	case 1: load = true; break;
	case 2: case 3: load = true; cached_cbp[cld & 1] = cbp; break;
	case 4: case 5: load = cached_cbp[cld & 1] != cbp; cached_cbp[cld & 1] = cbp; break;
	default: break;
	}
	if (!load)
		return;
	if (xfer.active)
		partial_flush_transfer();
	flush_render_pass();

	if (cpsm == CT32)
		csa &= 15u;
	const uint32_t csm = field(t0, 55, 1);
	const bool index4 = psm == T4 || psm == T4HL || psm == T4HH;
	const uint32_t cbw = csm ? field(regs.texclut, 0, 6) : 0;
	const uint32_t cou = csm ? field(regs.texclut, 6, 6) : 0;
	const uint32_t cov = csm ? field(regs.texclut, 12, 10) : 0;

	std::vector<uint16_t> tmp = clut;
	for (uint32_t lane = 0; lane < 256; lane++)
	{
		bool active;
		uint32_t write_coord = lane + csa * 16u;
		uint32_t read_coord = lane;
		if (index4)
			active = lane < 16;
		else if (csm == 0)
		{
			uint32_t csa_count = cpsm == CT32 ? 16u : 32u;
			active = lane < (csa_count - csa) * 16u;
			read_coord += csa * 16u;
		}
		else
			active = true;
		if (!active)
			continue;
		uint32_t x, y;
		if (csm)
		{
			x = uint32_t(float(read_coord) * 1.0f + 1.0f / 256.0f) + cou * 16u;
			y = cov;
		}
		else
		{
			x = (read_coord & 7u) | (((read_coord >> 4) & 1u) << 3);
			y = ((read_coord >> 3) & 1u) | (((read_coord >> 5) & 7u) << 1);
		}
		if (cpsm == CT32)
		{
			uint32_t a = swizzle(x, y, cbp, cbw, CT32);
			uint32_t d;
			memcpy(&d, vram.data() + a * 4u, 4);
			if (write_coord < 512)
				tmp[write_coord] = uint16_t(d & 0xffffu);
			if (write_coord + 256u < 512u)
				tmp[write_coord + 256u] = uint16_t(d >> 16);
		}
		else
		{
			uint32_t a = swizzle(x, y, cbp, cbw, cpsm);
			uint16_t d;
			memcpy(&d, vram.data() + a * 2u, 2);
			if (write_coord < 512)
				tmp[write_coord] = d;
		}
	}
	clut = tmp;
}

namespace
{
// Effective texture rectangle and texel addressing of TEX0 / CLAMP (gs/gs_util.cpp:31-79,
// gs/gs_renderer.cpp:1235-1265 and the upload addressing of gs/shaders/upload.comp:121-123).
// Level l of a mipmapped texture reads from its own TBP / TBW (MIPTBP1 for levels 1..3,
// MIPTBP2 for 4..6) with the origin, MSK and FIX shifted right by l
// (gs/gs_renderer.cpp:3502-3593).
struct TexLayout
{
	uint32_t psm, tw, th, levels;
	uint32_t tbp[7], tbw[7];
	uint32_t width, height, bx, by;
	uint32_t umsk, ufix, vmsk, vfix;
	uint32_t tex_x(uint32_t x, uint32_t l) const { return ((x & (umsk >> l)) + (bx >> l)) | (ufix >> l); }
	uint32_t tex_y(uint32_t y, uint32_t l) const { return ((y & (vmsk >> l)) + (by >> l)) | (vfix >> l); }
	uint32_t level_width(uint32_t l) const { return std::max(width >> l, 1u); }
	uint32_t level_height(uint32_t l) const { return std::max(height >> l, 1u); }
};

uint32_t effective_extent(uint32_t ext, uint32_t mode, uint32_t lo, uint32_t hi, uint32_t levels, uint32_t &base)
{
	base = 0;
	if (mode == 2)
	{
		lo = std::min(lo, hi);
		ext = std::max(hi, lo) - lo + 1u;
		if (levels > 1)
		{
			// Align the region to the coarsest level so every level has whole texels.
			uint32_t mask = (1u << (levels - 1u)) - 1u;
			ext += lo & mask;
			lo &= ~mask;
			ext = (ext + mask) & ~mask;
		}
		base = lo;
	}
	else if (mode == 3)
	{
		if (lo == 0)
		{
			ext = 1;
			base = hi;
		}
		else
		{
			uint32_t msb = 31u - uint32_t(__builtin_clz(lo));
			uint32_t lsb = hi ? uint32_t(__builtin_ctz(hi)) : 32u;
			if (lsb > msb)
			{
				ext = std::min(1u << (msb + 1u), ext);
				base = hi;
			}
		}
	}
	return ext;
}

// Levels of the texture image: MXL + 1 when TEX1.MMIN is a mipmapped filter (2..5), at most
// min(TW, TH) + 1 and 7 (gs/gs_renderer.cpp:1240-1243, gs/gs_interface.cpp:1539-1540).
uint32_t texture_levels(uint64_t tex0, uint64_t tex1)
{
	uint32_t mmin = field(tex1, 6, 3);
	if (mmin < 2 || mmin > 5)
		return 1;
	uint32_t tw = std::min(field(tex0, 26, 4), 10u), th = std::min(field(tex0, 30, 4), 10u);
	return std::min(std::min(field(tex1, 2, 3) + 1u, std::min(tw, th) + 1u), 7u);
}

TexLayout tex_layout(uint64_t tex0, uint64_t clamp, uint32_t levels, uint64_t mip1, uint64_t mip2)
{
	TexLayout l{};
	l.tbp[0] = field(tex0, 0, 14);
	l.tbw[0] = field(tex0, 14, 6);
	for (uint32_t i = 1; i < 7; i++)
	{
		uint64_t m = i <= 3 ? mip1 : mip2;
		uint32_t o = ((i - 1) % 3) * 20;
		l.tbp[i] = field(m, o, 14);
		l.tbw[i] = field(m, o + 14, 6);
	}
	l.psm = field(tex0, 20, 6);
	l.tw = std::min(field(tex0, 26, 4), 10u);
	l.th = std::min(field(tex0, 30, 4), 10u);
	l.levels = levels;
	const uint32_t wms = field(clamp, 0, 2), wmt = field(clamp, 2, 2);
	const uint32_t minu = field(clamp, 4, 10), maxu = field(clamp, 14, 10);
	const uint32_t minv = field(clamp, 24, 10), maxv = field(clamp, 34, 10);
	l.width = effective_extent(1u << l.tw, wms, minu, maxu, levels, l.bx);
	l.height = effective_extent(1u << l.th, wmt, minv, maxv, levels, l.by);
	l.umsk = wms == 3 ? minu : 0xffffffffu;
	l.ufix = wms == 3 ? maxu : 0u;
	l.vmsk = wmt == 3 ? minv : 0xffffffffu;
	l.vfix = wmt == 3 ? maxv : 0u;
	return l;
}
}

PageSpan page_span(uint32_t base_block, uint32_t stride, uint32_t psm, uint32_t x0, uint32_t y0, uint32_t x1,
                   uint32_t y1)
{
	// Page index is (y >> page_h) * stride_pages + (x >> page_w): non-decreasing in x and y,
	// so the rectangle's corners bound it. The Z formats' block XOR stays inside the
	// 32-block group, and an unaligned base can spill into the next page.
	uint32_t pw = 6, ph = 5;
	if (psm == CT16 || psm == CT16S || psm == Z16 || psm == Z16S) ph = 6;
	else if (psm == T8) { pw = 7; ph = 6; }
	else if (psm == T4) { pw = 7; ph = 7; }
	if (psm == T8 || psm == T4)
		stride >>= 1;
	uint32_t lo = (y0 >> ph) * stride + (x0 >> pw);
	uint32_t hi = (y1 >> ph) * stride + (x1 >> pw);
	return { (base_block >> 5) + lo, ((base_block + 31u) >> 5) + hi };
}

bool spans_overlap(PageSpan a, PageSpan b)
{
	const uint32_t pages = 512;
	if (a.last - a.first + 1u >= pages || b.last - b.first + 1u >= pages)
		return true;
	uint32_t a0 = a.first % pages, a1 = a0 + (a.last - a.first);
	uint32_t b0 = b.first % pages, b1 = b0 + (b.last - b.first);
	auto hit = [](uint32_t p0, uint32_t p1, uint32_t q0, uint32_t q1) { return p0 <= q1 && q0 <= p1; };
	return hit(a0, a1, b0, b1) || hit(a0 + pages, a1 + pages, b0, b1) || hit(a0, a1, b0 + pages, b1 + pages);
}

std::vector<PageSpan> GS::texture_span(uint32_t ctx) const
{
	const uint32_t levels = texture_levels(regs.tex0[ctx], regs.tex1[ctx]);
	const TexLayout l = tex_layout(regs.tex0[ctx], regs.clamp[ctx], levels, regs.miptbp1[ctx], regs.miptbp2[ctx]);
	std::vector<PageSpan> spans;
	for (uint32_t lv = 0; lv < levels; lv++)
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
		spans.push_back(page_span(l.tbp[lv], l.tbw[lv], l.psm, x0, y0, x1, y1));
	}
	return spans;
}

// Does the texture of context `ctx` (any of its levels) read pages the pending render pass
// writes?
bool GS::texture_hazard(uint32_t ctx) const
{
	const uint32_t x0 = uint32_t(rp_bb[0]), y0 = uint32_t(rp_bb[1]), x1 = uint32_t(rp_bb[2]), y1 = uint32_t(rp_bb[3]);
	const PageSpan fb = page_span(field(rp_frame, 0, 9) * 32u, field(rp_frame, 16, 6), field(rp_frame, 24, 6), x0, y0,
	                              x1, y1);
	const PageSpan z = page_span(field(rp_zbuf, 0, 9) * 32u, field(rp_frame, 16, 6), 0x30u | field(rp_zbuf, 24, 4), x0,
	                             y0, x1, y1);
	for (const PageSpan &t : texture_span(ctx))
		if (spans_overlap(t, fb) || (rp_zwrite && spans_overlap(t, z)))
			return true;
	return false;
}

// Finds or creates the render pass's texture slot for the current TEX0 / CLAMP / TEXA state
// and decodes the texture image from VRAM (gs/shaders/upload.comp:108-228).
// The page tracker's page rectangle walk (gs/gs_util.cpp:141-171 compute_page_rect): the page
// of the first pixel, then whole pages to the last pixel's, row by row.
void GS::mark_pages(uint32_t base_block, uint32_t stride, uint32_t psm, uint32_t x0, uint32_t y0, uint32_t x1,
                    uint32_t y1, bool fb)
{
	uint32_t pw = 6, ph = 5; // 32-bit formats: 64 x 32 pages
	if (psm == CT16 || psm == CT16S || psm == Z16 || psm == Z16S) { pw = 6; ph = 6; }
	else if (psm == T8) { pw = 7; ph = 6; }
	else if (psm == T4) { pw = 7; ph = 7; }
	if (psm == T8 || psm == T4)
		stride >>= 1;
	const uint32_t base = base_block / 32u + (x0 >> pw) + stride * (y0 >> ph);
	for (uint32_t y = 0; y <= (y1 >> ph) - (y0 >> ph); y++)
		for (uint32_t x = 0; x <= (x1 >> pw) - (x0 >> pw); x++)
			may_super_sample[(base + y * stride + x) % 512u] = fb ? 1 : 0;
}

uint32_t GS::texture_slot(uint32_t ctx)
{
	const uint64_t key0 = regs.tex0[ctx] & ((1ull << 61) - 1ull); // CLD does not affect the texture
	const uint64_t keyc = regs.clamp[ctx], keya = regs.texa;
	const uint32_t levels = texture_levels(key0, regs.tex1[ctx]);
	// Only the used levels' MIPTBP fields select the image.
	const uint32_t n1 = std::min(levels - 1u, 3u), n2 = levels > 4 ? levels - 4u : 0u;
	const uint64_t mip1 = regs.miptbp1[ctx] & ((1ull << (20 * n1)) - 1ull);
	const uint64_t mip2 = regs.miptbp2[ctx] & ((1ull << (20 * n2)) - 1ull);
	// Super-sampled when super-sampled textures are on, the primitive is a sprite and the texture
	// has one level of 8 bits per texel or more whose base page was last written by a render pass
	// (gs/gs_interface.cpp:1593-1603; the base page of the level-0 page rectangle).
	uint32_t samples = 0;
	const uint32_t tpsm = field(key0, 20, 6);
	const bool bpp8 = !(tpsm == T4 || tpsm == T4HL || tpsm == T4HH);
	if (super_sampled_textures && sample_log2 != 0 && levels == 1 && bpp8 && field(regs.prim, 0, 3) == 6)
	{
		const TexLayout l0 = tex_layout(key0, keyc, 1, 0, 0);
		uint32_t x0 = ~0u, y0 = ~0u;
		for (uint32_t x = 0; x < l0.level_width(0); x++)
			x0 = std::min(x0, l0.tex_x(x, 0));
		for (uint32_t y = 0; y < l0.level_height(0); y++)
			y0 = std::min(y0, l0.tex_y(y, 0));
		uint32_t pw = 6, ph = 5, stride = l0.tbw[0];
		if (tpsm == CT16 || tpsm == CT16S || tpsm == Z16 || tpsm == Z16S) { pw = 6; ph = 6; }
		else if (tpsm == T8) { pw = 7; ph = 6; stride >>= 1; }
		const uint32_t base = l0.tbp[0] / 32u + (x0 >> pw) + stride * (y0 >> ph);
		if (may_super_sample[base % 512u])
			samples = 1u << sample_log2;
	}
	for (uint32_t i = 0; i < rp_textures.size(); i++)
	{
		const TextureSlot &o = rp_textures[i];
		if (o.tex0 == key0 && o.clamp == keyc && o.texa == keya && o.levels == levels && o.mip1 == mip1 &&
		    o.mip2 == mip2 && o.samples == samples)
			return i;
	}

	TextureSlot t{};
	t.samples = samples;
	t.tex0 = key0;
	t.clamp = keyc;
	t.texa = keya;
	t.mip1 = mip1;
	t.mip2 = mip2;
	t.levels = levels;
	const TexLayout l = tex_layout(key0, keyc, levels, mip1, mip2);
	const uint32_t psm = l.psm;
	const uint32_t cpsm0 = field(key0, 51, 4);
	const bool clut16 = cpsm0 == CT16 || cpsm0 == CT16S;
	// A 32-bit CLUT only addresses the lower 16 banks (gs/gs_interface.cpp:1306-1311).
	const uint32_t csa = clut16 ? field(key0, 56, 5) : field(key0, 56, 5) & 15u;
	const uint32_t wms = field(keyc, 0, 2), wmt = field(keyc, 2, 2);
	const uint32_t aem = field(keya, 15, 1), ta0 = field(keya, 0, 8), ta1 = field(keya, 32, 8);

	t.width = l.width;
	t.height = l.height;
	t.region[0] = t.region[1] = t.region[2] = t.region[3] = 0.0f;
	if (wms == 1) { t.region[0] = 0.0f; t.region[2] = float(t.width) - 1.0f; }
	else if (wms == 2) { t.region[0] = float(field(keyc, 4, 10)); t.region[2] = float(field(keyc, 14, 10)); }
	if (wmt == 1) { t.region[1] = 0.0f; t.region[3] = float(t.height) - 1.0f; }
	else if (wmt == 2) { t.region[1] = float(field(keyc, 24, 10)); t.region[3] = float(field(keyc, 34, 10)); }
	t.size[0] = float(1u << l.tw);
	t.size[1] = float(1u << l.th);
	t.size[2] = 1.0f / float(t.width);
	t.size[3] = 1.0f / float(t.height);
	t.bias[0] = -float(l.bx) * t.size[2];
	t.bias[1] = -float(l.by) * t.size[3];

	auto clut_color = [&](uint32_t index) -> uint32_t {
		uint32_t bi = csa * 16u + index;
		if (psm == T8 || psm == T8H)
		{
			uint32_t bank = std::min(bi >> 4, clut16 ? 31u : 15u);
			bi = bank * 16u + (bi & 15u);
		}
		if (clut16)
			return rgba16_to_32(clut[bi & 511u], aem, ta0, ta1);
		return uint32_t(clut[bi & 511u]) | (uint32_t(clut[(bi + 256u) & 511u]) << 16);
	};

	size_t total = 0;
	for (uint32_t lv = 0; lv < levels; lv++)
	{
		t.level_offset[lv] = uint32_t(total);
		t.level_width[lv] = l.level_width(lv);
		t.level_height[lv] = l.level_height(lv);
		total += size_t(t.level_width[lv]) * t.level_height[lv];
	}
	const uint32_t layers = samples ? samples + 1u : 1u;
	t.rgba.resize(total * layers);
	const uint8_t *v = vram.data();
	for (uint32_t layer = 0; layer < layers; layer++)
	for (uint32_t lv = 0; lv < levels; lv++)
	for (uint32_t y = 0; y < t.level_height[lv]; y++)
		for (uint32_t x = 0; x < t.level_width[lv]; x++)
		{
			uint32_t a = swizzle(l.tex_x(x, lv), l.tex_y(y, lv), l.tbp[lv], l.tbw[lv], psm);
			// Layer k > 0 reads sample k - 1 where the reference plane marks the value
			// super-sampled (upload.comp:133-228 with SUPER_SAMPLED).
			const uint8_t *ref = v + VRAM_SIZE, *hi = v + size_t(layer + 1u) * VRAM_SIZE;
			uint32_t c = 0;
			switch (psm)
			{
			case CT24: case Z24:
			{
				uint32_t w, m = 0, h;
				memcpy(&w, v + a * 4u, 4);
				if (layer)
				{
					memcpy(&m, ref + a * 4u, 4);
					memcpy(&h, hi + a * 4u, 4);
					if (m != 0)
						w = (h & m) | (w & ~m);
				}
				c = rgb24_to_32(w, aem, ta0);
				break;
			}
			case CT16: case CT16S: case Z16: case Z16S:
			{
				uint16_t h, m = 0;
				memcpy(&h, v + a * 2u, 2);
				if (layer)
				{
					memcpy(&m, ref + a * 2u, 2);
					if (m == 0xffffu)
						memcpy(&h, hi + a * 2u, 2);
				}
				c = rgba16_to_32(h, aem, ta0, ta1);
				break;
			}
			case T8: c = clut_color((layer && ref[a] == 0xffu) ? hi[a] : v[a]); break;
			case T8H: c = clut_color((layer && ref[a * 4u + 3u] == 0xffu) ? hi[a * 4u + 3u] : v[a * 4u + 3u]); break;
			case T4: c = clut_color((v[a >> 1] >> ((a & 1u) * 4u)) & 0xfu); break;
			case T4HL: case T4HH:
			{
				uint32_t w;
				memcpy(&w, v + a * 4u, 4);
				c = clut_color((w >> (psm == T4HH ? 28 : 24)) & 0xfu);
				break;
			}
			default:
			{
				uint32_t m = 0;
				memcpy(&c, v + a * 4u, 4);
				if (layer)
				{
					memcpy(&m, ref + a * 4u, 4);
					if (m == ~0u)
						memcpy(&c, hi + a * 4u, 4);
				}
				break;
			}
			}
			t.rgba[size_t(layer) * total + t.level_offset[lv] + size_t(y) * t.level_width[lv] + x] = c;
		}

	rp_textures.push_back(std::move(t));
	return uint32_t(rp_textures.size() - 1);
}

namespace
{
// Texel (x, y) of an image level with REPEAT addressing.
inline uint32_t texel(const TextureSlot &t, uint32_t lv, int32_t x, int32_t y)
{
	const int32_t w = int32_t(t.level_width[lv]), h = int32_t(t.level_height[lv]);
	uint32_t xx = uint32_t(((x % w) + w) % w);
	uint32_t yy = uint32_t(((y % h) + h) % h);
	return t.rgba[t.level_offset[lv] + size_t(yy) * uint32_t(w) + xx];
}

inline uint32_t lerp8(uint32_t a, uint32_t b, uint32_t w)
{
	return uint32_t(int32_t(a) + ((int32_t(b) - int32_t(a)) * int32_t(w) + 128) / 256 -
	                (((int32_t(b) - int32_t(a)) * int32_t(w) + 128) % 256 < 0 ? 1 : 0));
}

// A Vulkan textureLod fetch with REPEAT addressing at integer level `lod` of the image, which
// reads level min(lod, levels - 1), for normalized coordinates (un, vn).
uint32_t fetch(const TextureSlot &t, uint32_t lod, float un, float vn, bool linear)
{
	const uint32_t lv = std::min(lod, t.levels - 1u);
	const float fw = float(t.level_width[lv]), fh = float(t.level_height[lv]);
	if (!linear)
	{
		float xs = std::floor(un * fw), ys = std::floor(vn * fh);
		return texel(t, lv, int32_t(xs), int32_t(ys));
	}
	float xs = un * fw - 0.5f, ys = vn * fh - 0.5f;
	float x0 = std::floor(xs), y0 = std::floor(ys);
	uint32_t wx = uint32_t((xs - x0) * 256.0f), wy = uint32_t((ys - y0) * 256.0f);
	uint32_t c00 = texel(t, lv, int32_t(x0), int32_t(y0)), c10 = texel(t, lv, int32_t(x0) + 1, int32_t(y0));
	uint32_t c01 = texel(t, lv, int32_t(x0), int32_t(y0) + 1), c11 = texel(t, lv, int32_t(x0) + 1, int32_t(y0) + 1);
	uint32_t c = 0;
	for (uint32_t ch = 0; ch < 4; ch++)
	{
		uint32_t sh = ch * 8u;
		uint32_t top = lerp8((c00 >> sh) & 0xffu, (c10 >> sh) & 0xffu, wx);
		uint32_t bot = lerp8((c01 >> sh) & 0xffu, (c11 >> sh) & 0xffu, wx);
		c |= lerp8(top, bot, wy) << sh;
	}
	return c;
}

// texture_clamp (ubershader.comp:947-953): the region scaled to the level.
inline float texture_clamp(float coord, float lo, float hi, float lod)
{
	float a = std::floor(std::ldexp(lo, -int(lod))) + 0.5f;
	float b = std::floor(std::ldexp(hi, -int(lod))) + 0.5f;
	return std::min(std::max(coord, a), b);
}
}

// The Q-derived level of detail ceil(ldexp(-log2|q|, L + 4) + K) (ubershader.comp:1076-1087)
// evaluated exactly rather than with a rounded log2 (the precision of log2 is the
// implementation's): with |q| = m * 2^e, m in [1, 2), it is K - 2^(L+4) e - floor(2^(L+4) log2 m).
// 2^(L+4) log2 m is irrational for m != 1, so the long-double floor is exact unless it lies
// within 2^-50 or so of an integer (the tests check every threshold). A zero, subnormal,
// infinite or NaN q gives a formula of 0, as the reference's non-finite flush does (a subnormal
// q taken as flushed to zero). Returns whether the formula is above 0.
bool lod_from_q(float q, int32_t k, uint32_t l, int32_t &n)
{
	const float a = std::fabs(q);
	if (!(a >= 0x1p-126f) || std::isinf(a))
	{
		n = 0;
		return false;
	}
	int e;
	const float m = std::frexp(a, &e) * 2.0f; // in [1, 2)
	e -= 1;
	const long double s = std::ldexp((long double)std::log2((long double)m), int(l) + 4);
	n = k - (16 << l) * e - int32_t(std::floor(s));
	return n > 0;
}

// Phase filter table of super-sampled textures for a 2^rate_x x 2^rate_y sampling grid
// (gs/gs_renderer.cpp:94-313 init_phase_lut): 64 x 64 LINEAR entries of two words (tap sample
// ids in bits 0..15, 2-bit texel offsets x, y in bits 16..31; unorm8 weights), then 32 x 32
// NEAREST entries (the nearest sample id).
void phase_lut(uint32_t rate_x, uint32_t rate_y, std::vector<uint32_t> &out)
{
	out.assign(2 * (64 * 64 + 32 * 32), 0u);
	if (rate_y == 0 || rate_x > rate_y || rate_y > 3)
		return;
	const int ns = 1 << (rate_x + rate_y);
	// compute_sample_points (gs_renderer.cpp:106-150) on a 64-step lattice per texel.
	auto points = [&](bool nearest, std::vector<std::pair<int, int>> &pts) {
		pts.clear();
		const int step = 64 >> rate_y;
		for (int i = 0; i < ns; i++)
		{
			int x, y;
			if (rate_y == rate_x + 2)
			{
				const int sparse[4] = { 0, 2, 3, 1 };
				x = 4 * (i / 8) + sparse[i % 4];
				y = i % 8;
			}
			else
			{
				y = (i & 1) + 2 * ((i >> 2) & 1);
				x = ((i >> 1) & 1) + 2 * ((i >> 3) & 1);
				if (rate_y == rate_x + 1)
					x = 2 * x + (i & 1);
			}
			x *= step;
			y *= step;
			if (nearest)
			{
				x += (step - 1) / 2;
				y += (step - 1) / 2;
			}
			pts.emplace_back(x, y);
		}
	};
	struct R { int id, tx, ty, dx, dy, m; };
	// compute_phase_lut_samples (gs_renderer.cpp:152-187): every sample point of the texels
	// within `reach`, sorted by Chebyshev distance then vertical distance (std::sort, as there).
	auto gather = [&](int x, int y, const std::vector<std::pair<int, int>> &pts, int reach) {
		std::vector<R> r;
		for (int i = 0; i < ns; i++)
			for (int ty = -reach; ty <= reach; ty++)
				for (int tx = -reach; tx <= reach; tx++)
				{
					R e;
					e.id = i;
					e.tx = tx;
					e.ty = ty;
					e.dx = std::abs(pts[size_t(i)].first + 64 * tx - x);
					e.dy = std::abs(pts[size_t(i)].second + 64 * ty - y);
					e.m = std::max(e.dx, e.dy);
					r.push_back(e);
				}
		std::sort(r.begin(), r.end(), [](const R &a, const R &b) { return a.m == b.m ? a.dy < b.dy : a.m < b.m; });
		return r;
	};
	std::vector<std::pair<int, int>> pts, near;
	points(false, pts);
	points(true, near);
	int falloff = 1 << 30;
	for (int y = 0; y < 64; y++)
		for (int x = 0; x < 64; x++)
			falloff = std::min(falloff, gather(x, y, pts, 1)[4].m);
	for (int y = 0; y < 64; y++)
		for (int x = 0; x < 64; x++)
		{
			std::vector<R> r = gather(x, y, pts, 1);
			if (r[0].tx != 0 || r[0].ty != 0)
				for (size_t i = 1; i < r.size(); i++)
					if (r[i].tx == 0 && r[i].ty == 0)
					{
						if (i < 4)
							std::swap(r[0], r[i]);
						else
						{
							std::swap(r[3], r[i]);
							std::swap(r[0], r[3]);
						}
						break;
					}
			uint32_t w0 = 0;
			float wt[4], sum = 0.0f;
			for (int i = 0; i < 4; i++)
			{
				w0 |= uint32_t(r[size_t(i)].id) << (4 * i) | (uint32_t(r[size_t(i)].tx) & 3u) << (16 + 4 * i) |
				      (uint32_t(r[size_t(i)].ty) & 3u) << (18 + 4 * i);
				const float ex = float(r[size_t(i)].dx) / float(falloff), ey = float(r[size_t(i)].dy) / float(falloff);
				wt[i] = std::max(0.0f, 1.0f - ex) * std::max(0.0f, 1.0f - ey);
				sum += wt[i];
			}
			if (sum < 0.001f)
			{
				wt[0] = wt[1] = wt[2] = wt[3] = 1.0f;
				sum = 4.0f;
			}
			uint32_t w1 = 0, used = 0;
			for (int i = 0; i < 3; i++)
			{
				uint32_t w = uint32_t(wt[i] * 255.0f / sum + 0.5f);
				w = std::min(w, 255u - used);
				used += w;
				w1 |= w << (8 * i);
			}
			w1 |= (255u - used) << 24;
			out[size_t(2 * (64 * y + x))] = w0;
			out[size_t(2 * (64 * y + x) + 1)] = w1;
		}
	for (int y = 0; y < 32; y++)
		for (int x = 0; x < 32; x++)
			out[size_t(2 * (4096 + 32 * y + x))] = uint32_t(gather(2 * x, 2 * y, near, 0)[0].id);
}

namespace
{
// Nearest texel of image `layer` of a super-sampled texture (REPEAT addressing).
uint32_t layer_texel(const TextureSlot &t, uint32_t layer, float un, float vn)
{
	const float fw = float(t.width), fh = float(t.height);
	const int32_t w = int32_t(t.width), h = int32_t(t.height);
	const int32_t x = int32_t(std::floor(un * fw)), y = int32_t(std::floor(vn * fh));
	const uint32_t xx = uint32_t(((x % w) + w) % w), yy = uint32_t(((y % h) + h) % h);
	return t.rgba[size_t(layer) * t.width * t.height + size_t(yy) * t.width + xx];
}
}

// Super-sampled texture read (ubershader.comp:1155-1211) at the clamped texel coordinate
// (u1, v1) for sample `smp`; `shift` is 0.5 less the mean sample offset of the shading grid.
uint32_t sample_texture_array(const TextureSlot &t, const uint32_t *lut, uint32_t tex, float u1, float v1, uint32_t smp,
                              float shift)
{
	const bool linear = (tex & (1u << 16)) != 0;
	const float u = u1 + 1.0f / 1024.0f, v = v1 + 1.0f / 1024.0f;
	const bool per_sample = (tex & (1u << 24)) != 0, mapping = (tex & (1u << 25)) != 0;
	const bool resolved = (tex & (1u << 26)) != 0;
	if (!per_sample || mapping || resolved)
	{
		// A NEAREST or LINEAR (8-bit weights, round-half-up lerp) fetch of one image, with
		// REPEAT addressing, at the unsnapped coordinate.
		const uint32_t layer = mapping ? smp + 1u : 0u;
		const float un = u * t.size[2] + t.bias[0], vn = v * t.size[3] + t.bias[1];
		const float fw = float(t.width), fh = float(t.height);
		auto at = [&](int32_t x, int32_t y) {
			const int32_t w = int32_t(t.width), h = int32_t(t.height);
			const uint32_t xx = uint32_t(((x % w) + w) % w), yy = uint32_t(((y % h) + h) % h);
			return t.rgba[size_t(layer) * t.width * t.height + size_t(yy) * t.width + xx];
		};
		if (!linear)
			return at(int32_t(std::floor(un * fw)), int32_t(std::floor(vn * fh)));
		const float xs = un * fw - 0.5f, ys = vn * fh - 0.5f;
		const float x0 = std::floor(xs), y0 = std::floor(ys);
		const uint32_t wx = uint32_t((xs - x0) * 256.0f), wy = uint32_t((ys - y0) * 256.0f);
		const int32_t ix = int32_t(x0), iy = int32_t(y0);
		const uint32_t c00 = at(ix, iy), c10 = at(ix + 1, iy), c01 = at(ix, iy + 1), c11 = at(ix + 1, iy + 1);
		uint32_t c = 0;
		for (uint32_t ch = 0; ch < 4; ch++)
		{
			const uint32_t sh = 8 * ch;
			const uint32_t top = lerp8((c00 >> sh) & 0xffu, (c10 >> sh) & 0xffu, wx);
			const uint32_t bot = lerp8((c01 >> sh) & 0xffu, (c11 >> sh) & 0xffu, wx);
			c |= lerp8(top, bot, wy) << sh;
		}
		return c;
	}
	const float su = u - shift, sv = v - shift;
	const float fu = std::floor(su), fv = std::floor(sv);
	const uint32_t n = linear ? 64u : 32u;
	const uint32_t ix = uint32_t((su - fu) * float(n)), iy = uint32_t((sv - fv) * float(n));
	const uint32_t *e = lut + 2u * (linear ? iy * 64u + ix : 4096u + iy * 32u + ix);
	const float cu = (fu + 0.5f) * t.size[2] + t.bias[0];
	const float cv = (fv + 0.5f) * t.size[3] + t.bias[1];
	const uint32_t c0 = layer_texel(t, (e[0] & 15u) + 1u, cu, cv);
	// The texel is on the clamp region's edge (a reversed region, min > max, clamps to [max, max],
	// GLSL's clamp being undefined there, so its edge is max).
	const bool clamp_any = (tex & ((1u << 19) | (1u << 20))) != 0;
	const float lo_u = std::min(t.region[0], t.region[2]), lo_v = std::min(t.region[1], t.region[3]);
	const bool edge = clamp_any && (fu == lo_u || fv == lo_v || fu == t.region[2] || fv == t.region[3]);
	if (!linear || edge)
		return c0;
	float acc[4];
	const float w0 = float(e[1] & 0xffu) / 255.0f;
	for (uint32_t ch = 0; ch < 4; ch++)
		acc[ch] = (float((c0 >> (8 * ch)) & 0xffu) / 255.0f) * w0;
	for (uint32_t i = 1; i < 4; i++)
	{
		const uint32_t wb = (e[1] >> (8 * i)) & 0xffu;
		if (!wb)
			continue;
		const float w = float(wb) / 255.0f;
		int32_t ox = int32_t((e[0] >> (16 + 4 * i)) & 3u), oy = int32_t((e[0] >> (18 + 4 * i)) & 3u);
		ox = ox >= 2 ? ox - 4 : ox;
		oy = oy >= 2 ? oy - 4 : oy;
		const uint32_t ci = layer_texel(t, ((e[0] >> (4 * i)) & 15u) + 1u, cu + float(ox) * t.size[2],
		                                cv + float(oy) * t.size[3]);
		for (uint32_t ch = 0; ch < 4; ch++)
			acc[ch] = acc[ch] + w * (float((ci >> (8 * ch)) & 0xffu) / 255.0f);
	}
	uint32_t c = 0;
	for (uint32_t ch = 0; ch < 4; ch++)
		c |= uint32_t(std::nearbyint(255.0f * acc[ch])) << (8 * ch);
	return c;
}

// ubershader.comp:1042-1258.
void sample_texture(const TextureSlot &t, uint32_t tex, uint32_t tex2, float u, float v, float q, bool perspective,
                    const uint32_t *lut, uint32_t smp, float shift, float out[4])
{
	const bool clamp_s = (tex & (1u << 19)) != 0, clamp_t = (tex & (1u << 20)) != 0;
	const bool mag_linear = (tex & (1u << 16)) != 0, min_linear = (tex & (1u << 17)) != 0;
	const bool trilinear = (tex & (1u << 18)) != 0;
	const int32_t max_miplevel = int32_t((tex >> 21) & 7u);

	// LOD (ubershader.comp:1066-1131).
	float lod = 0.0f;
	bool bilinear = mag_linear;
	if (max_miplevel > 0)
	{
		const int32_t k = int32_t(tex2 << 17) >> 20; // bits 3..14, signed
		if (!(tex2 & 1u))
		{
			int32_t n;
			const bool positive = lod_from_q(q, k, (tex2 >> 1) & 3u, n);
			lod = float(n) * 0.0625f;
			bilinear = positive ? min_linear : mag_linear;
		}
		else
		{
			lod = float(k >> 4);
			bilinear = lod > 0.0f ? min_linear : mag_linear;
		}
	}
	lod = std::min(std::max(lod, 0.0f), float(max_miplevel));
	float lod_1 = std::floor(lod), lod_2 = std::ceil(lod);
	float fract_lod = lod - lod_1;
	if (!trilinear)
	{
		fract_lod = 0.0f;
		lod_1 = float(int(lod + 0.5f));
		lod_2 = lod_1;
	}

	const float bs[2] = { perspective ? t.size[0] : 1.0f, perspective ? t.size[1] : 1.0f };
	float u1 = u * std::ldexp(bs[0], -int(lod_1));
	float v1 = v * std::ldexp(bs[1], -int(lod_1));
	if (clamp_s)
		u1 = texture_clamp(u1, t.region[0], t.region[2], lod_1);
	if (clamp_t)
		v1 = texture_clamp(v1, t.region[1], t.region[3], lod_1);
	if (t.samples)
	{
		const uint32_t c = sample_texture_array(t, lut, tex, u1, v1, smp, shift);
		for (uint32_t ch = 0; ch < 4; ch++)
			out[ch] = float((c >> (8 * ch)) & 0xffu);
		return;
	}
	const float inv1[2] = { std::ldexp(t.size[2], int(lod_1)), std::ldexp(t.size[3], int(lod_1)) };
	float un = std::floor(u1 * 16.0f + 1.0f / 64.0f);
	float vn = std::floor(v1 * 16.0f + 1.0f / 64.0f);
	un = (un + 16.0f / 1024.0f) * inv1[0] * 0.0625f + t.bias[0];
	vn = (vn + 16.0f / 1024.0f) * inv1[1] * 0.0625f + t.bias[1];
	uint32_t c = fetch(t, uint32_t(lod_1), un, vn, bilinear);

	if (fract_lod != 0.0f)
	{
		float u2 = u * std::ldexp(bs[0], -int(lod_2));
		float v2 = v * std::ldexp(bs[1], -int(lod_2));
		if (clamp_s)
			u2 = texture_clamp(u2, t.region[0], t.region[2], lod_2);
		if (clamp_t)
			v2 = texture_clamp(v2, t.region[1], t.region[3], lod_2);
		const float inv2[2] = { std::ldexp(t.size[2], int(lod_2)), std::ldexp(t.size[3], int(lod_2)) };
		// The second level is normalized without the rectangle origin (bias), as in the reference.
		float un2 = (std::floor(u2 * 16.0f + 1.0f / 64.0f) + 16.0f / 1024.0f) * inv2[0] * 0.0625f;
		float vn2 = (std::floor(v2 * 16.0f + 1.0f / 64.0f) + 16.0f / 1024.0f) * inv2[1] * 0.0625f;
		const uint32_t c2 = fetch(t, uint32_t(lod_2), un2, vn2, bilinear);
		const float weight = std::floor(fract_lod * 16.0f);
		uint32_t r = 0;
		for (uint32_t ch = 0; ch < 4; ch++)
		{
			const float a = float((c >> (8 * ch)) & 0xffu), b = float((c2 >> (8 * ch)) & 0xffu);
			r |= uint32_t(std::nearbyint(((16.0f - weight) * a + b * weight) * (1.0f / 16.0f))) << (8 * ch);
		}
		c = r;
	}
	for (uint32_t ch = 0; ch < 4; ch++)
		out[ch] = float((c >> (8 * ch)) & 0xffu);
}
}
