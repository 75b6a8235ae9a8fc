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
// Effective texture rectangle and texel addressing of TEX0 / CLAMP (gs/gs_util.cpp:31-79 and
// the upload addressing of gs/shaders/upload.comp:121-123).
struct TexLayout
{
	uint32_t tbp, tbw, psm, tw, th;
	uint32_t width, height, bx, by;
	uint32_t umsk, ufix, vmsk, vfix;
	uint32_t tex_x(uint32_t x) const { return ((x & umsk) + bx) | ufix; }
	uint32_t tex_y(uint32_t y) const { return ((y & vmsk) + by) | vfix; }
};

uint32_t effective_extent(uint32_t ext, uint32_t mode, uint32_t lo, uint32_t hi, uint32_t &base)
{
	base = 0;
	if (mode == 2)
	{
		lo = std::min(lo, hi);
		ext = std::max(hi, lo) - lo + 1u;
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

TexLayout tex_layout(uint64_t tex0, uint64_t clamp)
{
	TexLayout l{};
	l.tbp = field(tex0, 0, 14);
	l.tbw = field(tex0, 14, 6);
	l.psm = field(tex0, 20, 6);
	l.tw = std::min(field(tex0, 26, 4), 10u);
	l.th = std::min(field(tex0, 30, 4), 10u);
	const uint32_t wms = field(clamp, 0, 2), wmt = field(clamp, 2, 2);
	const uint32_t minu = field(clamp, 4, 10), maxu = field(clamp, 14, 10);
	const uint32_t minv = field(clamp, 24, 10), maxv = field(clamp, 34, 10);
	l.width = effective_extent(1u << l.tw, wms, minu, maxu, l.bx);
	l.height = effective_extent(1u << l.th, wmt, minv, maxv, l.by);
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

PageSpan GS::texture_span(uint32_t ctx) const
{
	const TexLayout l = tex_layout(regs.tex0[ctx], regs.clamp[ctx]);
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

// Does the texture of context `ctx` read pages the pending render pass writes?
bool GS::texture_hazard(uint32_t ctx) const
{
	const PageSpan t = texture_span(ctx);
	const uint32_t x0 = uint32_t(rp_bb[0]), y0 = uint32_t(rp_bb[1]), x1 = uint32_t(rp_bb[2]), y1 = uint32_t(rp_bb[3]);
	const PageSpan fb = page_span(field(rp_frame, 0, 9) * 32u, field(rp_frame, 16, 6), field(rp_frame, 24, 6), x0, y0,
	                              x1, y1);
	if (spans_overlap(t, fb))
		return true;
	if (rp_zwrite)
	{
		const PageSpan z = page_span(field(rp_zbuf, 0, 9) * 32u, field(rp_frame, 16, 6),
		                             0x30u | field(rp_zbuf, 24, 4), x0, y0, x1, y1);
		if (spans_overlap(t, z))
			return true;
	}
	return false;
}

// Finds or creates the render pass's texture slot for the current TEX0 / CLAMP / TEXA state
// and decodes the texture image from VRAM (gs/shaders/upload.comp:108-228).
uint32_t GS::texture_slot(uint32_t ctx)
{
	const uint64_t key0 = regs.tex0[ctx] & ((1ull << 61) - 1ull); // CLD does not affect the texture
	const uint64_t keyc = regs.clamp[ctx], keya = regs.texa;
	for (uint32_t i = 0; i < rp_textures.size(); i++)
		if (rp_textures[i].tex0 == key0 && rp_textures[i].clamp == keyc && rp_textures[i].texa == keya)
			return i;

	TextureSlot t{};
	t.tex0 = key0;
	t.clamp = keyc;
	t.texa = keya;
	const TexLayout l = tex_layout(key0, keyc);
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

	t.rgba.resize(size_t(t.width) * t.height);
	const uint8_t *v = vram.data();
	for (uint32_t y = 0; y < t.height; y++)
		for (uint32_t x = 0; x < t.width; x++)
		{
			uint32_t a = swizzle(l.tex_x(x), l.tex_y(y), l.tbp, l.tbw, psm);
			uint32_t c = 0;
			switch (psm)
			{
			case CT24: case Z24:
			{
				uint32_t w;
				memcpy(&w, v + a * 4u, 4);
				c = rgb24_to_32(w, aem, ta0);
				break;
			}
			case CT16: case CT16S: case Z16: case Z16S:
			{
				uint16_t h;
				memcpy(&h, v + a * 2u, 2);
				c = rgba16_to_32(h, aem, ta0, ta1);
				break;
			}
			case T8: c = clut_color(v[a]); break;
			case T8H: c = clut_color(v[a * 4u + 3u]); break;
			case T4: c = clut_color((v[a >> 1] >> ((a & 1u) * 4u)) & 0xfu); break;
			case T4HL: case T4HH:
			{
				uint32_t w;
				memcpy(&w, v + a * 4u, 4);
				c = clut_color((w >> (psm == T4HH ? 28 : 24)) & 0xfu);
				break;
			}
			default:
				memcpy(&c, v + a * 4u, 4);
				break;
			}
			t.rgba[size_t(y) * t.width + x] = c;
		}

	rp_textures.push_back(std::move(t));
	return uint32_t(rp_textures.size() - 1);
}

namespace
{
inline uint32_t texel(const TextureSlot &t, int32_t x, int32_t y)
{
	uint32_t xx = uint32_t(((x % int32_t(t.width)) + int32_t(t.width)) % int32_t(t.width));
	uint32_t yy = uint32_t(((y % int32_t(t.height)) + int32_t(t.height)) % int32_t(t.height));
	return t.rgba[size_t(yy) * t.width + xx];
}

inline uint32_t lerp8(uint32_t a, uint32_t b, uint32_t w)
{
	return uint32_t(int32_t(a) + ((int32_t(b) - int32_t(a)) * int32_t(w) + 128) / 256 -
	                (((int32_t(b) - int32_t(a)) * int32_t(w) + 128) % 256 < 0 ? 1 : 0));
}
}

// ubershader.comp:1042-1258 (non-array path, LOD 0).
void sample_texture(const TextureSlot &t, uint32_t tex, float u, float v, float, bool perspective, float out[4])
{
	const bool clamp_s = (tex & (1u << 19)) != 0, clamp_t = (tex & (1u << 20)) != 0;
	const bool linear = (tex & (1u << 16)) != 0;
	float u1 = perspective ? u * t.size[0] : u;
	float v1 = perspective ? v * t.size[1] : v;
	if (clamp_s)
		u1 = std::min(std::max(u1, std::floor(t.region[0]) + 0.5f), std::floor(t.region[2]) + 0.5f);
	if (clamp_t)
		v1 = std::min(std::max(v1, std::floor(t.region[1]) + 0.5f), std::floor(t.region[3]) + 0.5f);
	float un = std::floor(u1 * 16.0f + 1.0f / 64.0f);
	float vn = std::floor(v1 * 16.0f + 1.0f / 64.0f);
	un = un + 16.0f / 1024.0f;
	vn = vn + 16.0f / 1024.0f;
	un = un * t.size[2];
	vn = vn * t.size[3];
	un = un * 0.0625f;
	vn = vn * 0.0625f;
	un = un + t.bias[0];
	vn = vn + t.bias[1];

	const float fw = float(t.width), fh = float(t.height);
	uint32_t c;
	if (!linear)
	{
		float xs = std::floor(un * fw), ys = std::floor(vn * fh);
		c = texel(t, int32_t(xs), int32_t(ys));
	}
	else
	{
		float xs = un * fw - 0.5f, ys = vn * fh - 0.5f;
		float x0 = std::floor(xs), y0 = std::floor(ys);
		uint32_t wx = uint32_t((xs - x0) * 256.0f), wy = uint32_t((ys - y0) * 256.0f);
		uint32_t c00 = texel(t, int32_t(x0), int32_t(y0)), c10 = texel(t, int32_t(x0) + 1, int32_t(y0));
		uint32_t c01 = texel(t, int32_t(x0), int32_t(y0) + 1), c11 = texel(t, int32_t(x0) + 1, int32_t(y0) + 1);
		c = 0;
		for (uint32_t ch = 0; ch < 4; ch++)
		{
			uint32_t sh = ch * 8u;
			uint32_t top = lerp8((c00 >> sh) & 0xffu, (c10 >> sh) & 0xffu, wx);
			uint32_t bot = lerp8((c01 >> sh) & 0xffu, (c11 >> sh) & 0xffu, wx);
			c |= lerp8(top, bot, wy) << sh;
		}
	}
	for (uint32_t ch = 0; ch < 4; ch++)
		out[ch] = float((c >> (8 * ch)) & 0xffu);
}
}
