// CPU oracle: primitive assembly, triangle setup and per-sample shading
// (TEST INFRASTRUCTURE ONLY, see oracle_core.hpp).
//
// Restated (not copied) from the reference:
//   gs/gs_interface.cpp:858-899, 2036-2115, 2473-2928   vertex kick, degenerate draws, assembly
//   gs/gs_interface.cpp:1065-1128, 1804-1968            state vector and primitive state bits
//   gs/gs_util.cpp:237-377                              parallelogram fusion
//   gs/shaders/triangle_setup.comp:127-873              triangle setup
//   gs/shaders/intersect.h, math_utils.h, utils.h       coverage, reciprocals, colour packing
//   gs/shaders/ubershader.comp:435-1638                 per-sample shading and tile write-back
//
// The oracle renders a render pass as one pixel loop per primitive, in submission order,
// against per-pixel "register" copies of colour and depth that are loaded on first touch and
// written back when the pass ends. That is what the reference's tile shader computes for
// every pixel (its early-Z / deferred-shading bookkeeping only skips work whose result is
// overwritten), without any of its scheduling.
#include "oracle_render.hpp"
#include <cstring>
#include <cmath>
#include <algorithm>
#include <unordered_map>

namespace orc
{
namespace
{
// State bits (gs/shaders/data_structures.h:206-307).
enum : uint32_t
{
	S_Z_TEST = 1u << 12, S_Z_WRITE = 1u << 13, S_OPAQUE = 1u << 14, S_PARALLELOGRAM = 1u << 15,
	S_MULTISAMPLE = 1u << 16, S_Z_GREATER = 1u << 17, S_PERSPECTIVE = 1u << 18, S_IIP = 1u << 19,
	S_FIX = 1u << 20, S_SPRITE = 1u << 21, S_LINE = 1u << 22, S_SCANMSK_EVEN = 1u << 23,
	S_SCANMSK_ODD = 1u << 24, S_SNAP_RASTER = 1u << 25, S_SNAP_ATTRIBUTE = 1u << 26,
	S_PROVOKING_SHIFT = 27,
};
enum : uint32_t
{
	T_MAG_LINEAR = 1u << 16, T_MIN_LINEAR = 1u << 17, T_MIP_LINEAR = 1u << 18, T_CLAMP_S = 1u << 19,
	T_CLAMP_T = 1u << 20,
	// data_structures.h:251-253
	T_PER_SAMPLE = 1u << 24, T_SAMPLE_MAPPING = 1u << 25, T_SAMPLE_RESOLVED = 1u << 26,
};
enum : uint32_t
{
	C_TME = 1u << 2, C_FOG = 1u << 3, C_TCC = 1u << 4,
	B_ATE = 1u << 1, B_DATE = 1u << 7, B_DATM = 1u << 8, B_ABE = 1u << 9, B_PABE = 1u << 10,
	B_COLCLAMP = 1u << 19, B_FBA = 1u << 20, B_DTHE = 1u << 21,
};

inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline uint32_t field(uint64_t v, uint32_t off, uint32_t bits) { return uint32_t((v >> off) & ((1ull << bits) - 1ull)); }

// The 256-entry 8-bit seed of the fixed-point Newton reciprocal. It is the reference's
// hand-tuned table (gs/gs_renderer.cpp:3602-3636); the integer iteration below only gives the
// reference's results with exactly these seeds.
const uint8_t kRcpSeed[256] = {
	255, 253, 251, 249, 247, 245, 243, 241, 240, 238, 236, 234, 232, 230, 229, 227,
	225, 223, 221, 220, 218, 216, 215, 213, 211, 210, 208, 206, 205, 203, 201, 200,
	198, 197, 195, 194, 192, 191, 189, 188, 186, 185, 183, 182, 180, 179, 177, 176,
	174, 173, 172, 170, 169, 167, 166, 165, 164, 162, 161, 159, 158, 157, 156, 154,
	153, 152, 150, 149, 148, 147, 145, 144, 143, 142, 140, 139, 138, 137, 136, 135,
	134, 132, 131, 130, 129, 128, 127, 126, 124, 123, 122, 121, 120, 119, 118, 116,
	116, 115, 113, 113, 112, 111, 110, 109, 107, 107, 106, 104, 103, 103, 102, 101,
	100,  99,  98,  97,  96,  95,  94,  93,  92,  91,  91,  89,  88,  88,  86,  86,
	 85,  84,  83,  82,  81,  81,  80,  79,  78,  78,  76,  75,  75,  74,  72,  72,
	 72,  70,  70,  69,  68,  67,  67,  66,  65,  64,  63,  63,  62,  61,  60,  59,
	 59,  58,  57,  56,  56,  55,  54,  54,  53,  52,  52,  51,  50,  49,  48,  48,
	 47,  46,  46,  45,  45,  44,  43,  42,  42,  41,  40,  39,  39,  38,  38,  37,
	 36,  35,  35,  34,  33,  33,  33,  32,  31,  31,  30,  29,  29,  28,  28,  27,
	 26,  25,  25,  25,  24,  23,  23,  22,  22,  21,  20,  19,  19,  18,  18,  17,
	 17,  16,  15,  15,  14,  14,  14,  13,  12,  12,  11,  11,  10,  10,   9,   8,
	  8,   7,   7,   7,   6,   5,   5,   4,   4,   3,   3,   2,   2,   1,   1,   1,
};

inline uint32_t umulhi(uint32_t a, uint32_t b) { return uint32_t((uint64_t(a) * b) >> 32); }

// Integer helpers with the 32-bit wrap-around of GLSL int arithmetic.
inline int32_t wadd(int32_t a, int32_t b) { return int32_t(uint32_t(a) + uint32_t(b)); }
inline int32_t wmul(int32_t a, int32_t b) { return int32_t(uint32_t(a) * uint32_t(b)); }
inline int32_t idot3(const int32_t p[3], int32_t x, int32_t y) { return wadd(wadd(wmul(p[0], x), wmul(p[1], y)), p[2]); }

// packUnorm4x8 / unpackUnorm4x8: c / 255 and round(clamp(c) * 255) with ties to even.
inline float unorm(uint32_t c, uint32_t ch) { return float((c >> (8 * ch)) & 0xffu) / 255.0f; }
inline uint32_t pack_unorm(const float v[4])
{
	uint32_t r = 0;
	for (int i = 0; i < 4; i++)
	{
		float c = std::min(std::max(v[i], 0.0f), 1.0f);
		r |= uint32_t(std::nearbyint(c * 255.0f)) << (8 * i);
	}
	return r;
}

// utils.h:182-208
uint32_t rgba16_to_rgba32(uint32_t c16, uint32_t aem, uint32_t ta0, uint32_t ta1)
{
	uint32_t r = (c16 & 0x1fu) << 3, g = ((c16 >> 5) & 0x1fu) << 3, b = ((c16 >> 10) & 0x1fu) << 3;
	uint32_t abit = (c16 >> 15) & 1u;
	bool zero_alpha = aem && c16 == 0;
	uint32_t a = zero_alpha ? 0u : (abit ? ta1 : ta0);
	return r | g << 8 | b << 16 | std::min(a, 255u) << 24;
}
uint32_t rgba32_to_rgba16(uint32_t c)
{
	return ((c >> 3) & 0x1fu) | (((c >> 11) & 0x1fu) << 5) | (((c >> 19) & 0x1fu) << 10) | ((c >> 31) << 15);
}
uint32_t rgb24_to_rgba32(uint32_t c, uint32_t aem, uint32_t ta0)
{
	bool zero_alpha = aem && (c & 0xffffffu) == 0;
	return (c & 0xffffffu) | ((zero_alpha ? 0u : ta0) << 24);
}

bool psm16(uint32_t p) { return p == CT16 || p == CT16S || p == Z16 || p == Z16S; }
bool psm24(uint32_t p) { return p == CT24 || p == Z24; }
bool psm32(uint32_t p) { return p == CT32 || p == Z32; }
}

const uint8_t *rcp_table() { return kRcpSeed; }

// math_utils.h:27-61: two integer Newton steps from an 8-bit seed, rounded to FP32.
float rcp_fixed(uint32_t b)
{
	uint32_t lz = b ? uint32_t(__builtin_clz(b)) : 32u;
	uint32_t nb = lz < 32 ? (b << lz) : 0u;
	uint32_t idx = (nb >> 23) & 0xffu;
	uint32_t r = kRcpSeed[idx] | 0x100u;
	uint32_t f = (r * r) << 14;
	uint32_t p = umulhi(f, nb);
	r = (r << 23) - p;
	r <<= 1;
	p = umulhi(r, nb);
	r = umulhi(r, 0u - p);
	r = (r + 64u) >> 7;
	r -= 1u << 23;
	r += uint32_t(127 - 32 + int32_t(lz)) << 23;
	return bitsf(r);
}

// math_utils.h:65-101 with the seed LUT 1 / (1 + (i + 0.5) / 256) (gs_renderer.cpp:3638-3640).
float rcp_float(float v)
{
	static float lut[256];
	static bool init = false;
	if (!init)
	{
		for (int i = 0; i < 256; i++)
			lut[i] = 1.0f / (1.0f + (float(i) + 0.5f) / 256.0f);
		init = true;
	}
	uint32_t u = fbits(v);
	int32_t e = int32_t((u >> 23) & 0xffu);
	if (e == 0)
		return bitsf(0x7f800000u);
	int32_t shifted = 127 - e;
	float x = lut[(u >> 15) & 0xffu];
	float a = bitsf((u & 0x7fffffu) | (127u << 23));
	for (int it = 0; it < 4; it++)
	{
		float x2 = x * 2.0f;
		float sq = x * x;
		float sqa = sq * a;
		x = x2 - sqa;
	}
	return std::ldexp(x, shifted);
}

// ---------------------------------------------------------------------------------------
// Vertex kicks and primitive assembly.

void GS::write_tex0(uint32_t ctx, uint64_t v)
{
	set_reg(regs.tex0[ctx], v, true);
	load_clut(ctx);
}

// TEX2 rewrites the CLUT/PSM part of TEX0 only (gs/gs_interface.cpp:2950-2968).
void GS::write_tex2(uint32_t ctx, uint64_t v)
{
	const uint64_t keep = (0x3fffull) | (0x3full << 14) | (0xfull << 26) | (0xfull << 30) | (1ull << 34) |
	                      (3ull << 35);
	uint64_t merged = (regs.tex0[ctx] & keep) | (v & ~keep);
	write_tex0(ctx, merged);
}

void GS::push_vertex(int32_t x, int32_t y, uint32_t z, float fog)
{
	if (vq_count == 3)
	{
		vq[0] = vq[1];
		vq[1] = vq[2];
		vq_count = 2;
	}
	uint32_t ctxt = field(regs.prim, 9, 1);
	Vertex &v = vq[vq_count++];
	v.x = x - int32_t(field(regs.xyoffset[ctxt], 0, 16));
	v.y = y - int32_t(field(regs.xyoffset[ctxt], 32, 16));
	v.z = z;
	v.s = regs.s;
	v.t = regs.t;
	v.q = regs.rgbaq_q;
	v.rgba = regs.rgbaq;
	v.fog = fog;
	v.u = field(regs.uv, 0, 14);
	v.v = field(regs.uv, 16, 14);
}

void GS::kick_xyz(uint64_t v, bool adc)
{
	push_vertex(int32_t(v & 0xffff), int32_t((v >> 16) & 0xffff), uint32_t(v >> 32), float(regs.fog));
	drawing_kick(adc);
}

void GS::kick_xyzf(uint64_t v, bool adc)
{
	push_vertex(int32_t(v & 0xffff), int32_t((v >> 16) & 0xffff), uint32_t((v >> 32) & 0xffffff), float(v >> 56));
	drawing_kick(adc);
}

// gs/gs_interface.cpp:2838-2928
void GS::drawing_kick(bool adc)
{
	uint32_t type = field(regs.prim, 0, 3);
	static const uint32_t nverts[8] = { 1, 2, 2, 3, 3, 3, 2, 0 };
	uint32_t n = nverts[type];
	if (n == 0)
	{
		vq_count = 0;
		return;
	}
	if (vq_count < n)
		return;

	if (!adc && !draw_is_degenerate())
	{
		Vertex v[3];
		if (n == 1)
		{
			// Points: a one-pixel sprite whose bottom-right wins on half-pixel positions.
			v[0] = vq[vq_count - 1];
			v[0].x -= 7;
			v[0].y -= 7;
			v[1] = v[0];
			v[1].x += 16;
			v[1].y += 16;
			v[2] = v[0];
		}
		else if (n == 2)
		{
			v[0] = vq[vq_count - 1];
			v[1] = vq[vq_count - 2];
			v[2] = v[0];
		}
		else
		{
			v[0] = vq[2];
			v[1] = vq[1];
			v[2] = vq[0];
		}
		append(v, n, type == 0 || type == 6);
	}

	if (type == 5)
	{
		vq[1] = vq[2];
		vq_count = 2;
	}
	else if (type == 0 || type == 1 || type == 3 || type == 6)
		vq_count = 0;
}

// gs/gs_interface.cpp:2036-2115
bool GS::draw_is_degenerate() const
{
	uint32_t ctx = field(regs.prim, 9, 1);
	uint64_t sc = regs.scissor[ctx], test = regs.test[ctx], frame = regs.frame[ctx], zbuf = regs.zbuf[ctx];
	uint64_t alpha = regs.alpha[ctx];
	if (field(sc, 0, 11) > field(sc, 16, 11) || field(sc, 32, 11) > field(sc, 48, 11))
		return true;
	uint32_t ate = field(test, 0, 1), atst = field(test, 1, 3), afail = field(test, 12, 2);
	uint32_t zte = field(test, 16, 1), ztst = field(test, 17, 2);
	if (zte && ztst == 0)
		return true;
	if (ate && atst == 0 && afail == 0)
		return true;

	uint32_t fbmsk = uint32_t(frame >> 32);
	uint32_t zmsk = field(zbuf, 32, 1);
	uint32_t psm = field(frame, 24, 6);
	if (psm == CT24 || psm == Z24)
		fbmsk |= 0xffu << 24;
	if (ate && atst == 0)
	{
		if (afail == 2) fbmsk = ~0u;
		else if (afail == 0) { fbmsk = ~0u; zmsk = 1; }
		else if (afail == 1) zmsk = 1;
		else if (afail == 3) fbmsk |= 0xffu << 24;
	}
	uint32_t abe = field(regs.prim, 6, 1);
	if (abe && field(alpha, 6, 2) == 1 && field(alpha, 0, 2) == field(alpha, 2, 2))
		fbmsk |= 0xffffffu;
	bool ro_color = fbmsk == ~0u;
	bool ro_depth = zmsk || zte == 0;
	return ro_color && ro_depth;
}

namespace
{
// gs/gs_util.cpp:237-310
bool parallelogram_candidate(const Vertex *v, int32_t lox, int32_t loy, int32_t hix, int32_t hiy, uint64_t prim,
                             int32_t order[3])
{
	bool aa1 = field(prim, 7, 1) && field(prim, 0, 3) != 6 && field(prim, 0, 3) != 0;
	if (aa1)
		return false;
	bool fst = field(prim, 8, 1), tme = field(prim, 4, 1), fge = field(prim, 5, 1);
	if (!fst && tme && (v[0].q != v[1].q || v[1].q != v[2].q))
		return false;
	if (v[0].z != v[1].z || v[1].z != v[2].z)
		return false;
	if (fge && (v[0].fog != v[1].fog || v[1].fog != v[2].fog))
		return false;
	if (v[0].rgba != v[1].rgba || v[1].rgba != v[2].rgba)
		return false;
	int32_t abx = v[1].x - v[0].x, aby = v[1].y - v[0].y;
	int32_t acx = v[2].x - v[0].x, acy = v[2].y - v[0].y;
	int32_t bcx = v[2].x - v[1].x, bcy = v[2].y - v[1].y;
	int32_t area = std::abs(wadd(wmul(abx, acy), -wmul(aby, acx)));
	if (area != wmul(hix - lox, hiy - loy))
		return false;
	if (abx != 0 && aby != 0)
	{
		order[0] = 2; order[1] = 1; order[2] = 0;
		if (acx * bcx + acy * bcy != 0)
			return false;
	}
	else if (acx != 0 && acy != 0)
	{
		order[0] = 1; order[1] = 2; order[2] = 0;
		if (acx * abx + acy * aby != 0)
			return false;
	}
	else
	{
		order[0] = 0; order[1] = 1; order[2] = 2;
		if (acx * abx + acy * aby != 0)
			return false;
	}
	if (v[order[1]].x > v[order[2]].x)
		std::swap(order[1], order[2]);
	return true;
}

// gs/gs_util.cpp:312-377
bool triangles_form_parallelogram(const Vertex *v, const int32_t o[3], const Vertex *lv, const int32_t lo[3],
                                  uint64_t prim)
{
	int32_t p3x = v[o[1]].x + v[o[2]].x - v[o[0]].x, p3y = v[o[1]].y + v[o[2]].y - v[o[0]].y;
	if (p3x != lv[lo[0]].x || p3y != lv[lo[0]].y || v[o[1]].x != lv[lo[1]].x || v[o[1]].y != lv[lo[1]].y ||
	    v[o[2]].x != lv[lo[2]].x || v[o[2]].y != lv[lo[2]].y)
		return false;
	bool fst = field(prim, 8, 1), tme = field(prim, 4, 1), fge = field(prim, 5, 1);
	if (v[0].z != lv[0].z)
		return false;
	if (!fst && tme && v[0].q != lv[0].q)
		return false;
	if (fge && v[0].fog != lv[0].fog)
		return false;
	if (v[0].rgba != lv[0].rgba)
		return false;
	if (tme)
	{
		if (fst)
		{
			uint32_t u3 = (v[o[1]].u + v[o[2]].u - v[o[0]].u) & 0xffffu;
			uint32_t v3 = (v[o[1]].v + v[o[2]].v - v[o[0]].v) & 0xffffu;
			if (u3 != lv[lo[0]].u || v3 != lv[lo[0]].v || v[o[1]].u != lv[lo[1]].u || v[o[1]].v != lv[lo[1]].v ||
			    v[o[2]].u != lv[lo[2]].u || v[o[2]].v != lv[lo[2]].v)
				return false;
		}
		else
		{
			float es = std::fabs((v[o[2]].s + v[o[1]].s - v[o[0]].s - lv[lo[0]].s) / v[0].q);
			float et = std::fabs((v[o[2]].t + v[o[1]].t - v[o[0]].t - lv[lo[0]].t) / v[0].q);
			if (v[o[1]].s != lv[lo[1]].s || v[o[1]].t != lv[lo[1]].t || v[o[2]].s != lv[lo[2]].s ||
			    v[o[2]].t != lv[lo[2]].t || es > 1e-4f || et > 1e-4f)
				return false;
		}
	}
	return true;
}
}

// gs/gs_interface.cpp:2473-2836 (single frame-buffer instance per render pass; a frame or
// depth buffer change ends the pass).
// On-tile feedback texel (ubershader.comp:955-1040): the pixel's (or sample's) current colour or
// depth read as the texture's format. Palette formats take the index from the word's top bits
// (8H: 8 bits, 4HL / 4HH: the nibbles at 24 and 28) plus CSA * 16 and look the CLUT up (8-bit
// indices past the last bank clamp to it; a 32-bit CLUT has 16 banks and two halves 256 entries
// apart); 24-bit formats get TA0 as alpha, 16-bit ones TA0 or TA1 by the alpha bit, AEM zeroing
// alpha for black.
void GS::feedback_texture(const Prim &p, uint32_t in, float out[4]) const
{
	const uint32_t psm = p.fb_tex_psm;
	uint32_t cpsm = p.fb_tex_cpsm == CT24 ? CT32 : p.fb_tex_cpsm;
	const uint32_t aem = field(p.fb_texa, 15, 1), ta0 = field(p.fb_texa, 0, 8), ta1 = field(p.fb_texa, 32, 8);
	uint32_t c;
	if (psm == T8 || psm == T4 || psm == T8H || psm == T4HL || psm == T4HH)
	{
		uint32_t csa = p.fb_csa;
		if (cpsm == CT32)
			csa &= 15u;
		uint32_t idx = 0;
		if (psm == T8H) idx = in >> 24;
		else if (psm == T4HL) idx = (in >> 24) & 0xfu;
		else if (psm == T4HH) idx = in >> 28;
		idx += csa * 16u;
		if (psm == T8 || psm == T8H)
		{
			uint32_t bank = std::min(idx >> 4, cpsm == CT32 ? 15u : 31u);
			idx = bank * 16u + (idx & 15u);
		}
		if (cpsm == CT32)
			c = uint32_t(clut[idx]) | (uint32_t(clut[idx + 256u]) << 16);
		else
			c = rgba16_to_rgba32(clut[idx], aem, ta0, ta1);
	}
	else
	{
		c = in;
		const bool zero_alpha = aem && (c & 0xffffffu) == 0;
		if (psm == CT24 || psm == Z24)
			c = (c & 0xffffffu) | ((zero_alpha ? 0u : ta0) << 24);
		else if (psm == CT16 || psm == CT16S || psm == Z16 || psm == Z16S)
			c = (c & 0xffffffu) | ((zero_alpha ? 0u : ((c >> 24) == 0 ? ta0 : ta1)) << 24);
	}
	for (uint32_t ch = 0; ch < 4; ch++)
		out[ch] = float((c >> (8 * ch)) & 0xffu);
}

// Frame-buffer feedback classification (gs/gs_interface.cpp:2141-2307 and 2388-2470).
namespace
{
uint32_t compat_key(uint32_t psm)
{
	switch (psm)
	{
	case Z32: return 2;
	case CT16: return 3;
	case Z16: return 4;
	case CT16S: return 5;
	case Z16S: return 6;
	case T8: return 7;
	case T4: return 8;
	case Z24: return 9;
	default: return 1;
	}
}

uint32_t word_mask(uint32_t psm)
{
	switch (psm)
	{
	case CT24: case Z24: return 0xffffffu;
	case T8H: return 0xff000000u;
	case T4HL: return 0x0f000000u;
	case T4HH: return 0xf0000000u;
	default: return 0xffffffffu;
	}
}
}

uint32_t GS::feedback_kind(uint32_t ctx, const Vertex *v, uint32_t n, bool quad, const int32_t bb[4],
                           bool &depth) const
{
	depth = false;
	const uint64_t t0 = regs.tex0[ctx], t1 = regs.tex1[ctx], cl = regs.clamp[ctx];
	const uint64_t frame = regs.frame[ctx], zbuf = regs.zbuf[ctx], test = regs.test[ctx];
	const uint32_t mmin = field(t1, 6, 3);
	if (mmin >= 2 && mmin <= 5 && field(t1, 2, 3) != 0)
		return 0;
	const uint32_t psm = field(t0, 20, 6), tbw = field(t0, 14, 6), fpsm = field(frame, 24, 6);
	const uint32_t wms = field(cl, 0, 2), wmt = field(cl, 2, 2);
	const bool palette = psm == T8 || psm == T4 || psm == T8H || psm == T4HL || psm == T4HH;
	uint32_t base = field(t0, 0, 14);
	if (!palette)
	{
		// Page size of the texture format: 64 or 128 pixels wide, 32, 64 or 128 high.
		uint32_t pw = 6, ph = 5;
		if (psm == CT16 || psm == CT16S || psm == Z16 || psm == Z16S) ph = 6;
		if (wms == 2 && field(cl, 4, 10))
			base += 32u * (field(cl, 4, 10) >> pw);
		if (wmt == 2 && field(cl, 24, 10))
			base += 32u * tbw * (field(cl, 24, 10) >> ph);
	}
	bool z_sensitive = false;
	for (const Prim &q : rp_prims)
		z_sensitive |= (q.state & (S_Z_TEST | S_Z_WRITE)) != 0;
	if (field(test, 16, 1))
	{
		const bool implied_mask = field(test, 0, 1) && field(test, 1, 3) == 0 && field(test, 12, 2) != 2;
		if (field(test, 17, 2) > 1 || (!implied_mask && !field(zbuf, 32, 1)))
			z_sensitive = true;
	}
	const bool on_color = base == field(frame, 0, 9) * 32u;
	const bool on_depth = base == field(zbuf, 0, 9) * 32u && z_sensitive;
	const bool awkward = on_color && (compat_key(psm) != compat_key(fpsm) || tbw != field(frame, 16, 6));
	if (!awkward)
	{
		if (on_depth && (compat_key(psm) != compat_key(field(zbuf, 24, 4) | 0x30u) || tbw != field(frame, 16, 6)))
			return 0;
		if (wms == 3 || wmt == 3 || (!on_color && !on_depth))
			return 0;
	}
	const bool color_fb = on_color && (awkward || (word_mask(psm) & word_mask(fpsm)));
	const bool depth_fb = !awkward && on_depth && (word_mask(psm) & word_mask(field(zbuf, 24, 4)));
	if (!color_fb && !depth_fb)
		return 0;
	// Affine triangles and sprites only.
	if (!(n == 3 || (quad && n == 2)))
		return 0;
	if (awkward)
		return 2;
	const int32_t tw = 1 << field(t0, 26, 4), th = 1 << field(t0, 30, 4);
	const uint32_t nv = quad ? 2 : 3;
	int32_t uv[3][2];
	if (field(regs.prim, 8, 1))
	{
		for (uint32_t i = 0; i < nv; i++)
		{
			uv[i][0] = int32_t(v[i].u);
			uv[i][1] = int32_t(v[i].v);
		}
	}
	else
	{
		if (!quad && (v[0].q != v[1].q || v[1].q != v[2].q))
			return 0;
		for (uint32_t i = 0; i < nv; i++)
		{
			const float inv_q = 1.0f / v[i].q;
			const float fu = float(tw * 16) * (v[i].s * inv_q), fv = float(th * 16) * (v[i].t * inv_q);
			if (!(std::fabs(fu) < 1073741824.0f) || !(std::fabs(fv) < 1073741824.0f))
				return 0;
			uv[i][0] = int32_t(fu);
			uv[i][1] = int32_t(fv);
		}
	}
	if (wms == 2)
	{
		if (bb[0] < int32_t(field(cl, 4, 10)) || bb[2] > int32_t(field(cl, 14, 10)))
			return 0;
	}
	else if (bb[2] >= tw)
		return 0;
	if (wmt == 2)
	{
		if (bb[1] < int32_t(field(cl, 24, 10)) || bb[3] > int32_t(field(cl, 34, 10)))
			return 0;
	}
	else if (bb[3] >= th)
		return 0;
	int32_t lo = 1 << 30, hi = -(1 << 30);
	for (uint32_t i = 0; i < nv; i++)
	{
		const int32_t d[2] = { uv[i][0] - v[i].x, uv[i][1] - v[i].y };
		for (int32_t dd : d)
		{
			lo = std::min(lo, dd);
			hi = std::max(hi, dd);
		}
	}
	const bool ok = field(t1, 5, 1) ? (lo == 8 && hi == 8) : (lo >= 0 && hi < 16);
	if (!ok)
		return 0;
	depth = !color_fb;
	return 1;
}

void GS::append(const Vertex *vin, uint32_t n, bool quad)
{
	Vertex v[3] = { vin[0], vin[1], vin[2] };
	const uint64_t prim = regs.prim;
	const uint32_t ctx = field(prim, 9, 1);
	const bool is_line = !quad && n == 2;

	int32_t lox = std::min(v[0].x, v[1].x), loy = std::min(v[0].y, v[1].y);
	int32_t hix = std::max(v[0].x, v[1].x), hiy = std::max(v[0].y, v[1].y);
	if (!quad && !is_line)
	{
		lox = std::min(lox, v[2].x); loy = std::min(loy, v[2].y);
		hix = std::max(hix, v[2].x); hiy = std::max(hiy, v[2].y);
	}
	const int32_t pre_lo[2] = { lox, loy }, pre_hi[2] = { hix, hiy };
	hix -= 1; hiy -= 1;
	lox += (1 << (4 - int32_t(rate_y_log2()))) - 1;
	loy += (1 << (4 - int32_t(rate_y_log2()))) - 1;
	lox >>= 4; loy >>= 4; hix >>= 4; hiy >>= 4;
	bool aa1 = field(prim, 7, 1) && field(prim, 0, 3) != 6 && field(prim, 0, 3) != 0;
	if (is_line || aa1)
	{
		lox -= 1; loy -= 1; hix += 1; hiy += 1;
	}
	uint64_t sc = regs.scissor[ctx];
	lox = std::max(lox, int32_t(field(sc, 0, 11)));
	loy = std::max(loy, int32_t(field(sc, 32, 11)));
	hix = std::min(hix, int32_t(field(sc, 16, 11)));
	hiy = std::min(hiy, int32_t(field(sc, 48, 11)));
	if (hix < lox || hiy < loy)
		return;

	// A partial host->local transfer is flushed before a texture read
	// (gs/gs_interface.cpp:2640-2650).
	if (field(prim, 4, 1) && xfer.active)
		partial_flush_transfer();

	// Frame / depth buffer change ends the render pass.
	const uint64_t frame = regs.frame[ctx], zbuf = regs.zbuf[ctx];
	const uint64_t fb_key = frame & 0x3f3f01ffull;       // FBP, FBW, PSM
	const uint64_t z_key = zbuf & 0x0f0001ffull;         // ZBP, PSM
	if (!rp_prims.empty())
	{
		bool pass_z = false;
		for (auto &p : rp_prims)
			pass_z |= (p.state & (S_Z_TEST | S_Z_WRITE)) != 0;
		if ((rp_frame & 0x3f3f01ffull) != fb_key || ((rp_zbuf & 0x0f0001ffull) != z_key && pass_z))
			flush_render_pass();
	}

	bool candidate = false;
	int32_t order[3] = { 0, 1, 2 };
	if (n == 3)
	{
		candidate = parallelogram_candidate(v, pre_lo[0], pre_lo[1], pre_hi[0], pre_hi[1], prim, order);
		if (!state_dirty && candidate && last_parallelogram_candidate && !rp_prims.empty() &&
		    triangles_form_parallelogram(v, order, rp_prims.back().v, last_order, prim))
		{
			rp_prims.back().state |= S_PARALLELOGRAM | (uint32_t(last_order[0]) << S_PROVOKING_SHIFT);
			last_parallelogram_candidate = false;
			return;
		}
		memcpy(last_order, order, sizeof(order));
	}

	// A texture read of pages the pending pass writes renders the pass first, so the texture
	// sees every earlier primitive (the reference's page tracker, gs/page_tracker.cpp, orders
	// texture reads after pending writes the same way). Fused parallelograms skip the check,
	// as in the reference (gs/gs_interface.cpp:2570-2599).
	uint32_t fbk = 0;
	bool fb_depth = false;
	if (field(prim, 4, 1))
	{
		const int32_t bbox[4] = { lox, loy, hix, hiy };
		fbk = feedback_kind(ctx, v, n, quad, bbox, fb_depth);
	}
	if (fbk == 1)
	{
		// A pass holds one kind and texture format of on-tile feedback
		// (gs/gs_interface.cpp:909-933).
		const uint64_t t0 = regs.tex0[ctx];
		const uint32_t psm = field(t0, 20, 6);
		const bool palette = psm == T8 || psm == T4 || psm == T8H || psm == T4HL || psm == T4HH;
		const uint32_t cpsm = palette ? field(t0, 51, 4) : 0u;
		for (const Prim &q : rp_prims)
			if (q.feedback)
			{
				if (q.feedback != (fb_depth ? 2u : 1u) || q.fb_tex_psm != psm || (palette && q.fb_tex_cpsm != cpsm))
					flush_render_pass();
				break;
			}
	}
	else if (field(prim, 4, 1) && fbk == 0 && !rp_prims.empty() && texture_hazard(ctx))
		flush_render_pass();

	if (rp_prims.empty())
	{
		rp_frame = frame;
		rp_zbuf = zbuf;
	}
	rp_zbuf = (rp_zbuf & ~0x0f0001ffull) | (zbuf & 0x0f0001ffull);

	Prim p{};
	memcpy(p.v, v, sizeof(v));
	p.bb[0] = lox; p.bb[1] = loy; p.bb[2] = hix; p.bb[3] = hiy;

	const uint64_t test = regs.test[ctx], alpha = regs.alpha[ctx];
	const uint32_t ate = field(test, 0, 1), atst = field(test, 1, 3), aref = field(test, 4, 8);
	const uint32_t afail = field(test, 12, 2), date = field(test, 14, 1), datm = field(test, 15, 1);
	const uint32_t zte = field(test, 16, 1), ztst = field(test, 17, 2);
	const bool abe = field(prim, 6, 1), tme = field(prim, 4, 1);

	// State vector (gs/gs_interface.cpp:1065-1128).
	uint32_t blend = 0, comb = 0;
	p.dimx[0] = p.dimx[1] = 0;
	if (field(regs.dthe, 0, 1))
	{
		blend |= B_DTHE;
		p.dimx[0] = uint32_t(regs.dimx);
		p.dimx[1] = uint32_t(regs.dimx >> 32);
	}
	if (ate && atst != 1 && !(atst == 0 && afail == 1))
		blend |= B_ATE | (atst << 2) | (afail << 5);
	blend |= date ? B_DATE : 0;
	blend |= datm ? B_DATM : 0;
	if (abe || aa1)
		blend |= field(alpha, 0, 2) << 11 | field(alpha, 2, 2) << 13 | field(alpha, 4, 2) << 15 | field(alpha, 6, 2) << 17;
	if (abe)
		blend |= B_ABE | (field(regs.pabe, 0, 1) ? B_PABE : 0);
	blend |= field(regs.colclamp, 0, 1) ? B_COLCLAMP : 0;
	blend |= field(regs.fba[ctx], 0, 1) ? B_FBA : 0;
	if (tme)
	{
		uint64_t t0 = regs.tex0[ctx];
		comb |= C_TME | (field(t0, 34, 1) ? C_TCC : 0) | field(t0, 35, 2);
	}
	comb |= field(prim, 5, 1) ? C_FOG : 0;
	p.combiner = comb;
	p.blend_mode = blend;

	// Primitive state (gs/gs_interface.cpp:1804-1968).
	uint32_t state = 0;
	if (fbk == 1)
	{
		// The texel is the pixel's own colour or depth (ubershader.comp:955-1040).
		const uint64_t t0 = regs.tex0[ctx];
		p.feedback = fb_depth ? 2u : 1u;
		p.fb_tex_psm = field(t0, 20, 6);
		const bool palette = p.fb_tex_psm == T8 || p.fb_tex_psm == T4 || p.fb_tex_psm == T8H ||
		                     p.fb_tex_psm == T4HL || p.fb_tex_psm == T4HH;
		p.fb_tex_cpsm = palette ? field(t0, 51, 4) : 0u;
		p.fb_csa = field(t0, 56, 5);
		p.fb_texa = regs.texa;
		p.tex = 1u << 15;
		// Sampler bits as for a texture (triangle setup reads the magnification filter).
		const uint64_t t1 = regs.tex1[ctx], cl = regs.clamp[ctx];
		p.tex |= field(t1, 5, 1) ? T_MAG_LINEAR : 0;
		p.tex |= (field(cl, 0, 2) == 1 || field(cl, 0, 2) == 2) ? T_CLAMP_S : 0;
		p.tex |= (field(cl, 2, 2) == 1 || field(cl, 2, 2) == 2) ? T_CLAMP_T : 0;
		p.tex2 = 1u;
	}
	else if (tme)
	{
		p.tex_slot = texture_slot(ctx);
		uint64_t t1 = regs.tex1[ctx], cl = regs.clamp[ctx];
		p.tex = p.tex_slot;
		p.tex |= field(t1, 5, 1) ? T_MAG_LINEAR : 0;
		uint32_t wms = field(cl, 0, 2), wmt = field(cl, 2, 2);
		p.tex |= (wms == 1 || wms == 2) ? T_CLAMP_S : 0;
		p.tex |= (wmt == 1 || wmt == 2) ? T_CLAMP_T : 0;
		// Mipmapped filters (gs/gs_interface.cpp:1834-1863): LCM | L << 1 | K << 3 in tex2, MXL
		// at bit 21 of tex with the MMIN filter bits; otherwise a fixed LOD of 0.
		const uint32_t mmin = field(t1, 6, 3);
		if (mmin >= 2 && mmin <= 5)
		{
			p.tex2 = field(t1, 0, 1) | field(t1, 19, 2) << 1 | field(t1, 32, 12) << 3;
			p.tex |= field(t1, 2, 3) << 21;
			p.tex |= (mmin == 4 || mmin == 5) ? (1u << 17) : 0;
			p.tex |= (mmin == 3 || mmin == 5) ? (1u << 18) : 0;
		}
		else
			p.tex2 = 1u;
		// A super-sampled texture samples per sample; palette formats with REGION_REPEAT force
		// one layer per sample (gs/gs_interface.cpp:1783-1790, 1824-1831).
		if (rp_textures[p.tex_slot].samples)
		{
			const uint32_t psm = field(regs.tex0[ctx], 20, 6);
			const bool palette = psm == T8 || psm == T4 || psm == T8H || psm == T4HL || psm == T4HH;
			p.tex |= T_PER_SAMPLE;
			if (palette && (wms == 3 || wmt == 3))
				p.tex |= T_SAMPLE_MAPPING;
			rp_arrays = true;
		}
	}
	if (zte)
	{
		if (ztst > 1)
			state |= S_Z_TEST | (ztst == 3 ? S_Z_GREATER : 0);
		bool implied_z_mask = ate && atst == 0 && afail != 2;
		if (!implied_z_mask && field(zbuf, 32, 1) == 0)
			state |= S_Z_WRITE;
	}
	bool needs_prev = false;
	if (abe || aa1)
	{
		if (field(alpha, 0, 2) == 1 || field(alpha, 2, 2) == 1 || field(alpha, 4, 2) == 1 || field(alpha, 6, 2) == 1)
			needs_prev = true;
	}
	if (date)
		needs_prev = true;
	else if (ate && atst != 1)
	{
		bool can_write_z = (state & S_Z_WRITE) != 0;
		bool can_mask_color = afail != 1;
		bool can_mask_z = can_write_z && afail != 2;
		if (can_mask_color || can_mask_z)
			needs_prev = true;
	}
	if (fbk == 1)
		needs_prev = true; // reading the pixel's own value blends with it
	if (!needs_prev)
		state |= S_OPAQUE;
	if (aa1)
		state |= S_MULTISAMPLE;
	uint32_t msk = field(regs.scanmsk, 0, 2);
	if (msk >= 2)
	{
		state |= 1u << (23 + msk - 2);
		// gs/gs_interface.cpp:1947-1960.
		if (p.bb[1] + 400 < p.bb[3])
			complex_scanmsk_timeout = 4;
	}
	if (!field(prim, 8, 1))
		state |= S_PERSPECTIVE;
	if (field(prim, 3, 1))
		state |= S_IIP;
	if (field(prim, 10, 1))
		state |= S_FIX;

	p.fbmsk = uint32_t(frame >> 32);
	p.fogcol = uint32_t(regs.fogcol);
	p.alpha = (field(alpha, 32, 8) << 8) | aref;

	if (quad)
	{
		state |= S_PARALLELOGRAM | S_SPRITE | S_SNAP_RASTER | S_SNAP_ATTRIBUTE;
		state &= ~S_MULTISAMPLE;
	}
	else if (is_line)
	{
		state |= S_PARALLELOGRAM | S_LINE;
		if (state & S_MULTISAMPLE)
			state &= ~S_Z_WRITE;
	}
	if (n == 1)
		state |= S_FIX | S_SNAP_RASTER | S_SNAP_ATTRIBUTE;
	p.state = state;

	bool expands = rp_prims.empty();
	if (rp_prims.empty())
	{
		memcpy(rp_bb, p.bb, sizeof(rp_bb));
		rp_zwrite = false;
	}
	expands = expands || p.bb[0] < rp_bb[0] || p.bb[1] < rp_bb[1] || p.bb[2] > rp_bb[2] || p.bb[3] > rp_bb[3];
	rp_bb[0] = std::min(rp_bb[0], p.bb[0]);
	rp_bb[1] = std::min(rp_bb[1], p.bb[1]);
	rp_bb[2] = std::max(rp_bb[2], p.bb[2]);
	rp_bb[3] = std::max(rp_bb[3], p.bb[3]);
	rp_zwrite |= (state & S_Z_WRITE) != 0;
	rp_prims.push_back(p);
	// A draw growing the pass's box marks the box's frame-buffer pages, and its depth pages once
	// the pass writes depth, as possibly super-sampled (gs/gs_interface.cpp:2786-2816).
	if (expands && rp_bb[2] >= rp_bb[0] && rp_bb[3] >= rp_bb[1])
	{
		const uint32_t x0 = uint32_t(rp_bb[0]), y0 = uint32_t(rp_bb[1]), x1 = uint32_t(rp_bb[2]), y1 = uint32_t(rp_bb[3]);
		mark_pages(field(rp_frame, 0, 9) * 32u, field(rp_frame, 16, 6), field(rp_frame, 24, 6), x0, y0, x1, y1, true);
		if (rp_zwrite)
			mark_pages(field(rp_zbuf, 0, 9) * 32u, field(rp_frame, 16, 6), 0x30u | field(rp_zbuf, 24, 4), x0, y0, x1,
			           y1, true);
	}
	last_parallelogram_candidate = candidate;
	state_dirty = false;
	if (rp_prims.size() >= 64u * 1024u)
		flush_render_pass();
}

void GS::flush()
{
	partial_flush_transfer();
	flush_render_pass();
}

void GS::flush_render_pass()
{
	if (!rp_prims.empty())
	{
		render(rp_prims);
		rp_prims.clear();
		rp_textures.clear();
		rp_arrays = false;
		render_passes++;
	}
	state_dirty = true;
	last_parallelogram_candidate = false;
}

uint64_t GS::stat(uint32_t which) const
{
	switch (which)
	{
	case 0: return covered_samples;
	case 1: return primitives_drawn;
	case 2: return render_passes;
	default: return 0;
	}
}

// ---------------------------------------------------------------------------------------
// Triangle setup (gs/shaders/triangle_setup.comp:127-873).

namespace
{
bool tie_break(int32_t ex, int32_t ey) { return ey > 0 || (ey == 0 && ex < 0); }
bool order_less(int32_t ax, int32_t ay, int32_t bx, int32_t by, uint32_t az, uint32_t bz)
{
	if (bz != az)
		return az < bz;
	if (by != ay)
		return ay < by;
	return ax < bx;
}
int32_t quantize64(int64_t c) { return int32_t(uint32_t(uint64_t(c) >> 1)); }

// Returns false for a degenerate primitive (no coverage).
bool setup_primitive(Prim &p, Setup &s, uint32_t rate_log2, uint32_t opaque_fbmask, bool arrays = false,
                     uint32_t rate_x = 0, bool blur = false)
{
	const uint32_t st = p.state;
	const bool parallelogram = st & S_PARALLELOGRAM, perspective = st & S_PERSPECTIVE, iip = st & S_IIP;
	const bool fix = st & S_FIX, sprite = st & S_SPRITE, line = st & S_LINE, multisample = st & S_MULTISAMPLE;
	const uint32_t provoking = (st >> S_PROVOKING_SHIFT) & 3u;

	int32_t order[3] = { 0, 1, 2 };
	int32_t ax = p.v[0].x, ay = p.v[0].y, bx = p.v[1].x, by = p.v[1].y, cx = p.v[2].x, cy = p.v[2].y;
	uint32_t za = p.v[0].z, zb = p.v[1].z, zc = p.v[2].z;

	if (line)
	{
		bool sax = (ax & 15) == 0, say = (ay & 15) == 0, sbx = (bx & 15) == 0, sby = (by & 15) == 0;
		if ((sax || sbx) && bx > ax) { ax++; bx++; }
		if ((say || sby) && by > ay) { ay++; by++; }
	}

	if ((!parallelogram || line) && !fix)
	{
		if (!line && !order_less(bx, by, cx, cy, zb, zc))
		{
			std::swap(order[1], order[2]);
			std::swap(bx, cx); std::swap(by, cy); std::swap(zb, zc);
		}
		if (!order_less(ax, ay, bx, by, za, zb))
		{
			std::swap(order[0], order[1]);
			std::swap(ax, bx); std::swap(ay, by); std::swap(za, zb);
		}
	}
	else if (provoking == 1)
	{
		std::swap(order[0], order[1]);
		std::swap(ax, bx); std::swap(ay, by); std::swap(za, zb);
	}
	else if (provoking == 2)
	{
		std::swap(order[0], order[2]);
		std::swap(ax, cx); std::swap(ay, cy); std::swap(za, zc);
	}

	if (sprite)
	{
		cx = ax; cy = by;
		by = ay;
	}
	else if (line)
	{
		int32_t dx = bx - ax, dy = by - ay;
		if (std::abs(dx) >= std::abs(dy))
		{
			int32_t yd = -((dx > 0) - (dx < 0)) * 8;
			if (multisample) yd *= 2;
			cy = ay + yd; ay -= yd; by -= yd;
			if (!multisample) { ay++; by++; }
			cx = ax;
		}
		else
		{
			int32_t xd = ((dy > 0) - (dy < 0)) * 8;
			if (multisample) xd *= 2;
			cx = ax + xd; ax -= xd; bx -= xd;
			if (!multisample) { ax++; bx++; }
			cy = ay;
		}
	}

	int32_t abx = bx - ax, aby = by - ay, acx = cx - ax, acy = cy - ay;
	int64_t sa = int64_t(aby) * acx - int64_t(abx) * acy;
	if (sa == 0)
	{
		s.bb[0] = 0; s.bb[1] = 0; s.bb[2] = -1; s.bb[3] = -1;
		if (rate_log2)
			p.state |= S_SNAP_ATTRIBUTE;
		return false;
	}

	uint32_t area;
	if (sa < 0)
	{
		area = 0u - uint32_t(uint64_t(sa));
		std::swap(abx, acx); std::swap(aby, acy);
		std::swap(bx, cx); std::swap(by, cy); std::swap(zb, zc);
		std::swap(order[1], order[2]);
	}
	else
		area = uint32_t(uint64_t(sa));

	int32_t bcx = cx - bx, bcy = cy - by, cax = -acx, cay = -acy;
	int32_t dcdx = aby, dcdy = -abx, dadx = bcy, dady = -bcx, dbdx = cay, dbdy = -cax;

	int64_t kc = int64_t(-ax) * dcdx - int64_t(ay) * dcdy;
	int64_t ka = int64_t(-bx) * dadx - int64_t(by) * dady;
	int64_t kb = int64_t(-cx) * dbdx - int64_t(cy) * dbdy;
	bool ta = tie_break(bcx, bcy), tb = tie_break(cax, cay), tc = tie_break(abx, aby);
	int32_t ea = !ta, eb = !tb, ec = !tc;
	ka -= ea; kb -= eb; kc -= ec;
	ea += int32_t(uint32_t(uint64_t(ka)) & 1u);
	eb += int32_t(uint32_t(uint64_t(kb)) & 1u);
	ec += int32_t(uint32_t(uint64_t(kc)) & 1u);

	int32_t flip_b = 0, flip_c = 0;
	if (parallelogram)
	{
		int32_t off = 1 - int32_t(tb) - int32_t(tc);
		flip_b = (off + eb + ea + 2) >> 1;
		flip_c = (off + ec + ea + 2) >> 1;
	}

	float inv_area = rcp_fixed(area);
	float inv_area_scaled = inv_area * 2.0f;
	float error_i = float(eb) * inv_area;
	float error_j = float(ec) * inv_area;

	if (sprite) { zb = za; zc = za; }
	else if (line) zc = za;
	if (fix || sprite) { order[0] = 0; order[1] = 1; order[2] = 2; }

	const Vertex &a0 = p.v[order[0]], &a1 = p.v[order[1]];
	const Vertex &a2 = p.v[(!sprite && !line) ? order[2] : 0];
	float q0[4], q1[4], q2[4];
	auto load = [&](const Vertex &v, float *o) {
		if (perspective) { o[0] = v.s; o[1] = v.t; }
		else { o[0] = float(v.u) / 16.0f; o[1] = float(v.v) / 16.0f; }
		o[2] = v.q; o[3] = v.fog;
	};
	load(a0, q0);
	load(a1, q1);
	if (!sprite && !line)
		load(a2, q2);
	if (sprite)
	{
		if (sa < 0)
		{
			q2[0] = q1[0]; q2[1] = q0[1]; q2[2] = q0[2]; q2[3] = q0[3];
			q1[0] = q0[0]; q1[2] = q0[2]; q1[3] = q0[3];
		}
		else
		{
			q2[0] = q0[0]; q2[1] = q1[1]; q2[2] = q0[2]; q2[3] = q0[3];
			q1[1] = q0[1]; q1[2] = q0[2]; q1[3] = q0[3];
		}
	}
	else if (line)
		memcpy(q2, q0, sizeof(q0));

	if (perspective)
	{
		if (q0[2] < 0.0f)
			for (float *q : { q0, q1, q2 }) { q[0] = -q[0]; q[1] = -q[1]; }
		for (float *q : { q0, q1, q2 }) q[2] = std::fabs(q[2]);
		uint32_t qexp = (fbits(std::max(std::max(q0[2], q1[2]), q2[2])) >> 23) & 0xffu;
		for (float *q : { q0, q1, q2 }) q[2] = bitsf(fbits(q[2]) & ~0xffu);
		uint32_t sexp = (fbits(std::max(std::max(std::fabs(q0[0]), std::fabs(q1[0])), std::fabs(q2[0]))) >> 23) & 0xffu;
		uint32_t texp = (fbits(std::max(std::max(std::fabs(q0[1]), std::fabs(q1[1])), std::fabs(q2[1]))) >> 23) & 0xffu;
		uint32_t schop = std::min(9u + std::max(sexp, qexp) - sexp, 23u);
		uint32_t tchop = std::min(9u + std::max(texp, qexp) - texp, 23u);
		uint32_t smask = ~((1u << schop) - 1u), tmask = ~((1u << tchop) - 1u);
		for (float *q : { q0, q1, q2 })
		{
			q[0] = bitsf(fbits(q[0]) & smask);
			q[1] = bitsf(fbits(q[1]) & tmask);
		}
		bool needs_lod = (p.tex2 & 1u) == 0;
		if (!needs_lod && q0[2] == q1[2] && q0[2] == q2[2])
		{
			float inv_q = rcp_float(q0[2]);
			for (float *q : { q0, q1, q2 })
			{
				q[0] *= inv_q;
				q[1] *= inv_q;
				q[2] = 1.0f;
			}
		}
	}

	for (int k = 0; k < 4; k++)
	{
		s.stqf0[k] = q0[k];
		s.stqf1[k] = fix ? 0.0f : q1[k] - q0[k];
		s.stqf2[k] = fix ? 0.0f : q2[k] - q0[k];
	}

	if (iip && !sprite)
	{
		uint32_t r0 = a0.rgba, r1 = a1.rgba, r2 = a2.rgba;
		if (fix) r1 = r0;
		if (fix || line) r2 = r0;
		s.rgba[0] = r0; s.rgba[1] = r1; s.rgba[2] = r2;
	}
	else
		s.rgba[0] = s.rgba[1] = s.rgba[2] = p.v[0].rgba;

	s.a[0] = int32_t((uint32_t(dadx) << 6) | uint32_t(parallelogram) | (uint32_t(flip_b & 3) << 2) |
	                 (uint32_t(flip_c & 3) << 4));
	s.a[1] = dady; s.a[2] = quantize64(ka);
	s.b[0] = dbdx; s.b[1] = dbdy; s.b[2] = quantize64(kb);
	s.c[0] = dcdx; s.c[1] = dcdy; s.c[2] = quantize64(kc);

	int32_t dx = parallelogram ? bx + cx - ax : cx;
	int32_t dy = parallelogram ? by + cy - ay : cy;

	int32_t di = 0, dj = 0;
	if (sprite)
	{
		int32_t xd = dx - ax, yd = dy - ay;
		bool right = xd > 0 && (xd & (xd - 1)) != 0;
		bool bottom = yd > 0 && (yd & (yd - 1)) != 0;
		if (right) { di += dbdx; dj += dcdx; }
		if (bottom) { di += dbdy; dj += dcdy; }
	}
	else if (parallelogram && !perspective)
	{
		bool nearest = (p.tex & T_MAG_LINEAR) == 0;
		bool npot = (area & (area - 1u)) != 0;
		if (nearest && npot)
		{
			int32_t u0 = int32_t(a0.u) - ax, v0 = int32_t(a0.v) - ay;
			int32_t u1 = int32_t(a1.u) - bx, v1 = int32_t(a1.v) - by;
			int32_t u2 = int32_t(a2.u) - cx, v2 = int32_t(a2.v) - cy;
			if (u0 == u1 && v0 == v1 && u0 == u2 && v0 == v2)
			{
				bool mis_x = (std::min(ax, dx) & 15) != 0, mis_y = (std::min(ay, dy) & 15) != 0;
				float dudx = s.stqf1[0] * float(dbdx) + s.stqf2[0] * float(dcdx);
				float dvdx = s.stqf1[1] * float(dbdx) + s.stqf2[1] * float(dcdx);
				float dudy = s.stqf1[0] * float(dbdy) + s.stqf2[0] * float(dcdy);
				float dvdy = s.stqf1[1] * float(dbdy) + s.stqf2[1] * float(dcdy);
				bool adj_x = mis_x && ((dudx > 0.0f && dvdx == 0.0f) || (dvdx > 0.0f && dudx == 0.0f));
				bool adj_y = mis_y && ((dudy > 0.0f && dvdy == 0.0f) || (dvdy > 0.0f && dudy == 0.0f));
				if (adj_x) { di -= dbdx; dj -= dcdx; }
				if (adj_y) { di -= dbdy; dj -= dcdy; }
			}
		}
	}
	float k8 = 0.125f * inv_area_scaled;
	error_i = error_i + float(di) * k8;
	error_j = error_j + float(dj) * k8;

	s.inv_area = inv_area_scaled;
	s.error_i = error_i;
	s.error_j = error_j;
	s.z[0] = za;
	s.z[1] = fbits(float(zb - za));
	s.z[2] = fbits(float(zc - za));

	// pack_bounding_box (intersect.h:83-104), clipped to the scissor box.
	int32_t lox = std::min(std::min(ax, bx), std::min(cx, dx)), loy = std::min(std::min(ay, by), std::min(cy, dy));
	int32_t hix = std::max(std::max(ax, bx), std::max(cx, dx)), hiy = std::max(std::max(ay, by), std::max(cy, dy));
	lox = (lox + (15 >> rate_log2)) >> 4;
	loy = (loy + (15 >> rate_log2)) >> 4;
	hix = (hix - 1) >> 4;
	hiy = (hiy - 1) >> 4;
	if (multisample && !line) { lox--; loy--; hix++; hiy++; }
	s.bb[0] = std::max(lox, p.bb[0]);
	s.bb[1] = std::max(loy, p.bb[1]);
	s.bb[2] = std::min(hix, p.bb[2]);
	s.bb[3] = std::min(hiy, p.bb[3]);

	if (sprite && !perspective)
	{
		int32_t u0 = int32_t(a0.u) - ax, v0 = int32_t(a0.v) - ay;
		int32_t u1 = int32_t(a1.u) - dx, v1 = int32_t(a1.v) - dy;
		int32_t minu = std::min(u0, u1), minv = std::min(v0, v1), maxu = std::max(u0, u1), maxv = std::max(v0, v1);
		bool cs_u = (minu & 15) == 8, cs_v = (minv & 15) == 8;
		int32_t distx = std::abs(dx - ax), disty = std::abs(dy - ay);
		bool ns_u = (maxu & 15) == 9 && (distx & (distx - 1)) != 0;
		bool ns_v = (maxv & 15) == 9 && (disty & (disty - 1)) != 0;
		bool no_interp = (maxu - minu) <= 1 && (maxv - minv) <= 1;
		bool half_texel = (ns_u || cs_u) && (ns_v || cs_v);
		if (half_texel && no_interp)
			p.tex &= ~T_MAG_LINEAR;
	}

	// Per-sample texturing only where filtering does not break it (triangle_setup.comp:677-789,
	// FORCE_LINEAR_SUPER_SAMPLE off).
	if (rate_log2 != 0 && arrays && (p.tex & (T_PER_SAMPLE | T_SAMPLE_MAPPING)) == T_PER_SAMPLE)
	{
		const int32_t du = std::abs(int32_t(a1.u) - int32_t(a0.u)), dv = std::abs(int32_t(a1.v) - int32_t(a0.v));
		const int32_t distx = std::abs(dx - ax), disty = std::abs(dy - ay);
		if (!(p.tex & T_MAG_LINEAR))
		{
			if (perspective)
				p.tex |= T_SAMPLE_MAPPING;
			else if (du > distx + 1 || dv > disty + 1)
				p.tex &= ~T_PER_SAMPLE;
			else if (du >= distx - 16 || dv >= disty - 16)
			{
				const uint32_t bm = p.blend_mode;
				const bool blur_kernel = blur && (bm & B_ABE) && !(p.state & S_Z_TEST) && ((bm >> 11) & 3u) == 0u &&
				                         ((bm >> 13) & 3u) == 2u && ((bm >> 17) & 3u) == 1u;
				if (blur_kernel)
					p.tex &= ~T_PER_SAMPLE;
				else
					p.tex |= T_SAMPLE_MAPPING;
			}
		}
		else if (!perspective)
		{
			const bool down = du > distx + 1 && dv > disty + 1;
			const bool up = distx >= 2 * du && disty >= 2 * dv;
			if (down)
				p.tex &= ~T_PER_SAMPLE;
			else if (up)
				p.tex |= T_SAMPLE_RESOLVED;
		}
		if ((p.tex & (T_PER_SAMPLE | T_SAMPLE_MAPPING)) == T_PER_SAMPLE)
		{
			// get_average_sampling_offset (utils.h:108-125) times 2^3, on both axes.
			float avg;
			if (rate_x == rate_log2)
				avg = 0.5f * float((1 << rate_log2) - 1) / float(1 << rate_log2);
			else if (rate_log2 == 1)
				avg = 0.25f;
			else if (rate_x == 1 && rate_log2 == 2)
				avg = (0.125f + 0.625f) * 0.5f;
			else if (rate_log2 == 2)
				avg = 6.0f / 16.0f;
			else
				avg = (6.0f + 22.0f) / 64.0f;
			const float corner = 8.0f * avg;
			p.state &= ~S_SNAP_ATTRIBUTE;
			const float di_dc = float(s.b[0] + s.b[1]) * corner, dj_dc = float(s.c[0] + s.c[1]) * corner;
			s.error_i = s.error_i - di_dc * s.inv_area;
			s.error_j = s.error_j - dj_dc * s.inv_area;
		}
	}

	if (p.fbmsk != opaque_fbmask && (p.state & S_OPAQUE))
		p.state &= ~S_OPAQUE;

	if (rate_log2 != 0 && !sprite)
	{
		bool per_sample = multisample;
		if (!per_sample && perspective)
			per_sample = s.stqf1[2] != 0.0f || s.stqf2[2] != 0.0f;
		if (!per_sample)
			per_sample = za != zb || zb != zc;
		if (!per_sample && !line)
			p.state |= S_SNAP_ATTRIBUTE;
		if (!per_sample)
		{
			s.flat = true;
			s.flat_z = za;
		}
	}
	return true;
}
}

// ---------------------------------------------------------------------------------------
// Per-sample shading.

namespace
{
struct Regs
{
	uint32_t color = 0, depth = 0;
	bool loaded = false, cdirty = false, zdirty = false;
};

struct PassFormat
{
	uint32_t fb_psm, z_psm;   // z_psm = 0xffffffff when the pass does not touch depth
	uint32_t fbp, fbw, zbp;
};

// intersect.h:106-244: the coverage count of a sample, 0x80 where the sample centre is inside.
// With AA1 (multisample) a line's coverage falls off linearly across its two-pixel-wide
// parallelogram (128 - 256 |0.5 - j|), and a triangle sample that fails exactly one edge gets
// a feathered coverage from its distance to that edge, measured along the edge's major axis in
// 1/8-pixel units (never 0x80, so such samples never write depth).
int32_t evaluate_coverage(const Setup &s, int32_t cx, int32_t cy, uint32_t rate, float &i, float &j,
                          bool multisample = false, bool line = false)
{
	int32_t sx = cx >> rate, sy = cy >> rate;
	if (!(sx >= s.bb[0] && sy >= s.bb[1] && sx <= s.bb[2] && sy <= s.bb[3]))
		return 0;
	bool para = (s.a[0] & 1) != 0;
	int32_t offb = int32_t((uint32_t(s.a[0]) >> 2) & 3u) - 1;
	int32_t offc = int32_t((uint32_t(s.a[0]) >> 4) & 3u) - 1;
	int32_t a_eq[3] = { s.a[0] >> 6, s.a[1], s.a[2] };
	int32_t px = int32_t(uint32_t(cx) << (3 - rate)), py = int32_t(uint32_t(cy) << (3 - rate));
	int32_t A = idot3(a_eq, px, py), B = idot3(s.b, px, py), C = idot3(s.c, px, py);
	const int32_t A0 = A, B0 = B, C0 = C;
	float ti = float(B) * s.inv_area;
	float tj = float(C) * s.inv_area;
	i = ti + s.error_i;
	j = tj + s.error_j;
	if (para && A < 0)
	{
		B = wadd(B, wadd(A, offb));
		C = wadd(C, wadd(A, offc));
		A = 0;
	}
	const bool centre = A >= 0 && B >= 0 && C >= 0;
	if (multisample && line)
	{
		if (!centre)
			return 0;
		float d = std::fabs(0.5f - j);
		return int32_t(std::max(0.0f, 128.0f - 256.0f * d));
	}
	if (centre)
		return 0x80;
	if (!multisample || para)
		return 0;
	const bool ea = A0 >= 0, eb = B0 >= 0, ec = C0 >= 0;
	if (int(ea) + int(eb) + int(ec) != 2)
		return 0;
	float negative_bary;
	int32_t max_delta;
	if (!ea)
	{
		max_delta = std::max(std::abs(s.b[0] + s.c[0]), std::abs(s.b[1] + s.c[1]));
		negative_bary = (i + j) - 1.0f;
	}
	else if (!eb)
	{
		max_delta = std::max(std::abs(s.b[0]), std::abs(s.b[1]));
		negative_bary = -i;
	}
	else
	{
		max_delta = std::max(std::abs(s.c[0]), std::abs(s.c[1]));
		negative_bary = -j;
	}
	float scaled_delta = float(max_delta) * s.inv_area;
	float subpixels = negative_bary / scaled_delta;
	float coverage = 8.0f - subpixels;
	return int32_t(std::min(std::max(16.0f * coverage, 0.0f), 127.0f));
}

void barycentric(const Setup &s, int32_t cx, int32_t cy, uint32_t rate, float &i, float &j)
{
	int32_t px = int32_t(uint32_t(cx) << (3 - rate)), py = int32_t(uint32_t(cy) << (3 - rate));
	int32_t B = idot3(s.b, px, py), C = idot3(s.c, px, py);
	float ti = float(B) * s.inv_area;
	float tj = float(C) * s.inv_area;
	i = ti + s.error_i;
	j = tj + s.error_j;
}

bool alpha_test(uint32_t atst, uint32_t aref, float a)
{
	float r = float(aref);
	switch (atst)
	{
	case 0: return false;
	case 2: return a < r;
	case 3: return a <= r;
	case 4: return a == r;
	case 5: return a >= r;
	case 6: return a > r;
	case 7: return a != r;
	default: return true;
	}
}
// triangle_setup_intersects_tile (intersect.h:24-81): conservative test of a primitive against
// the pixel rectangle [lo, hi], clipped to the primitive's box first.
bool intersects_tile(const Setup &s, int32_t tlox, int32_t tloy, int32_t thix, int32_t thiy, bool multisampled_tri)
{
	int32_t lox = std::max(s.bb[0], tlox), loy = std::max(s.bb[1], tloy);
	int32_t hix = std::min(s.bb[2], thix), hiy = std::min(s.bb[3], thiy);
	if (hix < lox || hiy < loy)
		return false;
	const bool para = (s.a[0] & 1) != 0;
	const int32_t offb = int32_t((uint32_t(s.a[0]) >> 2) & 3u) - 1;
	const int32_t offc = int32_t((uint32_t(s.a[0]) >> 4) & 3u) - 1;
	const int32_t a_eq[3] = { s.a[0] >> 6, s.a[1], s.a[2] };
	lox <<= 3;
	loy <<= 3;
	hix = (hix << 3) | 7;
	hiy = (hiy << 3) | 7;
	auto pick = [&](int32_t e, int32_t lo, int32_t hi) { return e >= 0 ? hi : lo; };
	int32_t da = idot3(a_eq, pick(a_eq[0], lox, hix), pick(a_eq[1], loy, hiy));
	int32_t db = idot3(s.b, pick(s.b[0], lox, hix), pick(s.b[1], loy, hiy));
	int32_t dc = idot3(s.c, pick(s.c[0], lox, hix), pick(s.c[1], loy, hiy));
	if (multisampled_tri && !para)
	{
		da = wadd(da, 8 * std::max(std::abs(a_eq[0]), std::abs(a_eq[1])));
		db = wadd(db, 8 * std::max(std::abs(s.b[0]), std::abs(s.b[1])));
		dc = wadd(dc, 8 * std::max(std::abs(s.c[0]), std::abs(s.c[1])));
	}
	bool hit = da >= 0 && db >= 0 && dc >= 0;
	if (para && !hit)
	{
		const int32_t xa = pick(a_eq[0], lox, hix), ya = pick(a_eq[1], loy, hiy);
		int32_t pa = idot3(a_eq, xa, ya);
		int32_t pb = idot3(s.b, pick(-s.c[0], lox, hix), pick(-s.c[1], loy, hiy));
		int32_t pc = idot3(s.c, pick(-s.b[0], lox, hix), pick(-s.b[1], loy, hiy));
		pb = wadd(pb, wadd(pa, offb));
		pc = wadd(pc, wadd(pa, offc));
		const int32_t na = idot3(a_eq, a_eq[0] >= 0 ? lox : hix, a_eq[1] >= 0 ? loy : hiy);
		hit = -1 >= na && pb >= 0 && pc >= 0;
	}
	return hit;
}

// single_sample_heuristic.comp:148-290, run per active depth bucket. Every bucket sees the
// states as triangle setup left them; only the final step writes states, and only for
// primitives of its own bucket (the reference runs buckets as concurrent workgroups).
uint32_t hash_vertex(const Vertex &v)
{
	const uint32_t P = 1103515245u;
	uint32_t x = uint32_t(v.x), y = uint32_t(v.y), z = v.z;
	for (int r = 0; r < 3; r++)
	{
		const uint32_t nx = ((x >> 8) ^ y) * P, ny = ((y >> 8) ^ z) * P, nz = ((z >> 8) ^ x) * P;
		x = nx; y = ny; z = nz;
	}
	return x >> 16;
}

void single_sample_heuristic(std::vector<Prim> &prims, const std::vector<Setup> &setups, uint32_t depth_psm)
{
	uint32_t z_shift_setup, z_shift_bucket, z_shift, z_max;
	switch (depth_psm)
	{
	case Z16: case Z16S: z_shift_setup = 8; z_shift_bucket = 8; z_shift = 0; z_max = 0xffffu; break;
	case Z32: z_shift_setup = 24; z_shift_bucket = 16; z_shift = 8; z_max = 0xffffffffu; break;
	default: z_shift_setup = 16; z_shift_bucket = 16; z_shift = 0; z_max = 0xffffffu; break;
	}
	std::vector<bool> active(256, false);
	bool any = false;
	for (size_t i = 0; i < prims.size(); i++)
		if (setups[i].flat)
		{
			active[std::min(255u, setups[i].flat_z >> z_shift_setup)] = true;
			any = true;
		}
	if (!any)
		return;

	const uint32_t NO_FULL_3D = S_SPRITE | S_LINE | S_SNAP_ATTRIBUTE;
	std::vector<uint32_t> pre(prims.size());
	bool has_full_3d = false;
	uint32_t min3d = ~0u, max3d = 0;
	for (size_t i = 0; i < prims.size(); i++)
	{
		pre[i] = prims[i].state;
		if ((pre[i] & NO_FULL_3D) == 0)
		{
			has_full_3d = true;
			for (int k = 0; k < 3; k++)
			{
				min3d = std::min(min3d, prims[i].v[k].z);
				max3d = std::max(max3d, prims[i].v[k].z);
			}
		}
	}
	const bool force_single = !has_full_3d;
	auto snapped = [&](uint32_t z) { return std::min(z_max, z) >> z_shift; };

	std::vector<uint8_t> full3d(65536), cand(65536), shared(65536);
	for (uint32_t b = 0; b < 256; b++)
	{
		if (!active[b])
			continue;
		const uint32_t zlo = b << z_shift_bucket, zhi = ((b + 1u) << z_shift_bucket) - 1u;
		auto in_range = [&](uint32_t sz) { return sz >= zlo && sz <= zhi; };
		std::fill(full3d.begin(), full3d.end(), 0);
		std::fill(cand.begin(), cand.end(), 0);
		std::fill(shared.begin(), shared.end(), 0);
		for (size_t i = 0; i < prims.size(); i++)
		{
			const uint32_t st = pre[i];
			if ((st & NO_FULL_3D) == 0)
			{
				for (int k = 0; k < 3; k++)
				{
					const uint32_t sz = snapped(prims[i].v[k].z);
					if (in_range(sz))
						full3d[sz - zlo] = 1;
				}
			}
			else if ((st & (S_SNAP_ATTRIBUTE | S_SNAP_RASTER)) == S_SNAP_ATTRIBUTE || (st & S_LINE))
			{
				const uint32_t sz = snapped(prims[i].v[0].z);
				if (in_range(sz))
					cand[sz - zlo] = 1;
			}
		}
		for (size_t i = 0; i < prims.size(); i++)
		{
			if ((pre[i] & NO_FULL_3D) != 0)
				continue;
			for (int k = 0; k < 3; k++)
			{
				const uint32_t sz = snapped(prims[i].v[k].z);
				if (in_range(sz) && cand[sz - zlo])
					shared[hash_vertex(prims[i].v[k])] = 1;
			}
		}
		for (size_t i = 0; i < prims.size(); i++)
		{
			if ((pre[i] & (S_SNAP_ATTRIBUTE | S_SNAP_RASTER)) != S_SNAP_ATTRIBUTE)
				continue;
			const uint32_t sz = snapped(prims[i].v[0].z);
			if (!in_range(sz) || !full3d[sz - zlo])
				continue;
			if (shared[hash_vertex(prims[i].v[0])] || shared[hash_vertex(prims[i].v[1])] ||
			    shared[hash_vertex(prims[i].v[2])])
				cand[sz - zlo] = 0;
		}
		for (size_t i = 0; i < prims.size(); i++)
		{
			const uint32_t z = prims[i].v[0].z;
			const uint32_t sz = snapped(z);
			if (!in_range(sz))
				continue;
			uint32_t st = pre[i];
			const bool tri_cand = (st & (S_SNAP_ATTRIBUTE | S_SNAP_RASTER)) == S_SNAP_ATTRIBUTE;
			const bool line_cand = force_single && (st & S_LINE);
			if (!tri_cand && !line_cand)
				continue;
			if (cand[sz - zlo])
				prims[i].state = st | S_SNAP_RASTER;
			else if ((st & (S_Z_TEST | S_Z_WRITE)) && z > min3d && z < max3d)
				prims[i].state = st & ~S_SNAP_ATTRIBUTE;
		}
	}
}

// Hi-res sample offsets inside a pixel for sample index s (ubershader.comp:133-385: lane bits
// of the shading wave map to sample positions; sample index = lane & (samples - 1)).
void sample_offset(uint32_t rate_x, uint32_t rate_y, uint32_t s, int32_t &dx, int32_t &dy)
{
	if (rate_x == 0 && rate_y == 1) { dx = int32_t(s); dy = int32_t(s); }
	else if (rate_x == 1 && rate_y == 1) { dx = int32_t(s >> 1); dy = int32_t(s & 1u); }
	else if (rate_x == 1 && rate_y == 2)
	{
		dx = int32_t(2u * ((s >> 1) & 1u) + (s & 1u));
		dy = int32_t((s & 1u) + 2u * ((s >> 2) & 1u));
	}
	else if (rate_x == 2 && rate_y == 2)
	{
		dx = int32_t(((s >> 1) & 1u) + 2u * ((s >> 3) & 1u));
		dy = int32_t((s & 1u) + 2u * ((s >> 2) & 1u));
	}
	else if (rate_x == 0 && rate_y == 2)
	{
		// Sparse 1 x 4: one sample per quarter row at x quarters 0, 2, 3, 1 (the 0x78 offsets of
		// ubershader.comp:346-351).
		dx = int32_t((0x78u >> (2u * s)) & 3u);
		dy = int32_t(s);
	}
	else if (rate_x == 1 && rate_y == 3)
	{
		// Sparse 2 x 8: per eighth row two samples half a pixel apart, shifted by the same
		// four-row pattern.
		dx = int32_t(4u * ((s >> 3) & 1u) + ((0x78u >> (2u * (s & 3u))) & 3u));
		dy = int32_t(s & 7u);
	}
	else { dx = 0; dy = 0; }
}
// A render pass of a super-sampled context (planes: 0 = single-sampled VRAM, 1 = reference
// marking pixels whose super-samples are valid, 2 + s = sample s).
//
// Coarse tiles start at the pass bounding box (gs/gs_interface.cpp:148-187, 214-215); each tile
// is shaded super-sampled or single-sampled as binning decides (binning.comp:87-125, 272-312).
// begin_tile / end_tile follow ubershader.comp:435-780: samples load from their planes where
// the reference marks the pixel, and a dirty pixel whose samples differ stores every sample,
// the reference mask and the resolved value (float average in butterfly order, minimum depth).
//
// Only primitives lo..hi shade (the tile lists and tile modes still come from the whole pass);
// colour is read from csrc and depth from zsrc (copies of the planes, or v0 itself); bit 0 of
// skip leaves colour unstored, bit 1 depth.
template <typename Shade>
void render_super_sampled_pass(uint8_t *v0, const uint8_t *csrc, const uint8_t *zsrc, uint32_t lo, uint32_t hi,
                               uint32_t skip, const int32_t *bb, uint32_t context_rate_y, std::vector<Prim> &prims,
                               const std::vector<Setup> &setups, const std::vector<bool> &live, const PassFormat &f,
                               uint32_t rate_x, uint32_t rate_y, Shade &&shade)
{
	const uint32_t n = uint32_t(prims.size());
	const bool ssaa = rate_y != 0;
	const uint32_t S = 1u << (rate_x + rate_y);
	const bool zs = f.z_psm != 0xffffffffu;
	const bool fb16 = psm16(f.fb_psm), fb24 = psm24(f.fb_psm);
	const bool z16 = zs && psm16(f.z_psm), z24 = zs && psm24(f.z_psm);
	const uint32_t fb_base = f.fbp * 32u, z_base = f.zbp * 32u;
	const size_t P = size_t(4) << 20;

	// Coarse tile size (gs/gs_interface.cpp:148-187), in the reference's 32-bit arithmetic.
	const uint32_t w8 = uint32_t((bb[2] - bb[0]) >> 3) + 1u, h8 = uint32_t((bb[3] - bb[1]) >> 3) + 1u;
	const uint32_t cost = w8 * h8 * n;
	uint32_t tl = cost < 10000u ? 3u : cost < 10000000u ? 4u : cost < 100000000u ? 5u : 6u;
	if (context_rate_y != 0 && tl > 3)
		tl--;
	{
		const uint32_t cw = uint32_t((bb[2] - bb[0]) >> tl) + 1u, ch = uint32_t((bb[3] - bb[1]) >> tl) + 1u;
		uint64_t list_bytes = uint64_t(uint32_t(cw * ch * n)) * 2u;
		while (list_bytes > 48ull * 1024 * 1024)
		{
			tl++;
			list_bytes /= 4;
		}
	}
	const uint32_t cw = uint32_t((bb[2] - bb[0]) >> tl) + 1u, ch = uint32_t((bb[3] - bb[1]) >> tl) + 1u;
	const int32_t tsize = 1 << tl;

	auto rd32 = [&](size_t a) { uint32_t d; memcpy(&d, csrc + a, 4); return d; };
	auto rd16 = [&](size_t a) { uint16_t d; memcpy(&d, csrc + a, 2); return uint32_t(d); };
	auto zrd32 = [&](size_t a) { uint32_t d; memcpy(&d, zsrc + a, 4); return d; };
	auto zrd16 = [&](size_t a) { uint16_t d; memcpy(&d, zsrc + a, 2); return uint32_t(d); };
	auto wr32 = [&](size_t a, uint32_t d) { memcpy(v0 + a, &d, 4); };
	auto wr16 = [&](size_t a, uint32_t d) { uint16_t h = uint16_t(d); memcpy(v0 + a, &h, 2); };

	std::vector<uint32_t> list;
	Regs regs[16];
	for (uint32_t ty = 0; ty < ch; ty++)
	{
		for (uint32_t tx = 0; tx < cw; tx++)
		{
			const int32_t lox = bb[0] + int32_t(tx) * tsize, loy = bb[1] + int32_t(ty) * tsize;
			const int32_t hix = lox + tsize - 1, hiy = loy + tsize - 1;
			list.clear();
			bool req = false, req_zw = false, has_zw = false;
			for (uint32_t i = 0; i < n; i++)
			{
				if (!live[i])
					continue;
				const uint32_t st = prims[i].state;
				if (!intersects_tile(setups[i], lox, loy, hix, hiy, (st & S_MULTISAMPLE) && !(st & S_LINE)))
					continue;
				list.push_back(i);
				if ((st & (S_SNAP_RASTER | S_Z_TEST | S_OPAQUE)) != (S_SNAP_RASTER | S_OPAQUE))
					req_zw = true;
				if ((st & (S_SNAP_RASTER | S_OPAQUE)) != (S_SNAP_RASTER | S_OPAQUE))
					req = true;
				if (prims[i].tex & T_PER_SAMPLE) // binning.comp:123-124
					req = true;
				if (st & S_Z_WRITE)
					has_zw = true;
			}
			if (list.empty())
				continue;
			const bool super = ssaa && (req || (req_zw && has_zw));
			const uint32_t ns = super ? S : 1u;
			const uint32_t rate = super ? rate_y : 0u;

			for (int32_t py = loy; py <= hiy; py++)
			{
				for (int32_t px = lox; px <= hix; px++)
				{
					bool any = false;
					for (uint32_t i : list)
					{
						const Setup &s = setups[i];
						if (px >= s.bb[0] && px <= s.bb[2] && py >= s.bb[1] && py <= s.bb[3])
						{
							any = true;
							break;
						}
					}
					if (!any)
						continue;

					// begin_tile
					const size_t ca = fb16 ? size_t(swizzle(uint32_t(px), uint32_t(py), fb_base, f.fbw, f.fb_psm)) * 2u
					                       : size_t(swizzle(uint32_t(px), uint32_t(py), fb_base, f.fbw, f.fb_psm)) * 4u;
					size_t za = 0;
					if (zs)
						za = z16 ? size_t(swizzle(uint32_t(px), uint32_t(py), z_base, f.fbw, f.z_psm)) * 2u
						         : size_t(swizzle(uint32_t(px), uint32_t(py), z_base, f.fbw, f.z_psm)) * 4u;
					for (uint32_t smp = 0; smp < ns; smp++)
					{
						Regs &r = regs[smp];
						r = Regs{};
						r.loaded = true;
						if (fb16)
						{
							uint32_t c = rd16(ca);
							if (super && rd16(ca + P) == 0xffffu)
								c = rd16(ca + (2u + smp) * P);
							r.color = rgba16_to_rgba32(uint16_t(c), 0, 0, 0x80);
						}
						else
						{
							r.color = rd32(ca);
							if (super)
							{
								const uint32_t ref = rd32(ca + P), mask = fb24 ? 0xffffffu : ~0u;
								if ((ref & mask) == mask)
								{
									const uint32_t sc = rd32(ca + (2u + smp) * P);
									r.color = (fb24 && ref != ~0u) ? ((r.color & 0xff000000u) | (sc & 0xffffffu)) : sc;
								}
							}
						}
						if (zs)
						{
							if (z16)
							{
								r.depth = zrd16(za);
								if (super && zrd16(za + P) == 0xffffu)
									r.depth = zrd16(za + (2u + smp) * P);
							}
							else
							{
								const uint32_t mask = z24 ? 0xffffffu : ~0u;
								r.depth = zrd32(za) & mask;
								if (super && (zrd32(za + P) & mask) == mask)
									r.depth = zrd32(za + (2u + smp) * P) & mask;
							}
						}
					}

					for (uint32_t i : list)
					{
						const Setup &s = setups[i];
						if (i < lo || i > hi || !(px >= s.bb[0] && px <= s.bb[2] && py >= s.bb[1] && py <= s.bb[3]))
							continue;
						for (uint32_t smp = 0; smp < ns; smp++)
						{
							int32_t dx = 0, dy = 0;
							if (super)
								sample_offset(rate_x, rate_y, smp, dx, dy);
							const int32_t hx = (px << rate) + dx, hy = (py << rate) + dy;
							shade(prims[i], s, hx, hy, rate, super, smp, [&]() -> Regs & { return regs[smp]; });
						}
					}

					// end_tile
					bool req_c = false, req_d = false, cdirty = false, zdirty = false;
					for (uint32_t smp = 0; smp < ns; smp++)
					{
						req_c |= regs[smp].color != regs[0].color;
						req_d |= zs && regs[smp].depth != regs[0].depth;
						cdirty |= regs[smp].cdirty;
						zdirty |= zs && regs[smp].zdirty;
					}
					if (skip & 1u)
						cdirty = false;
					if (skip & 2u)
						zdirty = false;
					uint32_t color = regs[0].color, depth = regs[0].depth;
					if (req_c)
					{
						if (cdirty)
							for (uint32_t smp = 0; smp < ns; smp++)
							{
								if (fb16)
									wr16(ca + (2u + smp) * P, rgba32_to_rgba16(regs[smp].color));
								else
									wr32(ca + (2u + smp) * P, regs[smp].color);
							}
						float acc[16][4];
						for (uint32_t smp = 0; smp < ns; smp++)
							for (uint32_t k = 0; k < 4; k++)
								acc[smp][k] = unorm(regs[smp].color, k);
						for (uint32_t m = ns >> 1; m >= 1; m >>= 1)
						{
							float nx[16][4];
							for (uint32_t smp = 0; smp < ns; smp++)
								for (uint32_t k = 0; k < 4; k++)
									nx[smp][k] = acc[smp][k] + acc[smp ^ m][k];
							memcpy(acc, nx, sizeof(acc));
						}
						float avg[4];
						for (uint32_t k = 0; k < 4; k++)
							avg[k] = acc[0][k] / float(ns);
						color = pack_unorm(avg);
					}
					if (req_d)
					{
						if (zdirty)
							for (uint32_t smp = 0; smp < ns; smp++)
							{
								if (z16)
									wr16(za + (2u + smp) * P, regs[smp].depth);
								else
									wr32(za + (2u + smp) * P, regs[smp].depth);
							}
						for (uint32_t smp = 1; smp < ns; smp++)
							depth = std::min(depth, regs[smp].depth);
					}
					if (cdirty)
					{
						if (fb16)
						{
							wr16(ca + P, req_c ? 0xffffu : 0u);
							wr16(ca, rgba32_to_rgba16(color));
						}
						else
						{
							wr32(ca + P, req_c ? ~0u : 0u);
							wr32(ca, color);
						}
					}
					if (zdirty)
					{
						if (z16)
						{
							wr16(za + P, req_d ? 0xffffu : 0u);
							wr16(za, depth);
						}
						else
						{
							wr32(za + P, req_d ? (z24 ? 0xffffffu : ~0u) : 0u);
							if (z24)
								memcpy(v0 + za, &depth, 3);
							else
								wr32(za, depth);
						}
					}
				}
			}
		}
	}
}
}

// Texture sampling lives in oracle_texture.cpp.
void sample_texture(const TextureSlot &t, uint32_t tex, uint32_t tex2, float u, float v, float q, bool perspective,
                    const uint32_t *lut, uint32_t smp, float shift, float out[4]);

void GS::render(const std::vector<Prim> &prims_in)
{
	std::vector<Prim> prims = prims_in;
	const uint32_t ctx_frame_psm = field(rp_frame, 24, 6);
	PassFormat f;
	f.fb_psm = ctx_frame_psm;
	f.fbp = field(rp_frame, 0, 9);
	f.fbw = field(rp_frame, 16, 6);
	f.zbp = field(rp_zbuf, 0, 9);
	uint32_t zpsm = field(rp_zbuf, 24, 4) | 0x30u;
	// deduce_effectice_z_psm (gs/gs_renderer.cpp:2147-2170)
	if (f.fb_psm == Z32 || f.fb_psm == Z24 || f.fb_psm == Z16S)
	{
		if (zpsm == Z32) zpsm = CT32;
		else if (zpsm == Z24) zpsm = CT24;
		else if (zpsm == Z16S || zpsm == Z16) zpsm = CT16S;
	}
	bool z_sensitive = false, z_write_any = false;
	uint32_t write_mask = 0;
	for (auto &p : prims)
	{
		z_sensitive |= (p.state & (S_Z_TEST | S_Z_WRITE)) != 0;
		z_write_any |= (p.state & S_Z_WRITE) != 0;
		write_mask |= ~p.fbmsk;
	}
	f.z_psm = z_sensitive ? zpsm : 0xffffffffu;
	const uint32_t opaque_fbmask = ~write_mask;

	// Sampling rate of the pass (gs/gs_interface.cpp:212-260): x/y log2 from the context's
	// sample count and grid (2x: 0/1, 4x: 1/1 or sparse 0/2, 8x: 1/2, 16x: 2/2 or sparse 1/3);
	// channel-sliced 16-bit colour masks
	// and colour/depth aliasing with depth writes render single-sampled.
	uint32_t rate_x = rate_x_log2(), rate_y = rate_y_log2();
	if (sample_log2)
	{
		const bool fb_32_24 = f.fb_psm == CT32 || f.fb_psm == CT24 || f.fb_psm == Z32 || f.fb_psm == Z24;
		const uint32_t m = write_mask & 0xf8f8f8u;
		const bool channel_slice = !fb_32_24 && m != 0xf8f8f8u && m != 0;
		const bool aliased = z_write_any && f.zbp == f.fbp;
		bool z_feedback = false;
		for (const Prim &q : prims)
			z_feedback |= q.feedback == 2;
		// A pass sampling super-sampled textures keeps its rate (gs/gs_interface.cpp:224-226).
		if (!rp_arrays && (channel_slice || z_feedback || aliased))
			rate_x = rate_y = 0;
	}

	// A deduced blur (gs/gs_renderer.cpp:1848-1885): at most 64 primitives, one super-sampled
	// texture read at two phases (first vertex UV less position), none perspective.
	bool blur = false;
	if (rp_arrays && prims.size() <= 64u)
	{
		uint32_t last = ~0u;
		int32_t ph_u = 0, ph_v = 0;
		for (const Prim &p : prims)
		{
			if (!(p.tex & T_PER_SAMPLE))
				continue;
			if (p.state & S_PERSPECTIVE)
			{
				blur = false;
				break;
			}
			const int32_t pu = int32_t(p.v[0].u) - p.v[0].x, pv = int32_t(p.v[0].v) - p.v[0].y;
			if (p.tex_slot == last && (pu != ph_u || pv != ph_v))
			{
				blur = true;
				break;
			}
			if (p.tex_slot != last)
			{
				ph_u = pu;
				ph_v = pv;
			}
			last = p.tex_slot;
		}
	}
	static std::vector<uint32_t> luts[12];
	const uint32_t *lut = nullptr;
	if (rp_arrays && rate_y)
	{
		std::vector<uint32_t> &l = luts[rate_y * 3 + rate_x];
		if (l.empty())
			phase_lut(rate_x, rate_y, l);
		lut = l.data();
	}

	std::vector<Setup> setups(prims.size());
	std::vector<bool> live(prims.size());
	for (size_t i = 0; i < prims.size(); i++)
		live[i] = setup_primitive(prims[i], setups[i], rate_y, opaque_fbmask, rp_arrays, rate_x, blur);
	if (rate_y != 0)
		single_sample_heuristic(prims, setups, z_sensitive ? (field(rp_zbuf, 24, 4) | 0x30u) : uint32_t(Z24));

	const int32_t snap_mask = -(1 << rate_y);

	// One primitive at one sample: hi-res coordinate (hx, hy) of pixel (hx >> rate, hy >> rate).
	// 0.5 less the shading grid's mean sample offset (ubershader.comp:1167; 0.5 at one sample).
	float phase_shift = 0.5f;
	if (rate_y)
	{
		if (rate_x == rate_y)
			phase_shift = 0.5f - 0.5f * float((1 << rate_y) - 1) / float(1 << rate_y);
		else if (rate_y == 1)
			phase_shift = 0.25f;
		else if (rate_y == 2)
			phase_shift = 0.5f - 6.0f / 16.0f;
		else
			phase_shift = 0.5f - 28.0f / 64.0f;
	}
	auto shade = [&](const Prim &p, const Setup &s, int32_t hx, int32_t hy, uint32_t rate, bool super_sampled,
	                 uint32_t smp, auto &&get_regs) {
		const uint32_t st = p.state;
		const bool z_test = f.z_psm != 0xffffffffu && (st & S_Z_TEST);
		bool z_write = f.z_psm != 0xffffffffu && (st & S_Z_WRITE);
		const bool z_greater = st & S_Z_GREATER;
		const bool persp = st & S_PERSPECTIVE;
		// shade_early (ubershader.comp:782-910): snapped primitives rasterize at the
		// pixel's first sample position.
		const bool snap_raster = super_sampled && (st & S_SNAP_RASTER);
		const int32_t rx = snap_raster ? (hx & snap_mask) : hx, ry = snap_raster ? (hy & snap_mask) : hy;
		float fi, fj;
		const bool multisample = (st & S_MULTISAMPLE) != 0;
		int32_t cov = evaluate_coverage(s, rx, ry, rate, fi, fj, multisample, (st & S_LINE) != 0);
		if (cov == 0)
			return;
		if (st & (S_SCANMSK_EVEN | S_SCANMSK_ODD))
		{
			const int32_t sy = (st & S_SPRITE) ? (hy >> rate) : hy;
			if ((sy & 1) == 0 ? (st & S_SCANMSK_EVEN) : (st & S_SCANMSK_ODD))
				return;
		}
		// AA1 without full coverage does not write depth (ubershader.comp:848-851).
		if (multisample && cov < 0x80)
			z_write = false;
		covered_samples++;
		Regs &r = get_regs();

		// Z (ubershader.comp:852-871): no FMA contraction of dzdi * i + dzdj * j.
		uint32_t z = s.z[0];
		float t0 = bitsf(s.z[1]) * fi;
		float t1 = bitsf(s.z[2]) * fj;
		float zf = std::nearbyint(t0 + t1);
		zf = std::min(std::max(zf, 0.0f), 4294967040.0f);
		uint32_t dz = uint32_t(zf);
		dz = std::min(dz, 0xffffffffu - z);
		z += dz;
		if (f.z_psm != 0xffffffffu)
		{
			if (psm16(f.z_psm)) z = std::min(z, 0xffffu);
			else if (psm24(f.z_psm)) z = std::min(z, 0xffffffu);
		}

		if (z_test && !(z_greater ? z > r.depth : z >= r.depth))
			return;
		// SNAP_ATTRIBUTE without SNAP_RASTER: attributes at the pixel's first sample
		// (classic MSAA); depth stays per sample.
		if (super_sampled && (st & (S_SNAP_RASTER | S_SNAP_ATTRIBUTE)) == S_SNAP_ATTRIBUTE)
			barycentric(s, hx & snap_mask, hy & snap_mask, rate, fi, fj);

		// Combined colour (ubershader.comp:1260-1339).
		float col[4];
		for (uint32_t ch = 0; ch < 4; ch++)
		{
			float c0 = unorm(s.rgba[0], ch), c1 = unorm(s.rgba[1], ch), c2 = unorm(s.rgba[2], ch);
			float d1 = c1 - c0, d2 = c2 - c0;
			float e1 = d1 * fi, e2 = d2 * fj;
			float sum = c0 + e1;
			sum = sum + e2;
			col[ch] = std::floor(255.0f * sum + 1.0f / 256.0f);
		}
		float comb[4] = { col[0], col[1], col[2], col[3] };
		if (p.combiner & C_TME)
		{
			float sq[3];
			for (int k = 0; k < 3; k++)
			{
				float e1 = s.stqf1[k] * fi, e2 = s.stqf2[k] * fj;
				float sum = s.stqf0[k] + e1;
				sq[k] = sum + e2;
			}
			float u = sq[0], vv = sq[1];
			if (persp && sq[2] != 1.0f)
			{
				u = u / sq[2];
				vv = vv / sq[2];
			}
			// NClamp: a NaN coordinate clamps to the lower bound.
			u = std::fmin(std::fmax(u, -2047.0f), 2047.0f);
			vv = std::fmin(std::fmax(vv, -2047.0f), 2047.0f);
			float tx[4];
			if (p.feedback)
				feedback_texture(p, p.feedback == 2 ? r.depth : r.color, tx);
			else
				sample_texture(rp_textures[p.tex_slot], p.tex, p.tex2, u, vv, sq[2], persp, lut, smp,
				               super_sampled ? phase_shift : 0.5f, tx);
			const bool tcc = p.combiner & C_TCC;
			switch (p.combiner & 3u)
			{
			case 1: // DECAL
				for (int k = 0; k < 4; k++) comb[k] = tx[k];
				if (!tcc) comb[3] = col[3];
				break;
			case 0: // MODULATE
				for (int k = 0; k < 4; k++)
					comb[k] = std::min(std::max(std::floor(col[k] * tx[k] * (1.0f / 128.0f)), 0.0f), 255.0f);
				if (!tcc) comb[3] = col[3];
				break;
			default: // HIGHLIGHT / HIGHLIGHT2
			{
				for (int k = 0; k < 4; k++)
					comb[k] = std::min(std::max(std::floor(col[k] * tx[k] * (1.0f / 128.0f)), 0.0f), 255.0f);
				for (int k = 0; k < 3; k++) comb[k] += col[3];
				comb[3] = tcc ? ((p.combiner & 3u) == 2 ? tx[3] + col[3] : tx[3]) : col[3];
				for (int k = 0; k < 4; k++) comb[k] = std::min(std::max(comb[k], 0.0f), 255.0f);
				break;
			}
			}
		}
		if (p.combiner & C_FOG)
		{
			float e1 = s.stqf1[3] * fi, e2 = s.stqf2[3] * fj;
			float fog = std::nearbyint((s.stqf0[3] + e1) + e2);
			for (uint32_t k = 0; k < 3; k++)
			{
				float fc = unorm(p.fogcol, k) * 255.0f;
				float d = (comb[k] - fc) * fog;
				comb[k] = std::floor(fc + d * (1.0f / 256.0f) + 0.5f);
			}
		}

		// shade_resolve (ubershader.comp:1464-1607)
		float fb[4];
		for (uint32_t k = 0; k < 4; k++)
			fb[k] = std::nearbyint(unorm(r.color, k) * 255.0f);
		bool write_fb = true, write_zb = true, write_alpha = true;
		const uint32_t bm = p.blend_mode;
		bool blending = bm & B_ABE;
		// AA1: the coverage replaces alpha 0x80 (or any alpha when not blending) and blending is
		// forced on (ubershader.comp:1497-1505).
		if (multisample)
		{
			if (comb[3] == 128.0f || !blending)
				comb[3] = float(cov);
			blending = true;
		}
		bool keep_all = false;
		if (bm & B_ATE)
		{
			if (!alpha_test((bm >> 2) & 7u, p.alpha & 0xffu, comb[3]))
			{
				switch ((bm >> 5) & 3u)
				{
				case 0: keep_all = true; break;
				case 1: write_zb = false; break;
				case 2: write_fb = false; write_alpha = false; break;
				case 3:
					if (psm32(f.fb_psm)) write_alpha = false;
					write_zb = false;
					break;
				}
			}
		}
		if (keep_all)
			return;
		if (!psm24(f.fb_psm) && (bm & B_DATE))
		{
			uint32_t bit = fb[3] >= 128.0f ? 1u : 0u;
			if (uint32_t((bm & B_DATM) != 0) != bit)
				return;
		}
		if ((bm & B_PABE) && comb[3] < 128.0f)
			blending = false;
		float out[4] = { comb[0], comb[1], comb[2], comb[3] };
		if (blending)
		{
			float A[3] = { 0, 0, 0 }, B[3] = { 0, 0, 0 }, D[3] = { 0, 0, 0 }, C = 0.0f;
			uint32_t am = (bm >> 11) & 3u, bmd = (bm >> 13) & 3u, cm = (bm >> 15) & 3u, dm = (bm >> 17) & 3u;
			for (int k = 0; k < 3; k++)
			{
				A[k] = am == 0 ? comb[k] : am == 1 ? fb[k] : 0.0f;
				B[k] = bmd == 0 ? comb[k] : bmd == 1 ? fb[k] : 0.0f;
				D[k] = dm == 0 ? comb[k] : dm == 1 ? fb[k] : 0.0f;
			}
			if (cm == 0) C = comb[3];
			else if (cm == 1) C = psm24(f.fb_psm) ? 128.0f : fb[3];
			else if (cm == 2) C = float((p.alpha >> 8) & 0xffu);
			for (int k = 0; k < 3; k++)
				out[k] = std::floor((A[k] - B[k]) * C * (1.0f / 128.0f)) + D[k];
		}
		if (bm & B_DTHE)
		{
			int32_t px = (hx >> rate) & 3, py = (hy >> rate) & 3;
			int32_t bit_index = ((py * 4 + px) * 4) & 31;
			uint32_t dm = py < 2 ? p.dimx[0] : p.dimx[1];
			int32_t dv = int32_t(dm << (29 - bit_index)) >> 29;
			for (int k = 0; k < 3; k++)
				out[k] += float(dv);
		}
		if (psm24(f.fb_psm))
			write_alpha = false;
		if (!write_alpha)
			out[3] = fb[3];
		if (!write_fb)
			for (int k = 0; k < 3; k++) out[k] = fb[k];
		float norm[4];
		if (!(bm & B_COLCLAMP))
		{
			for (int k = 0; k < 4; k++)
			{
				float q = out[k] / 256.0f;
				norm[k] = (q - std::floor(q)) * (256.0f / 255.0f);
			}
		}
		else
			for (int k = 0; k < 4; k++)
				norm[k] = out[k] * (1.0f / 255.0f);
		uint32_t packed = pack_unorm(norm);
		packed |= uint32_t((bm & B_FBA) && write_alpha) << 31;

		if (p.fbmsk != ~0u && (write_fb || write_alpha))
		{
			r.color = p.fbmsk ? ((packed & ~p.fbmsk) | (r.color & p.fbmsk)) : packed;
			if (psm16(f.fb_psm))
				r.color &= 0x80f8f8f8u;
			r.cdirty = true;
		}
		if (z_write && write_zb)
		{
			r.depth = z;
			r.zdirty = true;
		}
	};

	for (size_t i = 0; i < prims.size(); i++)
		primitives_drawn += live[i] ? 1u : 0u;
	uint8_t *v0 = vram.data();
	const size_t plane_bytes = size_t(4) << 20;

	// One shading launch over primitives lo..hi: colour read from csrc, depth from zsrc
	// (copies of VRAM or VRAM itself), stores left out per skip (bit 0 colour, bit 1 depth).
	auto run = [&](uint32_t lo, uint32_t hi, const uint8_t *csrc, const uint8_t *zsrc, uint32_t skip) {
		if (sample_log2 != 0)
		{
			render_super_sampled_pass(v0, csrc, zsrc, lo, hi, skip, rp_bb, rate_y_log2(), prims, setups, live, f,
			                          rate_x, rate_y, shade);
			return;
		}
		const uint32_t fb_base = f.fbp * 32u, z_base = f.zbp * 32u;
		std::unordered_map<uint32_t, Regs> cache;
		std::vector<uint32_t> touch_order;

		auto load = [&](int32_t x, int32_t y) -> Regs & {
			uint32_t key = uint32_t(y) * 4096u + uint32_t(x);
			Regs &r = cache[key];
			if (!r.loaded)
			{
				r.loaded = true;
				touch_order.push_back(key);
				uint32_t a = swizzle(uint32_t(x), uint32_t(y), fb_base, f.fbw, f.fb_psm);
				if (psm16(f.fb_psm))
				{
					uint16_t c16;
					memcpy(&c16, csrc + a * 2u, 2);
					r.color = rgba16_to_rgba32(c16, 0, 0, 0x80);
				}
				else
					memcpy(&r.color, csrc + a * 4u, 4);
				if (f.z_psm != 0xffffffffu)
				{
					uint32_t za = swizzle(uint32_t(x), uint32_t(y), z_base, f.fbw, f.z_psm);
					if (psm16(f.z_psm))
					{
						uint16_t d16;
						memcpy(&d16, zsrc + za * 2u, 2);
						r.depth = d16;
					}
					else
					{
						memcpy(&r.depth, zsrc + za * 4u, 4);
						if (psm24(f.z_psm))
							r.depth &= 0xffffffu;
					}
				}
			}
			return r;
		};

		for (size_t pi = lo; pi <= hi && pi < prims.size(); pi++)
		{
			if (!live[pi])
				continue;
			const Prim &p = prims[pi];
			const Setup &s = setups[pi];
			for (int32_t y = s.bb[1]; y <= s.bb[3]; y++)
				for (int32_t x = s.bb[0]; x <= s.bb[2]; x++)
					shade(p, s, x, y, 0, false, 0u, [&]() -> Regs & { return load(x, y); });
		}

		// Write-back in 8x8 tile order: colour then depth per tile.
		std::sort(touch_order.begin(), touch_order.end(), [](uint32_t a, uint32_t b) {
			uint32_t ax = a % 4096u, ay = a / 4096u, bx = b % 4096u, by = b / 4096u;
			uint32_t ta = (ay >> 3) * 512u + (ax >> 3), tb = (by >> 3) * 512u + (bx >> 3);
			if (ta != tb) return ta < tb;
			return a < b;
		});
		size_t t0 = 0;
		while (t0 < touch_order.size())
		{
			size_t t1 = t0;
			auto tile = [](uint32_t k) { return ((k / 4096u) >> 3) * 512u + ((k % 4096u) >> 3); };
			while (t1 < touch_order.size() && tile(touch_order[t1]) == tile(touch_order[t0]))
				t1++;
			for (size_t t = t0; t < t1 && !(skip & 1u); t++)
			{
				uint32_t key = touch_order[t];
				Regs &r = cache[key];
				if (!r.cdirty)
					continue;
				uint32_t x = key % 4096u, y = key / 4096u;
				uint32_t a = swizzle(x, y, fb_base, f.fbw, f.fb_psm);
				if (psm16(f.fb_psm))
				{
					uint16_t c16 = uint16_t(rgba32_to_rgba16(r.color));
					memcpy(v0 + a * 2u, &c16, 2);
				}
				else
					memcpy(v0 + a * 4u, &r.color, 4);
			}
			for (size_t t = t0; t < t1 && !(skip & 2u); t++)
			{
				uint32_t key = touch_order[t];
				Regs &r = cache[key];
				if (!r.zdirty)
					continue;
				uint32_t x = key % 4096u, y = key / 4096u;
				uint32_t a = swizzle(x, y, z_base, f.fbw, f.z_psm);
				if (psm16(f.z_psm))
				{
					uint16_t d16 = uint16_t(r.depth);
					memcpy(v0 + a * 2u, &d16, 2);
				}
				else if (psm24(f.z_psm))
					memcpy(v0 + a * 4u, &r.depth, 3);
				else
					memcpy(v0 + a * 4u, &r.depth, 4);
			}
			t0 = t1;
		}
	};

	const uint32_t n = uint32_t(prims.size());
	// Colour and depth in one buffer (gs/gs_renderer.cpp:2525-2690): a primitive's colour may
	// land on another pixel's depth, so the pass shades in launches whose reads do not see
	// their own writes:
	// * primitives before the first depth-sensitive one together (depth only read by on-tile
	//   depth feedback, from a copy);
	// * with depth writes, each primitive from the first depth-sensitive one to the last depth
	//   writer alone, reading a copy of VRAM taken before it, all colour stored, then all depth;
	// * later depth-testing primitives in runs reading depth from a copy taken as the run
	//   starts, a run ending before a depth-sensitive sprite that overlaps the run's earlier
	//   sprites (gs/gs_renderer.cpp:2418-2475);
	// * with no later depth test, the rest without depth.
	if (!(f.z_psm != 0xffffffffu && f.fbp == f.zbp))
	{
		run(0, n - 1u, v0, v0, 0);
		return;
	}
	// Copy of the planes the launch reads (plane 0 only for single-sampled shading).
	std::vector<uint8_t> copy;
	auto take_copy = [&]() -> const uint8_t * {
		const size_t bytes = (sample_log2 != 0 && rate_y != 0) ? vram.size() : plane_bytes;
		copy.assign(vram.begin(), vram.begin() + ptrdiff_t(bytes));
		return copy.data();
	};
	uint32_t first_z = ~0u, last_zwrite = 0, last_z = 0;
	bool single_step = false;
	for (uint32_t i = 0; i < n; i++)
	{
		const uint32_t st = prims[i].state;
		if (st & S_Z_WRITE)
		{
			single_step = true;
			last_zwrite = i;
		}
		if (st & (S_Z_TEST | S_Z_WRITE))
		{
			if (first_z == ~0u)
				first_z = i;
			last_z = i;
		}
	}
	if (first_z == ~0u)
	{
		run(0, n - 1u, v0, v0, 0);
		return;
	}
	auto cached_runs = [&](uint32_t lo) {
		int32_t bb[4] = { INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN };
		for (uint32_t i = lo; i < n; i++)
		{
			const Prim &p = prims[i];
			if (!(p.state & S_SPRITE))
				continue;
			const bool overlap = std::max(bb[0], p.bb[0]) <= std::min(bb[2], p.bb[2]) &&
			                     std::max(bb[1], p.bb[1]) <= std::min(bb[3], p.bb[3]) &&
			                     (p.state & (S_Z_TEST | S_Z_WRITE)) != 0;
			if (overlap)
			{
				run(lo, i - 1u, v0, take_copy(), 0);
				lo = i;
				memcpy(bb, p.bb, sizeof(bb));
			}
			else
			{
				bb[0] = std::min(bb[0], p.bb[0]);
				bb[1] = std::min(bb[1], p.bb[1]);
				bb[2] = std::max(bb[2], p.bb[2]);
				bb[3] = std::max(bb[3], p.bb[3]);
			}
		}
		if (lo < n)
			run(lo, n - 1u, v0, take_copy(), 0);
	};
	if (first_z != 0)
		run(0, first_z - 1u, v0, take_copy(), 0);
	if (!single_step)
	{
		cached_runs(first_z);
		return;
	}
	for (uint32_t i = first_z; i <= last_zwrite; i++)
	{
		const uint8_t *c = take_copy();
		run(i, i, c, c, 2u);
		run(i, i, c, c, 1u);
	}
	if (last_zwrite + 1u < n)
	{
		if (last_z > last_zwrite)
			cached_runs(last_zwrite + 1u);
		else
		{
			f.z_psm = 0xffffffffu;
			run(last_zwrite + 1u, n - 1u, v0, v0, 0);
		}
	}
}

// Textures and CLUT are handled in oracle_texture.cpp.
}
