// Shading kernel for gfx950 (replaces ubershader.comp).
//
// One wave64 owns one 8x8 block of pixels: lane L shades pixel (L & 7, L >> 3) of the block.
// Colour and depth are loaded into VGPRs once (begin_tile, ubershader.comp:435-562), every
// primitive binned to the block's coarse tile is re-tested against the block with one lane per
// primitive (64 primitives per ballot), and the hits are applied in ascending primitive order
// with the primitive's setup read through the scalar cache. Dirty values are written back once
// at the end (end_tile, ubershader.comp:581-780).
//
// Workgroups of four waves cover 16x16 pixels. Workgroup ids are remapped so that the eight
// XCDs (which receive workgroups round-robin) each own a contiguous band of the screen and
// keep that band's primitive data in their own L2.
#include "gs_defs.hpp"
#include "gs_raster.hpp"
#include "lod_table.hpp"
#include "pgs_kernels.h"
#include "multi_launch.hpp"
#include <algorithm>
#include "multi_dispatch.hpp"

namespace pgs
{
namespace
{
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// A wave-uniform pointer in scalar registers (an operand of a scalar load). Free where the
// compiler already knows the pointer is uniform; needed where it cannot tell (a kernel holding
// the shading bodies of several sampling grids).
template <typename T>
__device__ __forceinline__ const T *uni_ptr(const T *p)
{
	const uint64_t v = reinterpret_cast<uint64_t>(p);
	const uint64_t lo = uni(uint32_t(v)), hi = uni(uint32_t(v >> 32));
	return reinterpret_cast<const T *>(hi << 32 | lo);
}
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return int32_t(uint32_t(a) + uint32_t(b)); }
__device__ __forceinline__ int32_t wmul(int32_t a, int32_t b) { return int32_t(uint32_t(a) * uint32_t(b)); }

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));

// A primitive's setup record (80 bytes) and state word, both wave-uniform, in one round trip
// through the scalar cache. Left to itself the compiler places each scalar load at the first
// use of its field, and since scalar loads return out of order every wait is a wait for all of
// them: one round trip per use. The three loads and their wait sit in one asm statement, so
// the outputs are complete when it ends; the statement has no side effects (the records are
// read-only while shading).
__device__ __forceinline__ void load_setup_and_state(const PrimitiveSetup *sp, const PrimitiveAttribute *pp,
                                                     PrimitiveSetup &s, uint32_t &state)
{
	u32x16 lo;
	u32x4 hi;
	uint32_t st;
	asm("s_load_dwordx16 %0, %3, 0x0\n\t"
	    "s_load_dwordx4 %1, %3, 0x40\n\t"
	    "s_load_dword %2, %4, 0x8\n\t"
	    "s_waitcnt lgkmcnt(0)"
	    : "=&s"(lo), "=&s"(hi), "=&s"(st)
	    : "s"(uni_ptr(sp)), "s"(uni_ptr(pp)));
	__builtin_memcpy(&s, &lo, 64);
	__builtin_memcpy(reinterpret_cast<uint8_t *>(&s) + 64, &hi, 16);
	state = st;
}

typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));

// The second round, once the state word is known: the attribute record's colour terms, the
// state vector, and the primitive's texture, alpha, frame-buffer mask and fog colour words.
struct PrimitiveShadeData
{
	float col0[4], dcol1[4], dcol2[4];
	StateVector sv;
	uint32_t tex, tex2, alpha, fbmsk, fogcol;
};
__device__ __forceinline__ void load_shade_data(const TransformedAttributes *tp, const StateVector *svp,
                                                const PrimitiveAttribute *pp, PrimitiveShadeData &d)
{
	u32x8 col_a;
	u32x4 col_b, sv, pa;
	uint32_t tex;
	asm("s_load_dwordx8 %0, %5, 0x40\n\t"
	    "s_load_dwordx4 %1, %5, 0x60\n\t"
	    "s_load_dwordx4 %2, %6, 0x0\n\t"
	    "s_load_dwordx4 %3, %7, 0x10\n\t"
	    "s_load_dword %4, %7, 0xc\n\t"
	    "s_waitcnt lgkmcnt(0)"
	    : "=&s"(col_a), "=&s"(col_b), "=&s"(sv), "=&s"(pa), "=&s"(tex)
	    : "s"(uni_ptr(tp)), "s"(uni_ptr(svp)), "s"(uni_ptr(pp)));
	__builtin_memcpy(&d.col0[0], &col_a, 32);
	__builtin_memcpy(&d.dcol2[0], &col_b, 16);
	__builtin_memcpy(&d.sv, &sv, 16);
	__builtin_memcpy(&d.tex2, &pa, 16);
	d.tex = tex;
}

// The third round, for a textured primitive: the S, T, Q terms of the attribute record and the
// texture descriptor of the primitive's slot.
__device__ __forceinline__ void load_texture_data(const TransformedAttributes *tp, const TextureDesc *tdp,
                                                  float stq[3][4], TextureDesc &td)
{
	u32x8 stq_a;
	u32x4 stq_b;
	u32x16 desc;
	asm("s_load_dwordx8 %0, %3, 0x0\n\t"
	    "s_load_dwordx4 %1, %3, 0x20\n\t"
	    "s_load_dwordx16 %2, %4, 0x0\n\t"
	    "s_waitcnt lgkmcnt(0)"
	    : "=&s"(stq_a), "=&s"(stq_b), "=&s"(desc)
	    : "s"(uni_ptr(tp)), "s"(uni_ptr(tdp)));
	__builtin_memcpy(&stq[0][0], &stq_a, 32);
	__builtin_memcpy(&stq[2][0], &stq_b, 16);
	__builtin_memcpy(&td, &desc, 64);
}
// Edge function at a sample, wrapping in 32 bits like the reference's int arithmetic. The
// products use the full-rate 24-bit multiply, whose low 32 bits equal the 32-bit product when
// both operands fit in signed 24 bits: the edge slopes are differences of vertex coordinates
// (16-bit register fields minus a 16-bit offset, moved by at most 17 for lines), so below 2^18
// in magnitude, and the sample coordinates are pixel positions (below 2^11) times 8.
__device__ __forceinline__ int32_t idot3(int32_t e0, int32_t e1, int32_t e2, int32_t x, int32_t y)
{
	return wadd(wadd(__mul24(e0, x), __mul24(e1, y)), e2);
}

// b / 255 correctly rounded for every byte (the constant folder divides in IEEE single
// precision, round to nearest even): for wave-uniform bytes the lookup is a scalar load where
// the division would be ten vector instructions.
struct Unorm8Table
{
	float v[256];
};
constexpr Unorm8Table make_unorm8_table()
{
	Unorm8Table t{};
	for (int i = 0; i < 256; i++)
		t.v[i] = float(i) / 255.0f;
	return t;
}
__constant__ Unorm8Table c_unorm8 = make_unorm8_table();
__device__ __forceinline__ float unorm8(uint32_t c, uint32_t ch) { return c_unorm8.v[(c >> (8 * ch)) & 0xffu]; }
// rint(unorm8(b) * 255) == b and unorm8(b) * 255 == b hold for all 256 bytes (exhaustive), so
// the frame buffer and fog colour channels are the bytes themselves.
__device__ __forceinline__ float byte_f(uint32_t c, uint32_t ch) { return float((c >> (8 * ch)) & 0xffu); }

__device__ __forceinline__ uint32_t pack_unorm(float r, float g, float b, float a)
{
	auto q = [](float c) { return uint32_t(rintf(__fmul_rn(fminf(fmaxf(c, 0.0f), 1.0f), 255.0f))); };
	return q(r) | (q(g) << 8) | (q(b) << 16) | (q(a) << 24);
}

__device__ __forceinline__ uint32_t rgba16_to_rgba32(uint32_t c16)
{
	return ((c16 & 0x1fu) << 3) | (((c16 >> 5) & 0x1fu) << 11) | (((c16 >> 10) & 0x1fu) << 19) |
	       (((c16 >> 15) & 1u) ? 0x80000000u : 0u);
}
__device__ __forceinline__ uint32_t rgba32_to_rgba16(uint32_t c)
{
	return ((c >> 3) & 0x1fu) | (((c >> 11) & 0x1fu) << 5) | (((c >> 19) & 0x1fu) << 10) | ((c >> 31) << 15);
}

__device__ __forceinline__ bool intersects_block(const PrimitiveSetup &s, int32_t tlox, int32_t tloy, int32_t thix,
                                                 int32_t thiy, bool multisampled_tri)
{
	int32_t lox = max(s.bb[0], tlox), loy = max(s.bb[1], tloy);
	int32_t hix = min(s.bb[2], thix), hiy = min(s.bb[3], thiy);
	if (hix < lox || hiy < loy)
		return false;
	const bool para = (s.a[0] & 1) != 0;
	const int32_t offb = int32_t((uint32_t(s.a[0]) >> 2) & 3u) - 1;
	const int32_t offc = int32_t((uint32_t(s.a[0]) >> 4) & 3u) - 1;
	const int32_t ax = s.a[0] >> 6;
	lox <<= 3; loy <<= 3;
	hix = (hix << 3) | 7; hiy = (hiy << 3) | 7;
	int32_t da = idot3(ax, s.a[1], s.a[2], ax >= 0 ? hix : lox, s.a[1] >= 0 ? hiy : loy);
	int32_t db = idot3(s.b[0], s.b[1], s.b[2], s.b[0] >= 0 ? hix : lox, s.b[1] >= 0 ? hiy : loy);
	int32_t dc = idot3(s.c[0], s.c[1], s.c[2], s.c[0] >= 0 ? hix : lox, s.c[1] >= 0 ? hiy : loy);
	if (multisampled_tri && !para)
	{
		da = wadd(da, wmul(8, max(abs(ax), abs(s.a[1]))));
		db = wadd(db, wmul(8, max(abs(s.b[0]), abs(s.b[1]))));
		dc = wadd(dc, wmul(8, max(abs(s.c[0]), abs(s.c[1]))));
	}
	bool hit = da >= 0 && db >= 0 && dc >= 0;
	if (para && !hit)
	{
		const bool mxa = ax >= 0, mya = s.a[1] >= 0;
		int32_t pa = idot3(ax, s.a[1], s.a[2], mxa ? hix : lox, mya ? hiy : loy);
		int32_t pb = idot3(s.b[0], s.b[1], s.b[2], -s.c[0] >= 0 ? hix : lox, -s.c[1] >= 0 ? hiy : loy);
		int32_t pc = idot3(s.c[0], s.c[1], s.c[2], -s.b[0] >= 0 ? hix : lox, -s.b[1] >= 0 ? hiy : loy);
		pb = wadd(pb, wadd(pa, offb));
		pc = wadd(pc, wadd(pa, offc));
		int32_t na = idot3(ax, s.a[1], s.a[2], mxa ? lox : hix, mya ? loy : hiy);
		hit = -1 >= na && pb >= 0 && pc >= 0;
	}
	return hit;
}

__device__ __forceinline__ bool alpha_test(uint32_t atst, uint32_t aref, float a)
{
	const float r = float(aref);
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

// Wrap of an integer-valued texel coordinate vf into [0, size) (REPEAT on the decoded image), for
// |vf| < 2^22: q * size is exact in float, floor(vf * (1 / size)) is off by at most one, and the
// two selects correct that, so the result is the mathematical remainder vf mod size for every
// size (seven full-rate instructions instead of an integer division). The sampler's coordinates
// are floors of texel positions below 2^22 in magnitude (|u| <= 2047 texture widths of at most
// 1024), so they arrive here as floats without an integer round trip.
__device__ __forceinline__ uint32_t wrap_texel(float vf, float size, float inv_size)
{
	const float q = floorf(__fmul_rn(vf, inv_size));
	float r = __fsub_rn(vf, __fmul_rn(q, size));
	r = r < 0.0f ? __fadd_rn(r, size) : r;
	r = r >= size ? __fsub_rn(r, size) : r;
	return uint32_t(r);
}

// Per byte of two RGBA8 words, a + floor(((b - a) * w + 128) / 256) for a weight w in 0..256
// (round-half-up lerp). Since 256 a + (b - a) w + 128 = a (256 - w) + b w + 128, which is at most
// 255 * 256 + 128 < 2^16, two bytes held 16 bits apart (mask 0x00ff00ff) are blended in one
// 24-bit multiply-add each without carrying into each other, and the field's bits 8..15 are
// the result.
__device__ __forceinline__ uint32_t lerp8x4(uint32_t a, uint32_t b, uint32_t w)
{
	constexpr uint32_t M = 0x00ff00ffu;
	const uint32_t iw = 256u - w;
	const uint32_t even = __umul24(a & M, iw) + __umul24(b & M, w) + 0x00800080u;
	const uint32_t odd = __umul24((a >> 8) & M, iw) + __umul24((b >> 8) & M, w) + 0x00800080u;
	return ((even >> 8) & M) | (odd & ~M);
}

// One level's texel fetch for normalized coordinates (un, vn) on a w x h image: NEAREST, or
// LINEAR with 8-bit fractional weights and an integer lerp (U, then V), both with REPEAT
// addressing. Images are at most 1024 texels on a side, so texel offsets fit the 24-bit
// multiply. fw, fh are the sizes as floats and inv_w, inv_h their rounded reciprocals.
__device__ __forceinline__ uint32_t fetch_level(const uint32_t *img, uint32_t w, uint32_t h, float fw, float fh,
                                                float inv_w, float inv_h, float un, float vn, bool linear)
{
	if (!linear)
	{
		const uint32_t x = wrap_texel(floorf(__fmul_rn(un, fw)), fw, inv_w);
		const uint32_t y = wrap_texel(floorf(__fmul_rn(vn, fh)), fh, inv_h);
		return img[__umul24(y, w) + x];
	}
	const float xs = __fsub_rn(__fmul_rn(un, fw), 0.5f), ys = __fsub_rn(__fmul_rn(vn, fh), 0.5f);
	const float x0 = floorf(xs), y0 = floorf(ys);
	const uint32_t wx = uint32_t(__fmul_rn(__fsub_rn(xs, x0), 256.0f));
	const uint32_t wy = uint32_t(__fmul_rn(__fsub_rn(ys, y0), 256.0f));
	// REPEAT on the decoded image; the second texel of each axis wraps from the first.
	const uint32_t ax = wrap_texel(x0, fw, inv_w), ay = wrap_texel(y0, fh, inv_h);
	const uint32_t bx = ax + 1u == w ? 0u : ax + 1u, by = ay + 1u == h ? 0u : ay + 1u;
	const uint32_t r0 = __umul24(ay, w), r1 = __umul24(by, w);
	const uint32_t c00 = img[r0 + ax], c10 = img[r0 + bx], c01 = img[r1 + ax], c11 = img[r1 + bx];
	// wx, wy: fractions below 1 times 256, so at most 256 (a fraction just below 1 can round
	// to 1).
	return lerp8x4(lerp8x4(c00, c10, wx), lerp8x4(c01, c11, wx), wy);
}

// The Q-derived level of detail in 1/16 units, ceil(K - 2^(L+4) * log2|q|) (ubershader.comp:
// 1076-1087), exactly: with |q| = m * 2^e, m in [1, 2), it is K - 2^(L+4) * e -
// floor(2^(L+4) * log2 m), and floor(128 * log2 m) is the largest k with m >= kLodThreshold[k].
// A zero, subnormal, infinite or NaN q gives 0 (the reference flushes non-finite LODs to 0; a
// subnormal q is taken as flushed to zero).
__device__ __forceinline__ int32_t lod_from_q(float q, int32_t k, uint32_t l)
{
	const uint32_t b = __float_as_uint(q) & 0x7fffffffu;
	const uint32_t e = b >> 23;
	if (e == 0u || e == 255u)
		return 0;
	const float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
	// An estimate from the hardware log, then corrected against the exact thresholds.
	int32_t j = min(max(int32_t(__fmul_rn(__log2f(m), 128.0f)), 0), 127);
	while (j > 0 && m < kLodThreshold[j])
		j--;
	while (j < 127 && m >= kLodThreshold[j + 1])
		j++;
	return k - (int32_t(e) - 127) * (16 << l) - (j >> (3u - l));
}

// Level `lod` of a mipmapped texture (ubershader.comp:1133-1254): coordinates scaled by 2^-lod,
// clamped to the region scaled to the level (texture_clamp, ubershader.comp:947-953), snapped
// to 1/16 texel, normalized with the level's 2^lod / width; the image level read is lod clamped
// to the image's last level. `bias` is false for the second level of a trilinear blend, which
// the reference normalizes without the rectangle origin.
__device__ __forceinline__ uint32_t sample_level(const TextureDesc &t, const uint32_t *images, uint32_t tex, float u,
                                                 float v, bool perspective, uint32_t lod, bool linear, bool bias)
{
	const int32_t nl = -int32_t(lod);
	float u1 = __fmul_rn(u, ldexpf(perspective ? t.size[0] : 1.0f, nl));
	float v1 = __fmul_rn(v, ldexpf(perspective ? t.size[1] : 1.0f, nl));
	// The descriptor's bounds are floor(region) + 0.5 (or infinite): region - 0.5 is the region.
	auto bound = [&](float r) { return __fadd_rn(floorf(ldexpf(__fsub_rn(r, 0.5f), nl)), 0.5f); };
	if (tex & TEX_CLAMP_S)
		u1 = __builtin_amdgcn_fmed3f(u1, bound(t.region[0]), bound(t.region[2]));
	if (tex & TEX_CLAMP_T)
		v1 = __builtin_amdgcn_fmed3f(v1, bound(t.region[1]), bound(t.region[3]));
	float un = floorf(__fadd_rn(__fmul_rn(u1, 16.0f), 1.0f / 64.0f));
	float vn = floorf(__fadd_rn(__fmul_rn(v1, 16.0f), 1.0f / 64.0f));
	un = __fmul_rn(__fmul_rn(__fadd_rn(un, 16.0f / 1024.0f), ldexpf(t.size[2], int32_t(lod))), 0.0625f);
	vn = __fmul_rn(__fmul_rn(__fadd_rn(vn, 16.0f / 1024.0f), ldexpf(t.size[3], int32_t(lod))), 0.0625f);
	if (bias)
	{
		un = __fadd_rn(un, t.bias[0]);
		vn = __fadd_rn(vn, t.bias[1]);
	}
	// The image level: lod, at most the last. Every level but a 1-texel-wide clamped one divides
	// the level-0 size exactly (sizes are powers of two, or multiples of 2^(levels - 1)), so the
	// level's reciprocal size is the level-0 one scaled.
	const uint32_t lv = min(lod, t.levels - 1u);
	uint32_t off = t.image_offset;
	for (uint32_t i = 0; i < lv; i++)
		off += max(t.width >> i, 1u) * max(t.height >> i, 1u);
	const uint32_t w = t.width >> lv, h = t.height >> lv;
	const float inv_w = w ? ldexpf(t.size[2], int32_t(lv)) : 1.0f, inv_h = h ? ldexpf(t.size[3], int32_t(lv)) : 1.0f;
	const uint32_t w1 = max(w, 1u), h1 = max(h, 1u);
	return fetch_level(images + off, w1, h1, float(w1), float(h1), inv_w, inv_h, un, vn, linear);
}

// Mipmapped sampling (ubershader.comp:1066-1131, 1233-1254): the level of detail from LCM / K
// or from Q, clamped to [0, MXL]; the filter is MMIN's when the LOD is above 0, MMAG's
// otherwise. Trilinear blends the two nearest levels with floor(fract * 16) / 16 weights,
// rounding half to even; otherwise the nearest level is read.
__device__ __forceinline__ uint32_t sample_mipmapped(const TextureDesc &t, const uint32_t *images, uint32_t tex,
                                                     uint32_t tex2, float u, float v, float q, bool perspective)
{
	const uint32_t mxl = (tex >> TEX_MAX_MIP_LEVEL_SHIFT) & 7u;
	const int32_t k = int32_t(tex2 << (32u - TEX2_K_SHIFT - 12u)) >> 20; // signed 12-bit K
	int32_t n;
	if (tex2 & TEX2_FIXED_LOD)
		n = (k >> 4) * 16;
	else
		n = lod_from_q(q, k, (tex2 >> TEX2_L_SHIFT) & 3u);
	const bool linear = (tex & (n > 0 ? TEX_MIN_LINEAR : TEX_MAG_LINEAR)) != 0;
	const uint32_t lod16 = uint32_t(min(max(n, 0), int32_t(mxl * 16u)));
	if (!(tex & TEX_MIP_LINEAR) || (lod16 & 15u) == 0u)
		return sample_level(t, images, tex, u, v, perspective, (tex & TEX_MIP_LINEAR) ? lod16 >> 4 : (lod16 + 8u) >> 4,
		                    linear, true);
	const uint32_t c1 = sample_level(t, images, tex, u, v, perspective, lod16 >> 4, linear, true);
	const uint32_t c2 = sample_level(t, images, tex, u, v, perspective, (lod16 + 15u) >> 4, linear, false);
	const uint32_t w = lod16 & 15u;
	uint32_t c = 0;
	for (uint32_t ch = 0; ch < 32u; ch += 8u)
	{
		const uint32_t x = ((c1 >> ch) & 0xffu) * (16u - w) + ((c2 >> ch) & 0xffu) * w; // x / 16, half to even
		c |= ((x + 7u + ((x >> 4) & 1u)) >> 4) << ch;
	}
	return c;
}

// Nearest texel of layer `layer` of a super-sampled texture's image (REPEAT addressing).
__device__ __forceinline__ uint32_t layer_texel(const TextureDesc &t, const uint32_t *images, uint32_t layer, float un,
                                                float vn)
{
	const uint32_t x = wrap_texel(floorf(__fmul_rn(un, t.fwidth)), t.fwidth, t.size[2]);
	const uint32_t y = wrap_texel(floorf(__fmul_rn(vn, t.fheight)), t.fheight, t.size[3]);
	return images[t.image_offset + __umul24(layer, t.width * t.height) + __umul24(y, t.width) + x];
}

// The raw clamp region of one bound of a super-sampled texture's descriptor: the stored
// floor(region) + 0.5 less 0.5 on a clamping axis, 0 on the others (whose stored bound is
// infinite), as the reference's TexInfo region (gs/gs_interface.cpp:1758-1777).
__device__ __forceinline__ float raw_region(float stored) { return __builtin_isinf(stored) ? 0.0f : __fsub_rn(stored, 0.5f); }

// Super-sampled texture (ubershader.comp:1155-1211): layer 0 holds the single-sampled image,
// layer k + 1 the image of sample k. PER_SAMPLE: a filter over the samples nearest to the
// coordinate's phase (the phase table's four taps and unorm8 weights) with nearest fetches,
// or the first tap alone when magnification is NEAREST or the texel is on a clamped edge;
// otherwise a NEAREST / LINEAR fetch of the wave's own sample's layer (SAMPLE_MAPPING) or of
// layer 0. The coordinates are not snapped to 1/16 texel. `smp` is the lane's sample and
// `shift` 0.5 less the grid's mean sample offset (0.5 in a single-sampled wave).
__device__ __forceinline__ uint32_t sample_arrayed(const TextureDesc &t, const uint32_t *images, const uint32_t *lut,
                                                   uint32_t tex, float u1, float v1, uint32_t smp, float shift)
{
	const float uu = __fadd_rn(u1, 1.0f / 1024.0f), vv = __fadd_rn(v1, 1.0f / 1024.0f);
	const bool linear = (tex & TEX_MAG_LINEAR) != 0;
	if ((tex & (TEX_PER_SAMPLE | TEX_SAMPLE_MAPPING | TEX_SAMPLE_RESOLVED)) != TEX_PER_SAMPLE)
	{
		const float un = __fadd_rn(__fmul_rn(uu, t.size[2]), t.bias[0]);
		const float vn = __fadd_rn(__fmul_rn(vv, t.size[3]), t.bias[1]);
		const uint32_t layer = (tex & TEX_SAMPLE_MAPPING) ? smp + 1u : 0u;
		return fetch_level(images + t.image_offset + __umul24(layer, t.width * t.height), t.width, t.height, t.fwidth,
		                   t.fheight, t.size[2], t.size[3], un, vn, linear);
	}
	const float su = __fsub_rn(uu, shift), sv = __fsub_rn(vv, shift);
	const float fu = floorf(su), fv = floorf(sv);
	// Phase table entry: 64 x 64 for LINEAR, 32 x 32 (after it) for NEAREST.
	const float n = linear ? 64.0f : 32.0f;
	const uint32_t ix = uint32_t(__fmul_rn(__fsub_rn(su, fu), n)), iy = uint32_t(__fmul_rn(__fsub_rn(sv, fv), n));
	const uint32_t e = linear ? iy * 64u + ix : 4096u + iy * 32u + ix;
	const uint2 entry = reinterpret_cast<const uint2 *>(lut)[e];
	const float cu = __fadd_rn(__fmul_rn(__fadd_rn(fu, 0.5f), t.size[2]), t.bias[0]);
	const float cv = __fadd_rn(__fmul_rn(__fadd_rn(fv, 0.5f), t.size[3]), t.bias[1]);
	const uint32_t c0 = layer_texel(t, images, (entry.x & 15u) + 1u, cu, cv);
	const bool on_edge = (tex & (TEX_CLAMP_S | TEX_CLAMP_T)) &&
	                     (fu == raw_region(t.region[0]) || fv == raw_region(t.region[1]) ||
	                      fu == raw_region(t.region[2]) || fv == raw_region(t.region[3]));
	if (!linear || on_edge)
		return c0;
	float acc[4];
	const float w0 = __fdiv_rn(float(entry.y & 0xffu), 255.0f);
	for (uint32_t ch = 0; ch < 4u; ch++)
		acc[ch] = __fmul_rn(__fdiv_rn(float((c0 >> (8u * ch)) & 0xffu), 255.0f), w0);
	for (uint32_t i = 1; i < 4u; i++)
	{
		const uint32_t wb = (entry.y >> (8u * i)) & 0xffu;
		if (wb == 0u)
			continue;
		const float w = __fdiv_rn(float(wb), 255.0f);
		// Signed 2-bit texel offsets of the tap.
		const int32_t ox = int32_t(entry.x << (14u - 4u * i)) >> 30, oy = int32_t(entry.x << (12u - 4u * i)) >> 30;
		const float tu = __fadd_rn(cu, __fmul_rn(float(ox), t.size[2]));
		const float tv = __fadd_rn(cv, __fmul_rn(float(oy), t.size[3]));
		const uint32_t ci = layer_texel(t, images, ((entry.x >> (4u * i)) & 15u) + 1u, tu, tv);
		for (uint32_t ch = 0; ch < 4u; ch++)
			acc[ch] = __fadd_rn(acc[ch], __fmul_rn(w, __fdiv_rn(float((ci >> (8u * ch)) & 0xffu), 255.0f)));
	}
	uint32_t c = 0;
	for (uint32_t ch = 0; ch < 4u; ch++)
		c |= uint32_t(rintf(__fmul_rn(255.0f, acc[ch]))) << (8u * ch);
	return c;
}

// sample_texture (ubershader.comp:1042-1258). At LOD 0 (no mipmap): UV snap to 1/16 texel with
// the reference's epsilons, then a NEAREST or LINEAR fetch; super-sampled textures take the
// array path after the clamp.
template <bool kArrays>
__device__ __forceinline__ void sample_texture(const TextureDesc &t, const uint32_t *images, uint32_t tex,
                                               uint32_t tex2, float u, float v, float q, bool perspective,
                                               const uint32_t *lut, uint32_t smp, float shift, float out[4])
{
	uint32_t c;
	if (tex & (7u << TEX_MAX_MIP_LEVEL_SHIFT))
		c = sample_mipmapped(t, images, tex, tex2, u, v, q, perspective);
	else
	{
		const uint32_t *img = images + t.image_offset;
		float u1 = perspective ? __fmul_rn(u, t.size[0]) : u;
		float v1 = perspective ? __fmul_rn(v, t.size[1]) : v;
		// CLAMP / REGION_CLAMP; the descriptor holds floor(region) + 0.5, or infinities on an
		// axis that does not clamp, with min <= max (a reversed region is stored as [max, max],
		// which min(max(x, lo), hi) also gives). u1 and v1 are finite (u and v were clamped to
		// +-2047), so the median of (x, lo, hi) is that clamp.
		u1 = __builtin_amdgcn_fmed3f(u1, t.region[0], t.region[2]);
		v1 = __builtin_amdgcn_fmed3f(v1, t.region[1], t.region[3]);
		if constexpr (kArrays)
		{
			if (tex & TEX_ARRAYED)
			{
				c = sample_arrayed(t, images, lut, tex, u1, v1, smp, shift);
				for (uint32_t ch = 0; ch < 4; ch++)
					out[ch] = float((c >> (8 * ch)) & 0xffu);
				return;
			}
		}
		float un = floorf(__fadd_rn(__fmul_rn(u1, 16.0f), 1.0f / 64.0f));
		float vn = floorf(__fadd_rn(__fmul_rn(v1, 16.0f), 1.0f / 64.0f));
		un = __fadd_rn(__fmul_rn(__fmul_rn(__fadd_rn(un, 16.0f / 1024.0f), t.size[2]), 0.0625f), t.bias[0]);
		vn = __fadd_rn(__fmul_rn(__fmul_rn(__fadd_rn(vn, 16.0f / 1024.0f), t.size[3]), 0.0625f), t.bias[1]);
		c = fetch_level(img, t.width, t.height, t.fwidth, t.fheight, t.size[2], t.size[3], un, vn,
		                (tex & TEX_MAG_LINEAR) != 0);
	}
	for (uint32_t ch = 0; ch < 4; ch++)
		out[ch] = float((c >> (8 * ch)) & 0xffu);
}

// On-tile frame-buffer feedback (ubershader.comp:955-1040): the texel is the pixel's (in a
// super-sampled pass the sample's) current colour, or depth, read as the texture's format:
// palette formats look up the CLUT with the index bits of the word (CSA banks as in the texture
// decode, 8-bit indices clamped to the last bank), 24-bit formats take TA0 as alpha, 16-bit ones
// TA0 / TA1 by the alpha bit, with AEM zeroing alpha for black.
__device__ __forceinline__ void feedback_texel(const pgs_shade_args &a, uint32_t tex, uint32_t tex2, uint32_t color,
                                               uint32_t depth, float out[4])
{
	const uint32_t in = (tex & TEX_FEEDBACK_DEPTH) ? depth : color;
	const uint32_t psm = (tex >> TEX_FEEDBACK_PSM_SHIFT) & 63u, cpsm = (tex >> TEX_FEEDBACK_CPSM_SHIFT) & 15u;
	const bool aem = (tex2 & TEX2_FEEDBACK_AEM) != 0;
	const uint32_t ta0 = (tex2 >> TEX2_FEEDBACK_TA0_SHIFT) & 0xffu, ta1 = (tex2 >> TEX2_FEEDBACK_TA1_SHIFT) & 0xffu;
	uint32_t c;
	if (psm == PSMT8 || psm == PSMT8H || psm == PSMT4 || psm == PSMT4HL || psm == PSMT4HH)
	{
		uint32_t csa = tex & 31u;
		if (cpsm == PSMCT32)
			csa &= 15u;
		uint32_t index = psm == PSMT8H ? in >> 24 : psm == PSMT4HL ? (in >> 24) & 15u : psm == PSMT4HH ? in >> 28 : 0u;
		index += csa * 16u;
		if (psm == PSMT8 || psm == PSMT8H)
			index = min(index >> 4, cpsm == PSMCT32 ? 15u : 31u) * 16u + (index & 15u);
		if (cpsm == PSMCT32)
			c = uint32_t(a.clut[index]) | (uint32_t(a.clut[index + 256u]) << 16);
		else
		{
			const uint32_t lo = a.clut[index];
			const uint32_t alpha = (aem && lo == 0u) ? 0u : ((lo >> 15) ? ta1 : ta0);
			c = ((lo & 0x1fu) << 3) | (((lo >> 5) & 0x1fu) << 11) | (((lo >> 10) & 0x1fu) << 19) | (alpha << 24);
		}
	}
	else
	{
		c = in;
		const bool zero_alpha = aem && (c & 0xffffffu) == 0u;
		if (psm == PSMCT24 || psm == PSMZ24)
			c = (c & 0xffffffu) | ((zero_alpha ? 0u : ta0) << 24);
		else if (psm == PSMCT16 || psm == PSMCT16S || psm == PSMZ16 || psm == PSMZ16S)
			c = (c & 0xffffffu) | ((zero_alpha ? 0u : ((c >> 24) == 0u ? ta0 : ta1)) << 24);
	}
	for (uint32_t ch = 0; ch < 4; ch++)
		out[ch] = float((c >> (8 * ch)) & 0xffu);
}

struct PassInfo
{
	uint32_t fb_psm, z_psm;
	bool fb16, fb24, fb32, z16, z24, zsens;
	uint32_t zmax; // largest depth the Z format holds
};

// AA1 coverage of a sample (intersect.h:156-244): a line's coverage falls off linearly across
// its two-pixel-wide parallelogram; a triangle sample outside exactly one edge gets a feathered
// coverage from its distance to that edge along the edge's major axis (below 0x80).
__device__ __forceinline__ uint32_t aa1_coverage(const PrimitiveSetup &s, bool line, bool centre, bool para,
                                                 bool ea, bool eb, bool ec, float fi, float fj)
{
	if (line)
	{
		const float d = fabsf(__fsub_rn(0.5f, fj));
		return centre ? uint32_t(int32_t(fmaxf(0.0f, __fsub_rn(128.0f, __fmul_rn(256.0f, d))))) : 0u;
	}
	if (centre)
		return 0x80u;
	if (para || int(ea) + int(eb) + int(ec) != 2)
		return 0u;
	int32_t max_delta;
	float negative_bary;
	if (!ea)
	{
		max_delta = max(abs(s.b[0] + s.c[0]), abs(s.b[1] + s.c[1]));
		negative_bary = __fsub_rn(__fadd_rn(fi, fj), 1.0f);
	}
	else if (!eb)
	{
		max_delta = max(abs(s.b[0]), abs(s.b[1]));
		negative_bary = -fi;
	}
	else
	{
		max_delta = max(abs(s.c[0]), abs(s.c[1]));
		negative_bary = -fj;
	}
	const float scaled_delta = __fmul_rn(float(max_delta), s.inv_area);
	const float cov = __fsub_rn(8.0f, __fdiv_rn(negative_bary, scaled_delta));
	return uint32_t(int32_t(fminf(fmaxf(__fmul_rn(16.0f, cov), 0.0f), 127.0f)));
}

// Applies one primitive to this lane's sample (ubershader.comp shade_early + shade_resolve).
// (hx, hy) is the sample's coordinate on the grid of 2^rate samples per pixel axis; the pixel
// is (hx >> rate, hy >> rate). Single-sampled shading passes rate 0 and super_sampled false.
// kPredicated: per-lane failures only clear `alive` and the wave leaves early when no lane is
// left (measured faster for single-sampled shading, where large primitives cover whole waves);
// otherwise every per-lane failure returns at once (faster for super-sampled shading of small
// primitives). Both give the same pixels.
// 0.5 less the mean sample offset of a sampling grid (utils.h:108-125): the phase of a
// super-sampled texture's filter is taken relative to it (ubershader.comp:1167).
__device__ __forceinline__ float phase_shift(uint32_t rate_x, uint32_t rate_y)
{
	if (rate_x == rate_y)
		return rate_y == 1u ? 0.25f : 0.125f;
	if (rate_y == 1u)
		return 0.25f;
	return rate_y == 2u ? 0.125f : 0.0625f;
}

template <bool kPredicated, bool kArrays = false>
__device__ __forceinline__ void shade_pixel(const pgs_shade_args &a, const PassInfo &pi, uint32_t prim_index,
                                            int32_t hx, int32_t hy, uint32_t rate, bool super_sampled, uint32_t smp,
                                            uint32_t &color, uint32_t &depth, bool &cdirty, bool &zdirty)
{
	const PrimitiveSetup *setups = static_cast<const PrimitiveSetup *>(a.setup);
	const PrimitiveAttribute *prims = static_cast<const PrimitiveAttribute *>(a.prims);
	uint32_t st;
	PrimitiveSetup s;
	load_setup_and_state(setups + prim_index, prims + prim_index, s, st);

	// Snapped primitives rasterize at the pixel's first sample (ubershader.comp:815-825).
	const int32_t snap_mask = -(1 << rate);
	const bool snap_raster = super_sampled && (st & ST_SNAP_RASTER);
	const int32_t rx = snap_raster ? (hx & snap_mask) : hx, ry = snap_raster ? (hy & snap_mask) : hy;
	const int32_t px = hx >> rate, py = hy >> rate;

	// Predicated form: per-lane conditions only clear `alive` (a divergent branch saves nothing
	// while some lane still runs it) and the colour and depth writes at the end are selects.
	// Nothing before them touches memory except loads of the primitive's data and in-bounds
	// texel loads.
	// Coverage (intersect.h:156-244).
	bool alive = (px >= s.bb[0]) & (py >= s.bb[1]) & (px <= s.bb[2]) & (py <= s.bb[3]);
	if (st & (ST_SCANMSK_EVEN | ST_SCANMSK_ODD))
	{
		// Sprites mask per pixel row, everything else per sample row.
		const int32_t sy_row = (st & ST_SPRITE) ? py : hy;
		alive = alive && !((sy_row & 1) == 0 ? (st & ST_SCANMSK_EVEN) : (st & ST_SCANMSK_ODD));
	}
	if constexpr (!kPredicated)
		if (!alive)
			return;
	const bool para = (s.a[0] & 1) != 0;
	const int32_t offb = int32_t((uint32_t(s.a[0]) >> 2) & 3u) - 1;
	const int32_t offc = int32_t((uint32_t(s.a[0]) >> 4) & 3u) - 1;
	const int32_t sx = int32_t(uint32_t(rx) << (3u - rate)), sy = int32_t(uint32_t(ry) << (3u - rate));
	int32_t A = idot3(s.a[0] >> 6, s.a[1], s.a[2], sx, sy);
	int32_t B = idot3(s.b[0], s.b[1], s.b[2], sx, sy);
	int32_t C = idot3(s.c[0], s.c[1], s.c[2], sx, sy);
	float fi = __fadd_rn(__fmul_rn(float(B), s.inv_area), s.error_i);
	float fj = __fadd_rn(__fmul_rn(float(C), s.inv_area), s.error_j);
	const bool edge_a = A >= 0, edge_b = B >= 0, edge_c = C >= 0;
	if (para && A < 0)
	{
		B = wadd(B, wadd(A, offb));
		C = wadd(C, wadd(A, offc));
		A = 0;
	}
	const bool centre = A >= 0 && B >= 0 && C >= 0;
	// AA1 (wave-uniform state bit): coverage below 0x80 on line fringes and near triangle edges.
	const bool multisample = (st & ST_MULTISAMPLE) != 0;
	uint32_t coverage = centre ? 0x80u : 0u;
	if (multisample)
		coverage = aa1_coverage(s, (st & ST_LINE) != 0, centre, para, edge_a, edge_b, edge_c, fi, fj);
	alive = alive && coverage != 0;

	// Z (ubershader.comp:852-871), before the coverage exit: it reads the rest of the setup
	// record, which then arrives with the first part.
	uint32_t z = s.z[0];
	{
		const float t0 = __fmul_rn(__uint_as_float(s.z[1]), fi);
		const float t1 = __fmul_rn(__uint_as_float(s.z[2]), fj);
		float zf = rintf(__fadd_rn(t0, t1));
		zf = fminf(fmaxf(zf, 0.0f), 4294967040.0f);
		uint32_t dz = uint32_t(zf);
		dz = min(dz, 0xffffffffu - z);
		z += dz;
		z = min(z, pi.zmax);
	}
	if constexpr (kPredicated)
	{
		if (__ballot(alive) == 0)
			return;
	}
	else if (!alive)
		return;

	const bool z_test = pi.zsens && (st & ST_Z_TEST);
	const bool z_write = pi.zsens && (st & ST_Z_WRITE);
	if (z_test)
	{
		// GREATER or GEQUAL as lane masks combined with the wave-uniform choice.
		const bool equal_passes = (st & ST_Z_GREATER) == 0;
		alive = alive & ((z > depth) | ((z == depth) & equal_passes));
		if constexpr (kPredicated)
		{
			if (__ballot(alive) == 0)
				return;
		}
		else if (!alive)
			return;
	}

	// The primitive's other records, read in one place: every interpolant is computed from them
	// here for each primitive, so their scalar loads leave together (one round trip) instead of
	// one wait at each first use further down.
	PrimitiveShadeData ta;
	load_shade_data(static_cast<const TransformedAttributes *>(a.tattr) + prim_index,
	                static_cast<const StateVector *>(a.states) + (st & STATE_INDEX_MASK), prims + prim_index, ta);
	const uint32_t sv_combiner = ta.sv.combiner, sv_blend = ta.sv.blend_mode;
	const uint32_t sv_dimx0 = ta.sv.dimx0, sv_dimx1 = ta.sv.dimx1;
	// Attribute snapping without raster snapping: attributes at the pixel's first sample,
	// depth per sample (ubershader.comp:888-898).
	float ai = fi, aj = fj;
	if (super_sampled && (st & (ST_SNAP_RASTER | ST_SNAP_ATTRIBUTE)) == ST_SNAP_ATTRIBUTE)
	{
		const int32_t bx_ = int32_t(uint32_t(hx & snap_mask) << (3u - rate));
		const int32_t by_ = int32_t(uint32_t(hy & snap_mask) << (3u - rate));
		ai = __fadd_rn(__fmul_rn(float(idot3(s.b[0], s.b[1], s.b[2], bx_, by_)), s.inv_area), s.error_i);
		aj = __fadd_rn(__fmul_rn(float(idot3(s.c[0], s.c[1], s.c[2], bx_, by_)), s.inv_area), s.error_j);
	}
	// Vertex colour: c0 + (c1 - c0) * i + (c2 - c0) * j with the per-primitive terms from setup
	// (ubershader.comp:1260-1339).
	float col[4];
	for (uint32_t ch = 0; ch < 4; ch++)
	{
		const float e1 = __fmul_rn(ta.dcol1[ch], ai);
		const float e2 = __fmul_rn(ta.dcol2[ch], aj);
		const float sum = __fadd_rn(__fadd_rn(ta.col0[ch], e1), e2);
		col[ch] = floorf(__fadd_rn(__fmul_rn(255.0f, sum), 1.0f / 256.0f));
	}
	const uint32_t p_tex = ta.tex, p_tex2 = ta.tex2, p_alpha = ta.alpha, p_fbmsk = ta.fbmsk, p_fogcol = ta.fogcol;
	const TransformedAttributes *tap = static_cast<const TransformedAttributes *>(a.tattr) + prim_index;

	// Combined colour (ubershader.comp:1260-1339).
	float comb[4] = { col[0], col[1], col[2], col[3] };
	if (sv_combiner & COMB_TME)
	{
		float tx[4];
		if (p_tex & TEX_FEEDBACK)
			feedback_texel(a, p_tex, p_tex2, color, depth, tx);
		else
		{
			float stq[3][4];
			TextureDesc td;
			load_texture_data(tap, static_cast<const TextureDesc *>(a.textures) + (p_tex & TEX_INDEX_MASK), stq, td);
			// S, T, Q at the attribute position.
			auto stqf = [&](int k) {
				return __fadd_rn(__fadd_rn(stq[0][k], __fmul_rn(stq[1][k], ai)), __fmul_rn(stq[2][k], aj));
			};
			const bool persp = (st & ST_PERSPECTIVE) != 0;
			float u = stqf(0), v = stqf(1);
			if (persp) // q == 1 divides exactly, as the reference's q != 1 test skips
			{
				const float q = stqf(2);
				u = __fdiv_rn(u, q);
				v = __fdiv_rn(v, q);
			}
			u = fminf(fmaxf(u, -2047.0f), 2047.0f);
			v = fminf(fmaxf(v, -2047.0f), 2047.0f);
			sample_texture<kArrays>(td, a.tex_images, p_tex, p_tex2, u, v, stqf(2), persp, a.phase_lut, smp,
			               super_sampled ? phase_shift(a.rate_x_log2, rate) : 0.5f, tx);
		}
		const bool tcc = (sv_combiner & COMB_TCC) != 0;
		const uint32_t mode = sv_combiner & COMB_MODE_MASK;
		if (mode == 1) // DECAL
		{
			for (int k = 0; k < 4; k++) comb[k] = tx[k];
			if (!tcc) comb[3] = col[3];
		}
		else
		{
			// MODULATE, HIGHLIGHT, HIGHLIGHT2 (ubershader.comp:918-945)
			for (int k = 0; k < 4; k++)
				comb[k] = fminf(fmaxf(floorf(__fmul_rn(__fmul_rn(col[k], tx[k]), 1.0f / 128.0f)), 0.0f), 255.0f);
			if (mode == 0)
			{
				if (!tcc) comb[3] = col[3];
			}
			else
			{
				for (int k = 0; k < 3; k++) comb[k] = __fadd_rn(comb[k], col[3]);
				comb[3] = tcc ? (mode == 2 ? __fadd_rn(tx[3], col[3]) : tx[3]) : col[3];
				for (int k = 0; k < 4; k++) comb[k] = fminf(fmaxf(comb[k], 0.0f), 255.0f);
			}
		}
	}
	if (sv_combiner & COMB_FOG)
	{
		const float fog = rintf(__fadd_rn(__fadd_rn(tap->stqf0[3], __fmul_rn(tap->stqf1[3], ai)),
		                                  __fmul_rn(tap->stqf2[3], aj)));
		for (uint32_t k = 0; k < 3; k++)
		{
			const float fc = byte_f(p_fogcol, k);
			const float d = __fmul_rn(__fsub_rn(comb[k], fc), fog);
			comb[k] = floorf(__fadd_rn(__fadd_rn(fc, __fmul_rn(d, 1.0f / 256.0f)), 0.5f));
		}
	}

	// shade_resolve (ubershader.comp:1464-1607).
	float fb[4];
	for (uint32_t k = 0; k < 4; k++)
		fb[k] = byte_f(color, k);
	bool write_fb = true, write_zb = true, write_alpha = true;
	const uint32_t bm = sv_blend;
	bool blending = bm & BLEND_ABE;
	if (multisample)
	{
		// AA1: coverage replaces alpha 0x80 (any alpha without blending), blending is forced on,
		// and partial coverage does not write depth (ubershader.comp:848-851, 1497-1505).
		comb[3] = (comb[3] == 128.0f || !blending) ? float(coverage) : comb[3];
		blending = true;
		write_zb = coverage >= 0x80u;
	}
	if (bm & BLEND_ATE)
	{
		const bool pass = alpha_test((bm >> 2) & 7u, p_alpha & 0xffu, comb[3]);
		switch ((bm >> 5) & 3u)
		{
		case 0:
			alive = alive && pass;
			if constexpr (!kPredicated)
				if (!alive)
					return;
			break;
		case 1: write_zb = write_zb && pass; break;
		case 2: write_fb = pass; write_alpha = pass; break;
		default:
			if (pi.fb32) write_alpha = pass;
			write_zb = write_zb && pass;
			break;
		}
	}
	if (!pi.fb24 && (bm & BLEND_DATE))
	{
		const uint32_t bit = fb[3] >= 128.0f ? 1u : 0u;
		alive = alive && uint32_t((bm & BLEND_DATM) != 0) == bit;
		if constexpr (!kPredicated)
			if (!alive)
				return;
	}
	float o[4] = { comb[0], comb[1], comb[2], comb[3] };
	if (blending)
	{
		// PABE turns blending off per lane (source alpha below 128): a select, not a branch.
		const bool lane_blends = !((bm & BLEND_PABE) && comb[3] < 128.0f);
		const uint32_t am = (bm >> 11) & 3u, bmd = (bm >> 13) & 3u, cm = (bm >> 15) & 3u, dm = (bm >> 17) & 3u;
		float C = 0.0f;
		if (cm == 0) C = comb[3];
		else if (cm == 1) C = pi.fb24 ? 128.0f : fb[3];
		else if (cm == 2) C = float((p_alpha >> 8) & 0xffu);
		for (int k = 0; k < 3; k++)
		{
			const float A_ = am == 0 ? comb[k] : am == 1 ? fb[k] : 0.0f;
			const float B_ = bmd == 0 ? comb[k] : bmd == 1 ? fb[k] : 0.0f;
			const float D_ = dm == 0 ? comb[k] : dm == 1 ? fb[k] : 0.0f;
			const float b = __fadd_rn(floorf(__fmul_rn(__fmul_rn(__fsub_rn(A_, B_), C), 1.0f / 128.0f)), D_);
			o[k] = lane_blends ? b : o[k];
		}
	}
	if (bm & BLEND_DTHE)
	{
		const int32_t dxp = px & 3, dyp = py & 3;
		const int32_t bit_index = ((dyp * 4 + dxp) * 4) & 31;
		const uint32_t dmx = dyp < 2 ? sv_dimx0 : sv_dimx1;
		const int32_t dv = int32_t(dmx << (29 - bit_index)) >> 29;
		for (int k = 0; k < 3; k++)
			o[k] = __fadd_rn(o[k], float(dv));
	}
	if (pi.fb24)
		write_alpha = false;
	if (!write_alpha)
		o[3] = fb[3];
	if (!write_fb)
	{
		o[0] = fb[0]; o[1] = fb[1]; o[2] = fb[2];
	}
	// Every o[k] is an integer-valued float of magnitude below 4096 (each path ends in floor,
	// rint, a byte or a sum of such). On those values the float wrap-or-clamp and pack of
	// shade_resolve, rint(clamp(fract(o / 256) * (256 / 255), 0, 1) * 255) and
	// rint(clamp(o * (1 / 255), 0, 1) * 255), equal int(o) & 255 and clamp(int(o), 0, 255)
	// (exhaustive over -4096..4096).
	uint32_t packed = 0;
	for (int k = 0; k < 4; k++)
	{
		const int32_t i = int32_t(o[k]);
		const uint32_t c = (bm & BLEND_COLCLAMP) ? uint32_t(min(max(i, 0), 255)) : uint32_t(i) & 0xffu;
		packed |= c << (8 * k);
	}
	packed |= uint32_t((bm & BLEND_FBA) && write_alpha) << 31;

	if (p_fbmsk != ~0u)
	{
		uint32_t c = p_fbmsk ? ((packed & ~p_fbmsk) | (color & p_fbmsk)) : packed;
		if (pi.fb16)
			c &= 0x80f8f8f8u;
		const bool wc = alive && (write_fb || write_alpha);
		color = wc ? c : color;
		cdirty = cdirty || wc;
	}
	if (z_write)
	{
		const bool wz = alive && write_zb;
		depth = wz ? z : depth;
		zdirty = zdirty || wz;
	}
}

__device__ __forceinline__ uint32_t waves_of(const pgs_shade_args &a) { return a.waves_per_group ? a.waves_per_group : 4u; }

// One wave shades 8x8 block (bx, by) of a single-sampled pass (or single-sampled tile).
__device__ __forceinline__ void shade_1x_block(const pgs_shade_args &a, uint32_t bx, uint32_t by)
{
	const uint32_t lane = threadIdx.x & 63u;
	if (bx >= a.blocks_x || by >= a.blocks_y)
		return;

	const int32_t blox = a.origin_x + int32_t(bx) * 8, bloy = a.origin_y + int32_t(by) * 8;
	const int32_t px = blox + int32_t(lane & 7u), py = bloy + int32_t(lane >> 3);
	const uint32_t shift = a.tile_log2 - 3u;
	const uint32_t tile = (by >> shift) * a.tiles_x + (bx >> shift);
	if (a.tile_modes && uni(a.tile_modes[tile]) != 1u)
		return;
	const uint32_t count = uni(a.counts[tile]);
	if (count == 0)
		return;
	const uint16_t *list = a.lists + size_t(tile) * a.list_stride;

	PassInfo pi;
	pi.fb_psm = a.fb_psm;
	pi.z_psm = a.z_psm;
	pi.fb16 = psm_is_16(a.fb_psm);
	pi.fb24 = a.fb_psm == PSMCT24 || a.fb_psm == PSMZ24;
	pi.fb32 = a.fb_psm == PSMCT32 || a.fb_psm == PSMZ32;
	pi.zsens = a.z_psm != 0xffffffffu;
	pi.z16 = pi.zsens && psm_is_16(a.z_psm);
	pi.z24 = pi.zsens && (a.z_psm == PSMCT24 || a.z_psm == PSMZ24);
	pi.zmax = pi.z16 ? 0xffffu : pi.z24 ? 0xffffffu : 0xffffffffu;

	// begin_tile. Colour and depth may come from copies of VRAM (a pass whose depth aliases
	// its colour buffer).
	const uint8_t *cv = a.src ? a.src : a.vram;
	const uint8_t *zv = a.zsrc ? a.zsrc : cv;
	const uint32_t caddr = swizzle_addr(uint32_t(px), uint32_t(py), a.fbp * 32u, a.fbw, a.fb_psm);
	uint32_t color;
	if (pi.fb16)
		color = rgba16_to_rgba32(*reinterpret_cast<const uint16_t *>(cv + caddr));
	else
		color = *reinterpret_cast<const uint32_t *>(cv + caddr);
	uint32_t zaddr = 0, depth = 0;
	if (pi.zsens)
	{
		zaddr = swizzle_addr(uint32_t(px), uint32_t(py), a.zbp * 32u, a.fbw, a.z_psm);
		if (pi.z16)
			depth = *reinterpret_cast<const uint16_t *>(zv + zaddr);
		else
		{
			depth = *reinterpret_cast<const uint32_t *>(zv + zaddr);
			if (pi.z24)
				depth &= 0xffffffu;
		}
	}
	bool cdirty = false, zdirty = false;
	const uint32_t prim_first = a.prim_first, prim_span = a.prim_count ? a.prim_count : ~0u;

	const PrimitiveSetup *setups = static_cast<const PrimitiveSetup *>(a.setup);
	const PrimitiveAttribute *prims = static_cast<const PrimitiveAttribute *>(a.prims);
	for (uint32_t base = 0; base < count; base += 64u)
	{
		const uint32_t li = base + lane;
		uint32_t idx = 0;
		bool hit = false;
		if (li < count)
		{
			idx = list[li];
			const uint32_t st = prims[idx].state;
			hit = idx - prim_first < prim_span &&
			      intersects_block(setups[idx], blox, bloy, blox + 7, bloy + 7, (st & ST_MULTISAMPLE) && !(st & ST_LINE));
		}
		uint64_t ballot = __ballot(hit);
		while (ballot)
		{
			const uint32_t bit = uint32_t(__builtin_ctzll(ballot));
			ballot &= ballot - 1u;
			const uint32_t prim = uint32_t(__builtin_amdgcn_readlane(int(idx), int(bit)));
			shade_pixel<true>(a, pi, prim, px, py, 0u, false, 0u, color, depth, cdirty, zdirty);
		}
	}

	// end_tile. In a super-sampled context the reference plane marks the pixel single-sampled.
	if (a.skip_stores & 1u)
		cdirty = false;
	if (a.skip_stores & 2u)
		zdirty = false;
	if (cdirty)
	{
		if (pi.fb16)
		{
			if (a.has_ref_plane)
				*reinterpret_cast<uint16_t *>(a.vram + kVramSize + caddr) = 0;
			*reinterpret_cast<uint16_t *>(a.vram + caddr) = uint16_t(rgba32_to_rgba16(color));
		}
		else
		{
			if (a.has_ref_plane)
				*reinterpret_cast<uint32_t *>(a.vram + kVramSize + caddr) = 0;
			*reinterpret_cast<uint32_t *>(a.vram + caddr) = color;
		}
	}
	if (zdirty)
	{
		if (a.has_ref_plane)
		{
			if (pi.z16)
				*reinterpret_cast<uint16_t *>(a.vram + kVramSize + zaddr) = 0;
			else
				*reinterpret_cast<uint32_t *>(a.vram + kVramSize + zaddr) = 0;
		}
		if (pi.z16)
			*reinterpret_cast<uint16_t *>(a.vram + zaddr) = uint16_t(depth);
		else if (pi.z24)
		{
			a.vram[zaddr] = uint8_t(depth);
			a.vram[zaddr + 1] = uint8_t(depth >> 8);
			a.vram[zaddr + 2] = uint8_t(depth >> 16);
		}
		else
			*reinterpret_cast<uint32_t *>(a.vram + zaddr) = depth;
	}
}

__device__ __forceinline__ void shade_1x_body(const pgs_shade_args &a, uint32_t wg, uint32_t nwg)
{
	// XCD-aware remap: workgroup w runs on XCD (w % 8); give each XCD a contiguous range.
	const uint32_t per_xcd = nwg / 8u;
	const uint32_t logical = (wg % 8u) * per_xcd + wg / 8u;
	// A workgroup of 1, 2 or 4 waves covers 1x1, 2x1 or 2x2 blocks of 8x8 pixels.
	const uint32_t waves = waves_of(a);
	const uint32_t gw = waves >= 2u ? 2u : 1u, gh = waves >= 4u ? 2u : 1u;
	const uint32_t groups_x = (a.blocks_x + gw - 1u) / gw;
	const uint32_t qx = logical % groups_x, qy = logical / groups_x;
	const uint32_t wave = threadIdx.x >> 6;
	shade_1x_block(a, qx * gw + (wave % gw), qy * gh + (wave / gw));
}

// Hi-res offset of sample s inside its pixel: the lane-bit layout of ubershader.comp:133-385
// (sample index = lane & (S - 1)), written out per sampling grid.
template <uint32_t RX, uint32_t RY>
__device__ __forceinline__ void sample_offset(uint32_t s, int32_t &dx, int32_t &dy)
{
	if (RX == 0 && RY == 1) { dx = int32_t(s); dy = int32_t(s); }
	else if (RX == 1 && RY == 1) { dx = int32_t(s >> 1); dy = int32_t(s & 1u); }
	else if (RX == 1 && RY == 2)
	{
		dx = int32_t(2u * ((s >> 1) & 1u) + (s & 1u));
		dy = int32_t((s & 1u) + 2u * ((s >> 2) & 1u));
	}
	else if (RX == 0 && RY == 2)
	{
		// Sparse 1 x 4: one sample per quarter row, at x quarters 0, 2, 3, 1 (the 0x78 offsets).
		dx = int32_t((0x78u >> (2u * s)) & 3u);
		dy = int32_t(s);
	}
	else if (RX == 1 && RY == 3)
	{
		// Sparse 2 x 8: two samples half a pixel apart per eighth row, with the same offsets.
		dx = int32_t(4u * ((s >> 3) & 1u) + ((0x78u >> (2u * (s & 3u))) & 3u));
		dy = int32_t(s & 7u);
	}
	else
	{
		dx = int32_t(((s >> 1) & 1u) + 2u * ((s >> 3) & 1u));
		dy = int32_t((s & 1u) + 2u * ((s >> 2) & 1u));
	}
}

// Super-sampled tiles: one wave64 holds the S samples of 64 / S pixels (lane = pixel * S + s),
// so the end-of-tile resolve is a butterfly over S consecutive lanes, summing in the
// reference's xor order (8, 4, 2, 1) and taking the minimum depth (ubershader.comp:581-780).
// One wave shades the 64 samples of sub-rectangle `sub` of 8x8 block (bx, by). In a launch
// with single_tiles, the first wave of a block of a single-sampled tile shades its 64 pixels at
// 1x through the same loop (the sampling rate is then a wave-uniform 0), as the single-sampled
// kernel would, and the block's other waves leave.
template <uint32_t RX, uint32_t RY, bool kArrays = false>
__device__ __forceinline__ void shade_ssaa_wave(const pgs_shade_args &a, uint32_t bx, uint32_t by, uint32_t sub)
{
	constexpr uint32_t SL = RX + RY, S = 1u << SL;
	constexpr uint32_t RW = SL <= 2 ? (SL == 1 ? 8u : 4u) : (SL == 3 ? 4u : 2u); // wave rectangle
	constexpr uint32_t RH = (64u >> SL) / RW;
	constexpr uint32_t SUB_X = 8u / RW;

	const uint32_t lane = threadIdx.x & 63u;
	if (bx >= a.blocks_x || by >= a.blocks_y)
		return;
	const uint32_t shift = a.tile_log2 - 3u;
	const uint32_t tile = (by >> shift) * a.tiles_x + (bx >> shift);
	const uint32_t mode = uni(a.tile_modes[tile]);
	const bool single = mode == 1u && a.single_tiles != 0;
	if (single ? sub != 0u : mode != 2u)
		return;
	const uint32_t count = uni(a.counts[tile]);
	if (count == 0)
		return;
	const uint16_t *list = a.lists + size_t(tile) * a.list_stride;

	int32_t rox, roy, rhx, rhy, px, py, hx, hy;
	uint32_t smp;
	if (single)
	{
		rox = a.origin_x + int32_t(bx * 8u);
		roy = a.origin_y + int32_t(by * 8u);
		rhx = rox + 7;
		rhy = roy + 7;
		px = rox + int32_t(lane & 7u);
		py = roy + int32_t(lane >> 3);
		hx = px;
		hy = py;
		smp = 0;
	}
	else
	{
		rox = a.origin_x + int32_t(bx * 8u + (sub % SUB_X) * RW);
		roy = a.origin_y + int32_t(by * 8u + (sub / SUB_X) * RH);
		rhx = rox + int32_t(RW) - 1;
		rhy = roy + int32_t(RH) - 1;
		smp = lane & (S - 1u);
		const uint32_t pix = lane >> SL;
		px = rox + int32_t(pix % RW);
		py = roy + int32_t(pix / RW);
		int32_t dx, dy;
		sample_offset<RX, RY>(smp, dx, dy);
		hx = (px << RY) + dx;
		hy = (py << RY) + dy;
	}
	const uint32_t rate = single ? 0u : RY;

	PassInfo pi;
	pi.fb_psm = a.fb_psm;
	pi.z_psm = a.z_psm;
	pi.fb16 = psm_is_16(a.fb_psm);
	pi.fb24 = a.fb_psm == PSMCT24 || a.fb_psm == PSMZ24;
	pi.fb32 = a.fb_psm == PSMCT32 || a.fb_psm == PSMZ32;
	pi.zsens = a.z_psm != 0xffffffffu;
	pi.z16 = pi.zsens && psm_is_16(a.z_psm);
	pi.z24 = pi.zsens && (a.z_psm == PSMCT24 || a.z_psm == PSMZ24);
	pi.zmax = pi.z16 ? 0xffffu : pi.z24 ? 0xffffffu : 0xffffffffu;

	// begin_tile: samples come from their planes where the reference plane marks the pixel; a
	// single-sampled tile reads plane 0 only. Colour and depth may come from copies of VRAM
	// with the same planes (a pass whose depth aliases its colour buffer).
	uint8_t *v = a.vram;
	const uint8_t *cv = a.src ? a.src : v;
	const uint8_t *zv = a.zsrc ? a.zsrc : cv;
	const size_t P = kVramSize;
	const uint32_t caddr = swizzle_addr(uint32_t(px), uint32_t(py), a.fbp * 32u, a.fbw, a.fb_psm);
	uint32_t color;
	if (pi.fb16)
	{
		uint32_t c = *reinterpret_cast<const uint16_t *>(cv + caddr);
		if (!single && *reinterpret_cast<const uint16_t *>(cv + P + caddr) == 0xffffu)
			c = *reinterpret_cast<const uint16_t *>(cv + (2u + smp) * P + caddr);
		color = rgba16_to_rgba32(c);
	}
	else
	{
		color = *reinterpret_cast<const uint32_t *>(cv + caddr);
		const uint32_t ref = single ? 0u : *reinterpret_cast<const uint32_t *>(cv + P + caddr);
		const uint32_t mask = pi.fb24 ? 0xffffffu : ~0u;
		if ((ref & mask) == mask)
		{
			const uint32_t sc = *reinterpret_cast<const uint32_t *>(cv + (2u + smp) * P + caddr);
			// A 24-bit frame keeps the single-sampled alpha unless the whole word is marked.
			color = (pi.fb24 && ref != ~0u) ? ((color & 0xff000000u) | (sc & 0xffffffu)) : sc;
		}
	}
	uint32_t zaddr = 0, depth = 0;
	if (pi.zsens)
	{
		zaddr = swizzle_addr(uint32_t(px), uint32_t(py), a.zbp * 32u, a.fbw, a.z_psm);
		if (pi.z16)
		{
			depth = *reinterpret_cast<const uint16_t *>(zv + zaddr);
			if (!single && *reinterpret_cast<const uint16_t *>(zv + P + zaddr) == 0xffffu)
				depth = *reinterpret_cast<const uint16_t *>(zv + (2u + smp) * P + zaddr);
		}
		else
		{
			const uint32_t mask = pi.z24 ? 0xffffffu : ~0u;
			depth = *reinterpret_cast<const uint32_t *>(zv + zaddr) & mask;
			if (!single && (*reinterpret_cast<const uint32_t *>(zv + P + zaddr) & mask) == mask)
				depth = *reinterpret_cast<const uint32_t *>(zv + (2u + smp) * P + zaddr) & mask;
		}
	}
	bool cdirty = false, zdirty = false;
	const uint32_t prim_first = a.prim_first, prim_span = a.prim_count ? a.prim_count : ~0u;

	const PrimitiveSetup *setups = static_cast<const PrimitiveSetup *>(a.setup);
	const PrimitiveAttribute *prims = static_cast<const PrimitiveAttribute *>(a.prims);
	for (uint32_t base = 0; base < count; base += 64u)
	{
		const uint32_t li = base + lane;
		uint32_t idx = 0;
		bool hit = false;
		if (li < count)
		{
			idx = list[li];
			const uint32_t st = prims[idx].state;
			hit = idx - prim_first < prim_span &&
			      intersects_block(setups[idx], rox, roy, rhx, rhy, (st & ST_MULTISAMPLE) && !(st & ST_LINE));
		}
		uint64_t ballot = __ballot(hit);
		while (ballot)
		{
			const uint32_t bit = uint32_t(__builtin_ctzll(ballot));
			ballot &= ballot - 1u;
			const uint32_t prim = uint32_t(__builtin_amdgcn_readlane(int(idx), int(bit)));
			shade_pixel<false, kArrays>(a, pi, prim, hx, hy, rate, !single, smp, color, depth, cdirty, zdirty);
		}
	}

	// end_tile. A single-sampled pixel is its own group of one lane: nothing differs from its
	// first sample, so it writes plane 0 and clears its reference-plane mark where dirty, as
	// the single-sampled kernel does.
	const uint32_t group = single ? lane : lane & ~(S - 1u);
	const uint64_t gmask = single ? (1ull << lane) : ((1ull << S) - 1ull) << group;
	const uint32_t first_color = uint32_t(__shfl(int(color), int(group)));
	const uint32_t first_depth = uint32_t(__shfl(int(depth), int(group)));
	const bool req_c = (__ballot(color != first_color) & gmask) != 0;
	const bool req_d = pi.zsens && (__ballot(depth != first_depth) & gmask) != 0;
	const bool cd = !(a.skip_stores & 1u) && (__ballot(cdirty) & gmask) != 0;
	const bool zd = pi.zsens && !(a.skip_stores & 2u) && (__ballot(zdirty) & gmask) != 0;

	float f[4];
	for (uint32_t k = 0; k < 4; k++)
		f[k] = unorm8(color, k);
	for (uint32_t m = S >> 1; m >= 1u; m >>= 1)
		for (uint32_t k = 0; k < 4; k++)
			f[k] = __fadd_rn(f[k], __shfl_xor(f[k], int(m)));
	for (uint32_t k = 0; k < 4; k++)
		f[k] = __fmul_rn(f[k], 1.0f / float(S)); // S is a power of two: exact, as the division
	const uint32_t resolved = pack_unorm(f[0], f[1], f[2], f[3]);
	uint32_t dmin = depth;
	for (uint32_t m = S >> 1; m >= 1u; m >>= 1)
		dmin = min(dmin, uint32_t(__shfl_xor(int(dmin), int(m))));

	if (req_c && cd)
	{
		if (pi.fb16)
			*reinterpret_cast<uint16_t *>(v + (2u + smp) * P + caddr) = uint16_t(rgba32_to_rgba16(color));
		else
			*reinterpret_cast<uint32_t *>(v + (2u + smp) * P + caddr) = color;
	}
	if (req_d && zd)
	{
		if (pi.z16)
			*reinterpret_cast<uint16_t *>(v + (2u + smp) * P + zaddr) = uint16_t(depth);
		else
			*reinterpret_cast<uint32_t *>(v + (2u + smp) * P + zaddr) = depth;
	}
	if (smp == 0)
	{
		const uint32_t out_color = req_c ? resolved : first_color;
		const uint32_t out_depth = req_d ? dmin : first_depth;
		if (cd)
		{
			if (pi.fb16)
			{
				*reinterpret_cast<uint16_t *>(v + P + caddr) = req_c ? 0xffffu : 0u;
				*reinterpret_cast<uint16_t *>(v + caddr) = uint16_t(rgba32_to_rgba16(out_color));
			}
			else
			{
				*reinterpret_cast<uint32_t *>(v + P + caddr) = req_c ? ~0u : 0u;
				*reinterpret_cast<uint32_t *>(v + caddr) = out_color;
			}
		}
		if (zd)
		{
			if (pi.z16)
			{
				*reinterpret_cast<uint16_t *>(v + P + zaddr) = req_d ? 0xffffu : 0u;
				*reinterpret_cast<uint16_t *>(v + zaddr) = uint16_t(out_depth);
			}
			else
			{
				*reinterpret_cast<uint32_t *>(v + P + zaddr) = req_d ? (pi.z24 ? 0xffffffu : ~0u) : 0u;
				if (pi.z24)
				{
					v[zaddr] = uint8_t(out_depth);
					v[zaddr + 1] = uint8_t(out_depth >> 8);
					v[zaddr + 2] = uint8_t(out_depth >> 16);
				}
				else
					*reinterpret_cast<uint32_t *>(v + zaddr) = out_depth;
			}
		}
	}
}

template <uint32_t RX, uint32_t RY, bool kArrays = false>
__device__ __forceinline__ void shade_ssaa_body(const pgs_shade_args &a, uint32_t wg, uint32_t nwg)
{
	constexpr uint32_t S = 1u << (RX + RY);
	const uint32_t per_xcd = nwg / 8u;
	const uint32_t logical = (wg % 8u) * per_xcd + wg / 8u;
	const uint32_t wave_global = logical * waves_of(a) + (threadIdx.x >> 6);
	const uint32_t block = wave_global / S, sub = wave_global % S;
	shade_ssaa_wave<RX, RY, kArrays>(a, block % a.blocks_x, block / a.blocks_x, sub);
}

__global__ void __launch_bounds__(256) k_shade_1x(pgs_shade_args a) { shade_1x_body(a, blockIdx.x, gridDim.x); }
// Held to 8 waves per SIMD (64 VGPRs, 78 SGPRs with a few spilled to VGPR lanes): faster than
// the 7 waves its unbounded 106 SGPRs allow (textured 4x: 86.5 against 91 us per launch).
// kArrays: the pass samples super-sampled textures (its own instantiation, so that passes
// without them keep the register allocation of the plain sampling path).
template <uint32_t RX, uint32_t RY, bool kArrays = false>
__global__ void __launch_bounds__(256) PGS_OCCUPANCY_8 k_shade_ssaa(pgs_shade_args a)
{
	shade_ssaa_body<RX, RY, kArrays>(a, blockIdx.x, gridDim.x);
}
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_1x_multi, shade_1x_body, pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_01_multi, (shade_ssaa_body<0, 1>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_11_multi, (shade_ssaa_body<1, 1>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_12_multi, (shade_ssaa_body<1, 2>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_22_multi, (shade_ssaa_body<2, 2>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_02_multi, (shade_ssaa_body<0, 2>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_13_multi, (shade_ssaa_body<1, 3>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_01_multi_arrays, (shade_ssaa_body<0, 1, true>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_11_multi_arrays, (shade_ssaa_body<1, 1, true>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_12_multi_arrays, (shade_ssaa_body<1, 2, true>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_22_multi_arrays, (shade_ssaa_body<2, 2, true>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_02_multi_arrays, (shade_ssaa_body<0, 2, true>), pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_ssaa_13_multi_arrays, (shade_ssaa_body<1, 3, true>), pgs_shade_args)

// Launches of different sampling grids in one multi-launch (independent passes of different
// contexts): each workgroup runs the body of its own launch's grid. One launch instead of one
// per grid leaves no partly idle tail between the grids. The bodies share one register
// allocation, so a kernel holding only the 1x, 2x and 4x bodies (kAllGrids false) spills less
// than one holding all five.
template <bool kAllGrids>
__device__ __forceinline__ void shade_any_body(const pgs_shade_args &a, uint32_t wg, uint32_t nwg)
{
	const uint32_t rx = a.rate_x_log2, ry = a.rate_log2;
	if (ry == 0)
		shade_1x_body(a, wg, nwg);
	else if (rx == 0 && ry == 1)
		shade_ssaa_body<0, 1>(a, wg, nwg);
	else if (rx == 1 && ry == 1)
		shade_ssaa_body<1, 1>(a, wg, nwg);
	else if constexpr (kAllGrids)
	{
		if (rx == 1 && ry == 2)
			shade_ssaa_body<1, 2>(a, wg, nwg);
		else if (rx == 2 && ry == 2)
			shade_ssaa_body<2, 2>(a, wg, nwg);
	}
}
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_mixed_multi, shade_any_body<true>, pgs_shade_args)
PGS_DEFINE_MULTI_ATTR(256, PGS_OCCUPANCY_8, k_shade_mixed_low_multi, shade_any_body<false>, pgs_shade_args)

// Whether a launch shading `want_mode` tiles works on a tile of `mode`: a super-sampled launch
// with single_tiles also takes the single-sampled tiles (with the same S waves per block, of
// which the first shades the block).
__device__ __forceinline__ bool tile_has_work(const pgs_shade_args &a, uint32_t mode, uint32_t want_mode)
{
	return mode == want_mode || (want_mode == 2u && mode == 1u && a.single_tiles);
}

// ---------------------------------------------------------------------------------------
// Work-queue shading by coarse tile: a grid of resident workgroups pops queue items (coarse
// tiles in the tile numbering of a multi-launch, where launch l owns tiles [prefix[l],
// prefix[l + 1]), or equal parts of a tile's waves) in the order `order` until counter[1]
// items have been taken; every wave of a workgroup sees the same item, so all leave together.
// Inside an item the waves take the tile's waves (blocks, and super-sampled each block's S
// sample waves) one at a time from an LDS counter, so the workgroup waits only for the last
// one. With `costs` set, thread 0 adds the item's shader-clock cycles to its tile's count.
template <uint32_t S, typename WaveFn>
__device__ __forceinline__ void tile_queue_loop(const pgs_queue_args &q, const pgs_shade_args *__restrict__ args,
                                                uint32_t want_mode, WaveFn fn)
{
	__shared__ uint32_t slot, next;
	// Item numbers, launch indices and wave numbers are wave-uniform: readfirstlane keeps them
	// (and every address derived from them) in scalar registers.
	const uint32_t limit = uni(q.counter[1]);
	const uint32_t lane = threadIdx.x & 63u;
	for (;;)
	{
		if (threadIdx.x == 0)
		{
			slot = atomicAdd(q.counter, 1u);
			next = 0;
		}
		__syncthreads();
		const uint32_t pos = uni(slot);
		__syncthreads();
		if (pos >= limit)
			return;
		const uint32_t e = uni(q.order[pos]);
		const uint32_t g = e >> 8, part = (e >> 4) & 15u, parts = (e & 15u) + 1u;
		const uint32_t l = uni(find_launch(q.prefix, q.num_launches, g));
		const pgs_shade_args &a = args[l];
		const uint32_t t = g - q.prefix[l];
		const uint64_t t0 = uint64_t(clock64());
		const bool work = a.tile_modes ? tile_has_work(a, a.tile_modes[t], want_mode) : a.counts[t] != 0;
		if (work)
		{
			const uint32_t B = 1u << (a.tile_log2 - 3u);
			const uint32_t tx = t % a.tiles_x, ty = t / a.tiles_x;
			const uint32_t W = B * B * S, begin = W * part / parts, end = W * (part + 1u) / parts;
			for (;;)
			{
				uint32_t k = 0;
				if (lane == 0)
					k = atomicAdd(&next, 1u);
				const uint32_t w = begin + uni(k);
				if (w >= end)
					break;
				const uint32_t block = w / S;
				fn(a, tx * B + block % B, ty * B + block / B, w % S);
			}
		}
		__syncthreads();
		if (q.costs && threadIdx.x == 0)
		{
			const uint64_t dt = uint64_t(clock64()) - t0;
			atomicAdd(&q.costs[g], dt > 0xffffffffull ? 0xffffffffu : uint32_t(dt));
		}
	}
}

// Ordered dispatch: one workgroup per chunk of waves_per_group consecutive waves of a tile,
// taken from `order` by workgroup index (the start order of the hardware dispatcher).
template <uint32_t S, typename WaveFn>
__device__ __forceinline__ void ordered_chunk(const pgs_queue_args &q, const pgs_shade_args *__restrict__ args,
                                              uint32_t want_mode, WaveFn fn)
{
	if (blockIdx.x >= uni(q.counter[1]))
		return;
	const uint32_t e = uni(q.order[blockIdx.x]);
	if (e == 0xffffffffu) // an unused place of a per-XCD order
		return;
	const uint32_t g = e >> 12, chunk = e & 0xfffu;
	const uint32_t l = uni(find_launch(q.prefix, q.num_launches, g));
	const pgs_shade_args &a = args[l];
	const uint32_t t = g - q.prefix[l];
	if (a.tile_modes ? !tile_has_work(a, a.tile_modes[t], want_mode) : a.counts[t] == 0)
		return;
	const uint32_t B = 1u << (a.tile_log2 - 3u);
	const uint32_t w = chunk * (blockDim.x >> 6) + uni(threadIdx.x >> 6);
	if (w >= B * B * S)
		return;
	const uint32_t tx = t % a.tiles_x, ty = t / a.tiles_x, block = w / S;
	fn(a, tx * B + block % B, ty * B + block / B, w % S);
}

// The argument blocks come in as a restrict kernel parameter: with it the compiler may load
// their fields through the scalar cache inside the loop, although the loop stores to VRAM.
__global__ void __launch_bounds__(256) PGS_OCCUPANCY_8 k_shade_1x_queue(pgs_queue_args q,
                                                                        const pgs_shade_args *__restrict__ args)
{
	tile_queue_loop<1>(q, args, 1u, [](const pgs_shade_args &a, uint32_t bx, uint32_t by, uint32_t) { shade_1x_block(a, bx, by); });
}
#define PGS_SSAA_QUEUE_KERNEL(NAME, RX, RY, KA)                                                         \
	__global__ void __launch_bounds__(256) PGS_OCCUPANCY_8 NAME(pgs_queue_args q,                  \
	                                                            const pgs_shade_args *__restrict__ args) \
	{                                                                                              \
		tile_queue_loop<(1u << (RX + RY))>(q, args, 2u, [](const pgs_shade_args &a, uint32_t bx, uint32_t by, \
		                                               uint32_t sub) { shade_ssaa_wave<RX, RY, KA>(a, bx, by, sub); }); \
	}
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_01_queue, 0, 1, false)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_01_queue_arrays, 0, 1, true)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_11_queue, 1, 1, false)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_11_queue_arrays, 1, 1, true)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_12_queue, 1, 2, false)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_12_queue_arrays, 1, 2, true)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_22_queue, 2, 2, false)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_22_queue_arrays, 2, 2, true)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_02_queue, 0, 2, false)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_02_queue_arrays, 0, 2, true)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_13_queue, 1, 3, false)
PGS_SSAA_QUEUE_KERNEL(k_shade_ssaa_13_queue_arrays, 1, 3, true)

__global__ void __launch_bounds__(256) PGS_OCCUPANCY_8 k_shade_1x_ordered(pgs_queue_args q,
                                                                          const pgs_shade_args *__restrict__ args)
{
	ordered_chunk<1>(q, args, 1u, [](const pgs_shade_args &a, uint32_t bx, uint32_t by, uint32_t) { shade_1x_block(a, bx, by); });
}
#define PGS_SSAA_ORDERED_KERNEL(NAME, RX, RY, KA)                                                       \
	__global__ void __launch_bounds__(256) PGS_OCCUPANCY_8 NAME(pgs_queue_args q,                  \
	                                                            const pgs_shade_args *__restrict__ args) \
	{                                                                                              \
		ordered_chunk<(1u << (RX + RY))>(q, args, 2u, [](const pgs_shade_args &a, uint32_t bx, uint32_t by, \
		                                                 uint32_t sub) { shade_ssaa_wave<RX, RY, KA>(a, bx, by, sub); }); \
	}
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_01_ordered, 0, 1, false)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_01_ordered_arrays, 0, 1, true)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_11_ordered, 1, 1, false)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_11_ordered_arrays, 1, 1, true)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_12_ordered, 1, 2, false)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_12_ordered_arrays, 1, 2, true)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_22_ordered, 2, 2, false)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_22_ordered_arrays, 2, 2, true)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_02_ordered, 0, 2, false)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_02_ordered_arrays, 0, 2, true)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_13_ordered, 1, 3, false)
PGS_SSAA_ORDERED_KERNEL(k_shade_ssaa_13_ordered_arrays, 1, 3, true)

__device__ __forceinline__ float wave_sum(float v)
{
	for (int m = 32; m >= 1; m >>= 1)
		v += __shfl_xor(v, m);
	return v;
}

// Test hook (pgs_k_lod_eval): the Q-derived level of detail lod_from_q for K = k and L = l over
// `count` consecutive float bit patterns from `first`, summed into *hash as the wrapping sum of
// ((n << 32) | input) * 0x9E3779B97F4A7C15 (n as 32-bit two's complement).
__global__ void __launch_bounds__(256) k_lod_eval(uint32_t l, int32_t k, uint32_t first, uint64_t count,
                                                  unsigned long long *hash)
{
	uint64_t h = 0;
	const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
	for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
	{
		const uint32_t in = first + uint32_t(i);
		const uint32_t n = uint32_t(lod_from_q(__uint_as_float(in), k, l));
		h += ((uint64_t(n) << 32) | in) * 0x9E3779B97F4A7C15ull;
	}
	for (int m = 32; m >= 1; m >>= 1)
		h += __shfl_xor(h, m);
	if ((threadIdx.x & 63u) == 0)
		atomicAdd(hash, (unsigned long long)h);
}

// Per-tile features of a shading multi-launch's work for the tile-cost model (one wave per
// coarse tile): primitive count, covered area and the state mix of the tile's primitives.
// Tiles without work for this launch (no primitives, or the other shading mode) get all-zero
// features except the constant input. Layout: PGS_MLP_INPUTS floats per tile.
__global__ void __launch_bounds__(256) k_tile_features(pgs_features_args f)
{
	const uint32_t item = blockIdx.x * 4u + (threadIdx.x >> 6);
	const uint32_t lane = threadIdx.x & 63u;
	if (item >= f.total)
		return;
	const pgs_shade_args *args = static_cast<const pgs_shade_args *>(f.args);
	const uint32_t l = find_launch(f.prefix, f.num_launches, item);
	const pgs_shade_args &a = args[l];
	const uint32_t t = item - f.prefix[l];
	float *out = f.features + size_t(item) * PGS_MLP_INPUTS;
	const uint32_t want = a.rate_log2 ? 2u : 1u;
	const bool work = a.tile_modes ? tile_has_work(a, a.tile_modes[t], want) : a.counts[t] != 0;
	const uint32_t count = work ? a.counts[t] : 0u;
	const uint32_t B = 1u << (a.tile_log2 - 3u);
	const uint32_t tx = t % a.tiles_x, ty = t / a.tiles_x;
	const uint32_t bx1 = min((tx + 1u) * B, a.blocks_x), by1 = min((ty + 1u) * B, a.blocks_y);
	const int32_t x0 = a.origin_x + int32_t(tx * B * 8u), y0 = a.origin_y + int32_t(ty * B * 8u);
	const int32_t x1 = a.origin_x + int32_t(bx1 * 8u) - 1, y1 = a.origin_y + int32_t(by1 * 8u) - 1;
	const PrimitiveSetup *setups = static_cast<const PrimitiveSetup *>(a.setup);
	const PrimitiveAttribute *prims = static_cast<const PrimitiveAttribute *>(a.prims);
	const StateVector *states = static_cast<const StateVector *>(a.states);
	const uint16_t *list = a.lists + size_t(t) * a.list_stride;
	float acc[16] = {};
	for (uint32_t base = 0; base < count; base += 64u)
	{
		const uint32_t li = base + lane;
		if (li >= count)
			break;
		const uint32_t idx = list[li];
		const PrimitiveSetup &s = setups[idx];
		const int32_t ix0 = max(s.bb[0], x0), iy0 = max(s.bb[1], y0), ix1 = min(s.bb[2], x1), iy1 = min(s.bb[3], y1);
		if (ix1 < ix0 || iy1 < iy0)
			continue;
		const PrimitiveAttribute &pa = prims[idx];
		const uint32_t st = pa.state;
		const StateVector &sv = states[st & STATE_INDEX_MASK];
		const float bbw = float(s.bb[2] - s.bb[0] + 1), bbh = float(s.bb[3] - s.bb[1] + 1);
		acc[0] += 1.0f;
		acc[1] += float((ix1 - ix0 + 1) * (iy1 - iy0 + 1));
		acc[2] += (sv.combiner & COMB_TME) ? 1.0f : 0.0f;
		acc[3] += ((sv.combiner & COMB_TME) && (pa.tex & TEX_MAG_LINEAR)) ? 1.0f : 0.0f;
		acc[4] += (st & ST_PERSPECTIVE) ? 1.0f : 0.0f;
		acc[5] += (st & ST_Z_TEST) ? 1.0f : 0.0f;
		acc[6] += (st & ST_Z_WRITE) ? 1.0f : 0.0f;
		acc[7] += (sv.blend_mode & BLEND_ABE) ? 1.0f : 0.0f;
		acc[8] += (st & ST_SPRITE) ? 1.0f : 0.0f;
		acc[9] += (st & ST_SNAP_RASTER) ? 1.0f : 0.0f;
		acc[10] += (sv.blend_mode & BLEND_ATE) ? 1.0f : 0.0f;
		acc[11] += (sv.combiner & COMB_FOG) ? 1.0f : 0.0f;
		acc[12] += pa.fbmsk ? 1.0f : 0.0f;
		acc[13] += (ix0 <= x0 && iy0 <= y0 && ix1 >= x1 && iy1 >= y1) ? 1.0f : 0.0f;
		acc[14] += log2f(bbw * bbh);
		acc[15] += (st & ST_PARALLELOGRAM) ? 1.0f : 0.0f;
	}
	for (uint32_t k = 0; k < 16; k++)
		acc[k] = wave_sum(acc[k]);
	if (lane != 0)
		return;
	if (f.tile_waves)
		f.tile_waves[item] = work ? B * B * (1u << (a.rate_x_log2 + a.rate_log2)) : 0u;
	if (!work)
	{
		for (uint32_t k = 0; k < PGS_MLP_INPUTS - 1u; k++)
			out[k] = 0.0f;
		out[PGS_MLP_INPUTS - 1u] = 1.0f;
		return;
	}
	const float hits = acc[0], inv = hits > 0.0f ? 1.0f / hits : 0.0f;
	const float samples = float(1u << (a.rate_x_log2 + a.rate_log2));
	const float tile_px = float((x1 - x0 + 1) * (y1 - y0 + 1));
	out[0] = log2f(1.0f + float(count));
	out[1] = log2f(1.0f + hits);
	out[2] = log2f(1.0f + acc[1]);
	out[3] = log2f(1.0f + acc[1] * samples);
	for (uint32_t k = 2; k <= 13; k++)
		out[2 + k] = acc[k] * inv;
	out[16] = hits > 0.0f ? acc[14] * inv / 16.0f : 0.0f;
	out[17] = acc[15] * inv;
	out[18] = log2f(samples) / 4.0f;
	out[19] = log2f(tile_px * samples) / 16.0f;
	out[20] = psm_is_16(a.fb_psm) ? 1.0f : 0.0f;
	out[21] = a.z_psm != 0xffffffffu ? 1.0f : 0.0f;
	out[22] = log2f(1.0f + float(a.list_stride)) / 16.0f;
	out[23] = 1.0f;
}
}
}

// Workgroups of one shading launch (padded to a multiple of 8 for the XCD remap).
uint32_t pgs_grid_shade(const pgs_shade_args &a)
{
	const uint32_t waves = a.waves_per_group ? a.waves_per_group : 4u;
	uint64_t groups;
	if (a.rate_log2 == 0)
	{
		const uint32_t gw = waves >= 2u ? 2u : 1u, gh = waves >= 4u ? 2u : 1u;
		groups = uint64_t((a.blocks_x + gw - 1u) / gw) * ((a.blocks_y + gh - 1u) / gh);
	}
	else
	{
		const uint64_t total_waves = uint64_t(a.blocks_x) * a.blocks_y << (a.rate_x_log2 + a.rate_log2);
		groups = (total_waves + waves - 1u) / waves;
	}
	return groups ? uint32_t((groups + 7u) & ~uint64_t(7)) : 0u;
}

static bool shade_args_valid(const pgs_shade_args *args)
{
	if (!args || !args->vram || !args->setup || !args->tattr || !args->prims || !args->states || !args->lists ||
	    !args->counts)
		return false;
	if (args->tile_log2 < 3 || args->tile_log2 > 7)
		return false;
	if (args->waves_per_group != 0 && args->waves_per_group != 1 && args->waves_per_group != 2 &&
	    args->waves_per_group != 4)
		return false;
	if (args->rate_log2 != 0 && !args->tile_modes)
		return false;
	if (args->single_tiles > 1u || (args->single_tiles && args->rate_log2 == 0))
		return false;
	if (args->skip_stores > 3u)
		return false;
	return true;
}

// A super-sampled pass that samples super-sampled textures.
static int launch_shade_arrays(const pgs_shade_args &a, uint32_t rx, uint32_t ry, uint32_t groups, uint32_t threads,
                               hipStream_t stream)
{
	if (rx == 0 && ry == 1)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<0, 1, true>), dim3(groups), dim3(threads), 0, stream, a);
	else if (rx == 1 && ry == 1)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<1, 1, true>), dim3(groups), dim3(threads), 0, stream, a);
	else if (rx == 1 && ry == 2)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<1, 2, true>), dim3(groups), dim3(threads), 0, stream, a);
	else if (rx == 2 && ry == 2)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<2, 2, true>), dim3(groups), dim3(threads), 0, stream, a);
	else if (rx == 0 && ry == 2)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<0, 2, true>), dim3(groups), dim3(threads), 0, stream, a);
	else if (rx == 1 && ry == 3)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<1, 3, true>), dim3(groups), dim3(threads), 0, stream, a);
	else
		return -int(hipErrorNotSupported);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

extern "C" int pgs_k_shade(const pgs_shade_args *args, hipStream_t stream)
{
	if (!shade_args_valid(args))
		return -int(hipErrorInvalidValue);
	const uint32_t groups = pgs_grid_shade(*args);
	if (groups == 0)
		return 0;
	const uint32_t threads = 64u * (args->waves_per_group ? args->waves_per_group : 4u);
	const uint32_t rx = args->rate_x_log2, ry = args->rate_log2;
	if (ry != 0 && args->phase_lut)
		return launch_shade_arrays(*args, rx, ry, groups, threads, stream);
	if (ry == 0)
		hipLaunchKernelGGL(pgs::k_shade_1x, dim3(groups), dim3(threads), 0, stream, *args);
	else if (rx == 0 && ry == 1)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<0, 1>), dim3(groups), dim3(threads), 0, stream, *args);
	else if (rx == 1 && ry == 1)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<1, 1>), dim3(groups), dim3(threads), 0, stream, *args);
	else if (rx == 1 && ry == 2)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<1, 2>), dim3(groups), dim3(threads), 0, stream, *args);
	else if (rx == 2 && ry == 2)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<2, 2>), dim3(groups), dim3(threads), 0, stream, *args);
	else if (rx == 0 && ry == 2)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<0, 2>), dim3(groups), dim3(threads), 0, stream, *args);
	else if (rx == 1 && ry == 3)
		hipLaunchKernelGGL((pgs::k_shade_ssaa<1, 3>), dim3(groups), dim3(threads), 0, stream, *args);
	else
		return -int(hipErrorNotSupported);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

// Every launch of a multi-launch shares rates and waves per group (the caller groups them).
int pgs_multi_shade(const pgs_shade_args *args, const uint32_t *prefix, uint32_t n, uint32_t total, uint32_t rate_x,
                    uint32_t rate_y, uint32_t waves, hipStream_t stream)
{
	const uint32_t threads = 64u * (waves ? waves : 4u);
	const bool arrays = (rate_x & kRateArrays) != 0 && rate_y != 0;
	rate_x &= ~kRateArrays;
#define PGS_MULTI(K) hipLaunchKernelGGL(pgs::K, dim3(total), dim3(threads), 0, stream, args, prefix, n)
	if (arrays)
	{
		if (rate_x == 0 && rate_y == 1) PGS_MULTI(k_shade_ssaa_01_multi_arrays);
		else if (rate_x == 1 && rate_y == 1) PGS_MULTI(k_shade_ssaa_11_multi_arrays);
		else if (rate_x == 1 && rate_y == 2) PGS_MULTI(k_shade_ssaa_12_multi_arrays);
		else if (rate_x == 2 && rate_y == 2) PGS_MULTI(k_shade_ssaa_22_multi_arrays);
		else if (rate_x == 0 && rate_y == 2) PGS_MULTI(k_shade_ssaa_02_multi_arrays);
		else if (rate_x == 1 && rate_y == 3) PGS_MULTI(k_shade_ssaa_13_multi_arrays);
		else return -int(hipErrorNotSupported);
		return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
	}
#undef PGS_MULTI
	if (rate_y == 0)
		hipLaunchKernelGGL(pgs::k_shade_1x_multi, dim3(total), dim3(threads), 0, stream, args, prefix, n);
	else if (rate_x == 0 && rate_y == 1)
		hipLaunchKernelGGL(pgs::k_shade_ssaa_01_multi, dim3(total), dim3(threads), 0, stream, args, prefix, n);
	else if (rate_x == 1 && rate_y == 1)
		hipLaunchKernelGGL(pgs::k_shade_ssaa_11_multi, dim3(total), dim3(threads), 0, stream, args, prefix, n);
	else if (rate_x == 1 && rate_y == 2)
		hipLaunchKernelGGL(pgs::k_shade_ssaa_12_multi, dim3(total), dim3(threads), 0, stream, args, prefix, n);
	else if (rate_x == 2 && rate_y == 2)
		hipLaunchKernelGGL(pgs::k_shade_ssaa_22_multi, dim3(total), dim3(threads), 0, stream, args, prefix, n);
	else if (rate_x == 0 && rate_y == 2)
		hipLaunchKernelGGL(pgs::k_shade_ssaa_02_multi, dim3(total), dim3(threads), 0, stream, args, prefix, n);
	else if (rate_x == 1 && rate_y == 3)
		hipLaunchKernelGGL(pgs::k_shade_ssaa_13_multi, dim3(total), dim3(threads), 0, stream, args, prefix, n);
	else
		return -int(hipErrorNotSupported);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

// Launches of any supported sampling grids; every launch shares the waves per group.
int pgs_multi_shade_mixed(const pgs_shade_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                          uint32_t waves, bool all_grids, hipStream_t stream)
{
	const uint32_t threads = 64u * (waves ? waves : 4u);
	if (all_grids)
		hipLaunchKernelGGL(pgs::k_shade_mixed_multi, dim3(total), dim3(threads), 0, stream, args, prefix, n);
	else
		hipLaunchKernelGGL(pgs::k_shade_mixed_low_multi, dim3(total), dim3(threads), 0, stream, args, prefix, n);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

// Grid of resident workgroups for the work-queue shading kernels.
static uint32_t queue_grid(const void *kernel, uint32_t threads)
{
	int dev = 0, cus = 0, per_cu = 0;
	(void)hipGetDevice(&dev);
	(void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
	(void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, int(threads), 0);
	return uint32_t(std::max(1, cus) * std::max(1, per_cu));
}

static const void *queue_kernel(uint32_t rate_x, uint32_t rate_y)
{
	if ((rate_x & kRateArrays) && rate_y != 0)
	{
		rate_x &= ~kRateArrays;
		if (rate_x == 0 && rate_y == 1) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_01_queue_arrays);
		if (rate_x == 1 && rate_y == 1) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_11_queue_arrays);
		if (rate_x == 1 && rate_y == 2) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_12_queue_arrays);
		if (rate_x == 2 && rate_y == 2) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_22_queue_arrays);
		if (rate_x == 0 && rate_y == 2) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_02_queue_arrays);
		if (rate_x == 1 && rate_y == 3) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_13_queue_arrays);
		return nullptr;
	}
	rate_x &= ~kRateArrays;
	if (rate_y == 0) return reinterpret_cast<const void *>(pgs::k_shade_1x_queue);
	if (rate_x == 0 && rate_y == 1) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_01_queue);
	if (rate_x == 1 && rate_y == 1) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_11_queue);
	if (rate_x == 1 && rate_y == 2) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_12_queue);
	if (rate_x == 2 && rate_y == 2) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_22_queue);
	if (rate_x == 0 && rate_y == 2) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_02_queue);
	if (rate_x == 1 && rate_y == 3) return reinterpret_cast<const void *>(pgs::k_shade_ssaa_13_queue);
	return nullptr;
}

uint32_t pgs_queue_slots(uint32_t rate_x, uint32_t rate_y, uint32_t waves_per_group, uint32_t total)
{
	const void *k = queue_kernel(rate_x, rate_y);
	const uint32_t waves = waves_per_group ? waves_per_group : 4u;
	if (!k || (waves != 1 && waves != 2 && waves != 4))
		return 0;
	return std::min(total, queue_grid(k, 64u * waves));
}

extern "C" int pgs_k_shade_queue(const pgs_queue_args *q, hipStream_t stream)
{
	if (!q || !q->args || !q->prefix || !q->counter || !q->order || q->num_launches == 0)
		return -int(hipErrorInvalidValue);
	if (q->total == 0)
		return 0;
	const uint32_t waves = q->waves_per_group ? q->waves_per_group : 4u;
	if (waves != 1 && waves != 2 && waves != 4)
		return -int(hipErrorInvalidValue);
	const void *k = queue_kernel(q->rate_x, q->rate_y);
	if (!k)
		return -int(hipErrorNotSupported);
	const uint32_t grid = pgs_queue_slots(q->rate_x, q->rate_y, waves, q->total);
	const pgs_shade_args *shade_args = static_cast<const pgs_shade_args *>(q->args);
	void *params[] = { const_cast<pgs_queue_args *>(q), &shade_args };
	if (hipLaunchKernel(k, dim3(grid), dim3(64u * waves), params, 0, stream) != hipSuccess)
		return -int(hipErrorLaunchFailure);
	return 0;
}

extern "C" int pgs_k_shade_ordered(const pgs_queue_args *q, uint32_t grid, hipStream_t stream)
{
	if (!q || !q->args || !q->prefix || !q->counter || !q->order || q->num_launches == 0)
		return -int(hipErrorInvalidValue);
	if (grid == 0)
		return 0;
	const uint32_t waves = q->waves_per_group ? q->waves_per_group : 4u;
	if (waves != 1 && waves != 2 && waves != 4)
		return -int(hipErrorInvalidValue);
	const void *k = nullptr;
	const uint32_t rx = q->rate_x & ~kRateArrays;
	if ((q->rate_x & kRateArrays) && q->rate_y != 0)
	{
		if (rx == 0 && q->rate_y == 1) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_01_ordered_arrays);
		else if (rx == 1 && q->rate_y == 1) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_11_ordered_arrays);
		else if (rx == 1 && q->rate_y == 2) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_12_ordered_arrays);
		else if (rx == 2 && q->rate_y == 2) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_22_ordered_arrays);
		else if (rx == 0 && q->rate_y == 2) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_02_ordered_arrays);
		else if (rx == 1 && q->rate_y == 3) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_13_ordered_arrays);
		else return -int(hipErrorNotSupported);
	}
	else if (q->rate_y == 0) k = reinterpret_cast<const void *>(pgs::k_shade_1x_ordered);
	else if (rx == 0 && q->rate_y == 1) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_01_ordered);
	else if (rx == 1 && q->rate_y == 1) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_11_ordered);
	else if (rx == 1 && q->rate_y == 2) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_12_ordered);
	else if (rx == 2 && q->rate_y == 2) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_22_ordered);
	else if (rx == 0 && q->rate_y == 2) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_02_ordered);
	else if (rx == 1 && q->rate_y == 3) k = reinterpret_cast<const void *>(pgs::k_shade_ssaa_13_ordered);
	else return -int(hipErrorNotSupported);
	const pgs_shade_args *shade_args = static_cast<const pgs_shade_args *>(q->args);
	void *params[] = { const_cast<pgs_queue_args *>(q), &shade_args };
	if (hipLaunchKernel(k, dim3(grid), dim3(64u * waves), params, 0, stream) != hipSuccess)
		return -int(hipErrorLaunchFailure);
	return 0;
}

extern "C" int pgs_k_tile_features(const pgs_features_args *f, hipStream_t stream)
{
	if (!f || !f->args || !f->prefix || !f->features)
		return -int(hipErrorInvalidValue);
	if (f->total == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_tile_features, dim3((f->total + 3u) / 4u), dim3(256), 0, stream, *f);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

extern "C" int pgs_k_lod_eval(uint32_t l, int32_t k, uint32_t first, uint64_t count, uint64_t *hash, hipStream_t stream)
{
	if (!hash || l > 3u || k < -2048 || k > 2047 || count > (uint64_t(1) << 32))
		return -int(hipErrorInvalidValue);
	if (count == 0)
		return 0;
	const uint64_t blocks = std::min<uint64_t>(4096, (count + 255u) / 256u);
	hipLaunchKernelGGL(pgs::k_lod_eval, dim3(uint32_t(blocks)), dim3(256), 0, stream, l, k, first, count,
	                   reinterpret_cast<unsigned long long *>(hash));
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}
