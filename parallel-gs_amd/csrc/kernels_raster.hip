// Raster kernels for gfx950: triangle setup, coarse binning, shading.
//
// Replaces triangle_setup.comp, binning.comp and ubershader.comp. Integer edge setup and all
// float expressions follow the reference formulas operation by operation; the translation unit
// is compiled with -ffp-contract=off so every float operation rounds as written (the reference
// marks the setup and barycentric expressions `precise`; the remaining expressions are
// evaluated without contraction as well).
#include "gs_defs.hpp"
#include "gs_raster.hpp"
#include "pgs_kernels.h"
#include "multi_launch.hpp"
#include "multi_dispatch.hpp"
#include "transfer_bodies.hpp"
#include "texture_bodies.hpp"

namespace pgs
{
// Seed table of the fixed-point reciprocal and the float reciprocal LUT, uploaded at init
// (pgs_raster_init_tables).
__constant__ uint8_t c_rcp_seed[256];
__constant__ float c_rcp_float_lut[256];

namespace
{
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int32_t uni(int32_t v) { return int32_t(__builtin_amdgcn_readfirstlane(uint32_t(v))); }
__device__ __forceinline__ float unif(float v) { return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v))); }

__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return int32_t(uint32_t(a) + uint32_t(b)); }
// Edge function at a pixel corner, wrapping in 32 bits. The 24-bit multiply gives the 32-bit
// product's low bits because both operands fit in signed 24 bits: edge slopes are differences of
// vertex coordinates (16-bit fields minus a 16-bit offset, below 2^18 in magnitude) and corners
// are pixel positions (below 2^11) times 8.
__device__ __forceinline__ int32_t idot3(int32_t e0, int32_t e1, int32_t e2, int32_t x, int32_t y)
{
	return wadd(wadd(__mul24(e0, x), __mul24(e1, y)), e2);
}

// math_utils.h:27-61
__device__ float rcp_fixed(uint32_t b)
{
	uint32_t lz = __clz(b);
	uint32_t nb = b << lz;
	uint32_t r = uint32_t(c_rcp_seed[(nb >> 23) & 0xffu]) | 0x100u;
	uint32_t f = (r * r) << 14;
	uint32_t p = __umulhi(f, nb);
	r = (r << 23) - p;
	r <<= 1;
	p = __umulhi(r, nb);
	r = __umulhi(r, 0u - p);
	r = (r + 64u) >> 7;
	r -= 1u << 23;
	r += uint32_t(127 - 32 + int32_t(lz)) << 23;
	return __uint_as_float(r);
}

// math_utils.h:65-101
__device__ float rcp_float(float v)
{
	uint32_t u = __float_as_uint(v);
	int32_t e = int32_t((u >> 23) & 0xffu);
	if (e == 0)
		return __uint_as_float(0x7f800000u);
	float x = c_rcp_float_lut[(u >> 15) & 0xffu];
	float a = __uint_as_float((u & 0x7fffffu) | (127u << 23));
	for (int it = 0; it < 4; it++)
	{
		float x2 = __fmul_rn(x, 2.0f);
		float sq = __fmul_rn(x, x);
		float sqa = __fmul_rn(sq, a);
		x = __fsub_rn(x2, sqa);
	}
	return ldexpf(x, 127 - e);
}

// Test hook (pgs_k_rcp_eval): evaluates rcp_fixed (mode 0) or rcp_float (mode 1, inputs are float
// bits) over `count` consecutive inputs from `first` and adds the wrapping sum of
// ((output bits << 32) | input) * 0x9E3779B97F4A7C15 to *hash.
__global__ void __launch_bounds__(256) k_rcp_eval(uint32_t mode, uint32_t first, uint64_t count,
                                                  unsigned long long *hash)
{
	uint64_t h = 0;
	const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
	for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < count; k += stride)
	{
		const uint32_t in = first + uint32_t(k);
		const float r = mode ? rcp_float(__uint_as_float(in)) : rcp_fixed(in);
		h += ((uint64_t(__float_as_uint(r)) << 32) | in) * 0x9E3779B97F4A7C15ull;
	}
	for (int m = 32; m >= 1; m >>= 1)
		h += __shfl_xor(h, m);
	if ((threadIdx.x & 63u) == 0)
		atomicAdd(hash, (unsigned long long)h);
}

__device__ __forceinline__ bool tie_break(int32_t ex, int32_t ey) { return ey > 0 || (ey == 0 && ex < 0); }
__device__ __forceinline__ bool order_less(int32_t ax, int32_t ay, int32_t bx, int32_t by, uint32_t az, uint32_t bz)
{
	if (bz != az)
		return az < bz;
	if (by != ay)
		return ay < by;
	return ax < bx;
}
__device__ __forceinline__ int32_t quantize64(int64_t c) { return int32_t(uint32_t(uint64_t(c) >> 1)); }

template <typename T> __device__ __forceinline__ void swp(T &a, T &b) { T t = a; a = b; b = t; }

// 8 x the mean sample offset of a sampling grid, in pixels (utils.h:108-125
// get_average_sampling_offset, the same on both axes): 2 at 2x and ordered 4x, 3 at ordered
// 16x, 8x and sparse 4x, 3.5 at sparse 16x.
__device__ __forceinline__ float mean_sample_offset8(uint32_t rate_x, uint32_t rate_y)
{
	if (rate_x == rate_y)
		return 4.0f * float((1u << rate_y) - 1u) / float(1u << rate_y);
	if (rate_y == 1u)
		return 2.0f;
	if (rate_y == 2u)
		return 3.0f;
	return 3.5f;
}

// ---------------------------------------------------------------------------------------
// Triangle setup (gs/shaders/triangle_setup.comp:127-873), one thread per primitive.
__device__ __forceinline__ void triangle_setup_one(const pgs_setup_args &args, uint32_t index)
{
	const VertexPosition *pos = static_cast<const VertexPosition *>(args.positions) + 3u * index;
	const VertexAttribute *attr = static_cast<const VertexAttribute *>(args.attributes) + 3u * index;
	PrimitiveAttribute &prim = static_cast<PrimitiveAttribute *>(args.prims)[index];
	PrimitiveSetup &out = static_cast<PrimitiveSetup *>(args.setup)[index];
	TransformedAttributes &ta = static_cast<TransformedAttributes *>(args.tattr)[index];
	const uint32_t rate = args.rate_log2;

	uint32_t state = prim.state;
	uint32_t tex = prim.tex;
	const uint32_t tex2 = prim.tex2;
	const bool parallelogram = state & ST_PARALLELOGRAM, perspective = state & ST_PERSPECTIVE;
	const bool iip = state & ST_IIP, fix = state & ST_FIX, sprite = state & ST_SPRITE, line = state & ST_LINE;
	const bool multisample = state & ST_MULTISAMPLE;
	const uint32_t provoking = (state >> ST_PROVOKING_SHIFT) & 3u;

	int32_t o0 = 0, o1 = 1, o2 = 2;
	int32_t ax = pos[0].x, ay = pos[0].y, bx = pos[1].x, by = pos[1].y, cx = pos[2].x, cy = pos[2].y;
	uint32_t za = pos[0].z, zb = pos[1].z, zc = pos[2].z;

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
			swp(o1, o2); swp(bx, cx); swp(by, cy); swp(zb, zc);
		}
		if (!order_less(ax, ay, bx, by, za, zb))
		{
			swp(o0, o1); swp(ax, bx); swp(ay, by); swp(za, zb);
		}
	}
	else if (provoking == 1)
	{
		swp(o0, o1); swp(ax, bx); swp(ay, by); swp(za, zb);
	}
	else if (provoking == 2)
	{
		swp(o0, o2); swp(ax, cx); swp(ay, cy); swp(za, zc);
	}

	if (sprite)
	{
		cx = ax; cy = by;
		by = ay;
	}
	else if (line)
	{
		int32_t dx = bx - ax, dy = by - ay;
		if (abs(dx) >= abs(dy))
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
	const int64_t sa = int64_t(aby) * acx - int64_t(abx) * acy;
	if (sa == 0)
	{
		out.bb[0] = 0; out.bb[1] = 0; out.bb[2] = -1; out.bb[3] = -1;
		// Degenerate primitives must not count as 3D geometry for the heuristic.
		if (rate)
			prim.state = state | ST_SNAP_ATTRIBUTE;
		if (args.state_copy)
			args.state_copy[index] = prim.state;
		return;
	}

	uint32_t area;
	if (sa < 0)
	{
		area = 0u - uint32_t(uint64_t(sa));
		swp(abx, acx); swp(aby, acy); swp(bx, cx); swp(by, cy); swp(zb, zc); swp(o1, o2);
	}
	else
		area = uint32_t(uint64_t(sa));

	const int32_t bcx = cx - bx, bcy = cy - by, cax = -acx, cay = -acy;
	const int32_t dcdx = aby, dcdy = -abx, dadx = bcy, dady = -bcx, dbdx = cay, dbdy = -cax;
	int64_t kc = int64_t(-ax) * dcdx - int64_t(ay) * dcdy;
	int64_t ka = int64_t(-bx) * dadx - int64_t(by) * dady;
	int64_t kb = int64_t(-cx) * dbdx - int64_t(cy) * dbdy;
	const bool ta_ = tie_break(bcx, bcy), tb_ = tie_break(cax, cay), tc_ = tie_break(abx, aby);
	int32_t ea = !ta_, eb = !tb_, ec = !tc_;
	ka -= ea; kb -= eb; kc -= ec;
	ea += int32_t(uint32_t(uint64_t(ka)) & 1u);
	eb += int32_t(uint32_t(uint64_t(kb)) & 1u);
	ec += int32_t(uint32_t(uint64_t(kc)) & 1u);

	int32_t flip_b = 0, flip_c = 0;
	if (parallelogram)
	{
		int32_t off = 1 - int32_t(tb_) - int32_t(tc_);
		flip_b = (off + eb + ea + 2) >> 1;
		flip_c = (off + ec + ea + 2) >> 1;
	}

	const float inv_area = rcp_fixed(area);
	const float inv_area_scaled = __fmul_rn(inv_area, 2.0f);
	float error_i = __fmul_rn(float(eb), inv_area);
	float error_j = __fmul_rn(float(ec), inv_area);

	if (sprite) { zb = za; zc = za; }
	else if (line) zc = za;
	if (fix || sprite) { o0 = 0; o1 = 1; o2 = 2; }

	const VertexAttribute a0 = attr[o0], a1 = attr[o1];
	const VertexAttribute a2 = attr[(!sprite && !line) ? o2 : 0];
	float q0[4], q1[4], q2[4];
	auto load = [&](const VertexAttribute &v, float *o) {
		if (perspective) { o[0] = v.s; o[1] = v.t; }
		else { o[0] = __fdiv_rn(float(v.uv & 0xffffu), 16.0f); o[1] = __fdiv_rn(float(v.uv >> 16), 16.0f); }
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
	{
		for (int k = 0; k < 4; k++) q2[k] = q0[k];
	}

	if (perspective)
	{
		if (q0[2] < 0.0f)
		{
			q0[0] = -q0[0]; q0[1] = -q0[1]; q1[0] = -q1[0]; q1[1] = -q1[1]; q2[0] = -q2[0]; q2[1] = -q2[1];
		}
		q0[2] = fabsf(q0[2]); q1[2] = fabsf(q1[2]); q2[2] = fabsf(q2[2]);
		const uint32_t qexp = (__float_as_uint(fmaxf(fmaxf(q0[2], q1[2]), q2[2])) >> 23) & 0xffu;
		q0[2] = __uint_as_float(__float_as_uint(q0[2]) & ~0xffu);
		q1[2] = __uint_as_float(__float_as_uint(q1[2]) & ~0xffu);
		q2[2] = __uint_as_float(__float_as_uint(q2[2]) & ~0xffu);
		const uint32_t sexp = (__float_as_uint(fmaxf(fmaxf(fabsf(q0[0]), fabsf(q1[0])), fabsf(q2[0]))) >> 23) & 0xffu;
		const uint32_t texp = (__float_as_uint(fmaxf(fmaxf(fabsf(q0[1]), fabsf(q1[1])), fabsf(q2[1]))) >> 23) & 0xffu;
		const uint32_t schop = min(9u + max(sexp, qexp) - sexp, 23u);
		const uint32_t tchop = min(9u + max(texp, qexp) - texp, 23u);
		const uint32_t smask = ~((1u << schop) - 1u), tmask = ~((1u << tchop) - 1u);
		q0[0] = __uint_as_float(__float_as_uint(q0[0]) & smask); q0[1] = __uint_as_float(__float_as_uint(q0[1]) & tmask);
		q1[0] = __uint_as_float(__float_as_uint(q1[0]) & smask); q1[1] = __uint_as_float(__float_as_uint(q1[1]) & tmask);
		q2[0] = __uint_as_float(__float_as_uint(q2[0]) & smask); q2[1] = __uint_as_float(__float_as_uint(q2[1]) & tmask);
		const bool needs_lod = (tex2 & TEX2_FIXED_LOD) == 0;
		if (!needs_lod && q0[2] == q1[2] && q0[2] == q2[2])
		{
			const float inv_q = rcp_float(q0[2]);
			q0[0] = __fmul_rn(q0[0], inv_q); q0[1] = __fmul_rn(q0[1], inv_q); q0[2] = 1.0f;
			q1[0] = __fmul_rn(q1[0], inv_q); q1[1] = __fmul_rn(q1[1], inv_q); q1[2] = 1.0f;
			q2[0] = __fmul_rn(q2[0], inv_q); q2[1] = __fmul_rn(q2[1], inv_q); q2[2] = 1.0f;
		}
	}

	float d1[4], d2[4];
	for (int k = 0; k < 4; k++)
	{
		d1[k] = fix ? 0.0f : __fsub_rn(q1[k], q0[k]);
		d2[k] = fix ? 0.0f : __fsub_rn(q2[k], q0[k]);
		ta.stqf0[k] = q0[k];
		ta.stqf1[k] = d1[k];
		ta.stqf2[k] = d2[k];
	}

	if (iip && !sprite)
	{
		uint32_t r0 = a0.rgba, r1 = a1.rgba, r2 = a2.rgba;
		if (fix) r1 = r0;
		if (fix || line) r2 = r0;
		ta.rgba[0] = r0; ta.rgba[1] = r1; ta.rgba[2] = r2;
	}
	else
	{
		const uint32_t flat = attr[0].rgba;
		ta.rgba[0] = flat; ta.rgba[1] = flat; ta.rgba[2] = flat;
	}
	ta.pad = 0;
	for (uint32_t ch = 0; ch < 4; ch++)
	{
		auto unorm = [&](uint32_t c) { return __fdiv_rn(float((c >> (8u * ch)) & 0xffu), 255.0f); };
		const float c0 = unorm(ta.rgba[0]);
		ta.col0[ch] = c0;
		ta.dcol1[ch] = __fsub_rn(unorm(ta.rgba[1]), c0);
		ta.dcol2[ch] = __fsub_rn(unorm(ta.rgba[2]), c0);
	}

	const int32_t pa0 = int32_t((uint32_t(dadx) << 6) | uint32_t(parallelogram) | (uint32_t(flip_b & 3) << 2) |
	                            (uint32_t(flip_c & 3) << 4));

	const int32_t dx = parallelogram ? bx + cx - ax : cx;
	const int32_t dy = parallelogram ? by + cy - ay : cy;
	int32_t di = 0, dj = 0;
	if (sprite)
	{
		const int32_t xd = dx - ax, yd = dy - ay;
		if (xd > 0 && (xd & (xd - 1)) != 0) { di += dbdx; dj += dcdx; }
		if (yd > 0 && (yd & (yd - 1)) != 0) { di += dbdy; dj += dcdy; }
	}
	else if (parallelogram && !perspective)
	{
		const bool nearest = (tex & TEX_MAG_LINEAR) == 0;
		const bool npot = (area & (area - 1u)) != 0;
		if (nearest && npot)
		{
			const int32_t u0 = int32_t(a0.uv & 0xffffu) - ax, v0 = int32_t(a0.uv >> 16) - ay;
			const int32_t u1 = int32_t(a1.uv & 0xffffu) - bx, v1 = int32_t(a1.uv >> 16) - by;
			const int32_t u2 = int32_t(a2.uv & 0xffffu) - cx, v2 = int32_t(a2.uv >> 16) - cy;
			if (u0 == u1 && v0 == v1 && u0 == u2 && v0 == v2)
			{
				const bool mis_x = (min(ax, dx) & 15) != 0, mis_y = (min(ay, dy) & 15) != 0;
				const float dudx = __fadd_rn(__fmul_rn(d1[0], float(dbdx)), __fmul_rn(d2[0], float(dcdx)));
				const float dvdx = __fadd_rn(__fmul_rn(d1[1], float(dbdx)), __fmul_rn(d2[1], float(dcdx)));
				const float dudy = __fadd_rn(__fmul_rn(d1[0], float(dbdy)), __fmul_rn(d2[0], float(dcdy)));
				const float dvdy = __fadd_rn(__fmul_rn(d1[1], float(dbdy)), __fmul_rn(d2[1], float(dcdy)));
				const bool adj_x = mis_x && ((dudx > 0.0f && dvdx == 0.0f) || (dvdx > 0.0f && dudx == 0.0f));
				const bool adj_y = mis_y && ((dudy > 0.0f && dvdy == 0.0f) || (dvdy > 0.0f && dudy == 0.0f));
				if (adj_x) { di -= dbdx; dj -= dcdx; }
				if (adj_y) { di -= dbdy; dj -= dcdy; }
			}
		}
	}
	const float k8 = __fmul_rn(0.125f, inv_area_scaled);
	error_i = __fadd_rn(error_i, __fmul_rn(float(di), k8));
	error_j = __fadd_rn(error_j, __fmul_rn(float(dj), k8));

	PrimitiveSetup s;
	s.a[0] = pa0; s.a[1] = dady; s.a[2] = quantize64(ka);
	s.b[0] = dbdx; s.b[1] = dbdy; s.b[2] = quantize64(kb);
	s.c[0] = dcdx; s.c[1] = dcdy; s.c[2] = quantize64(kc);
	s.inv_area = inv_area_scaled;
	s.error_i = error_i;
	s.error_j = error_j;
	s.z[0] = za;
	s.z[1] = __float_as_uint(__uint2float_rn(zb - za));
	s.z[2] = __float_as_uint(__uint2float_rn(zc - za));
	s.z[3] = 0;

	int32_t lox = min(min(ax, bx), min(cx, dx)), loy = min(min(ay, by), min(cy, dy));
	int32_t hix = max(max(ax, bx), max(cx, dx)), hiy = max(max(ay, by), max(cy, dy));
	lox = (lox + (15 >> rate)) >> 4;
	loy = (loy + (15 >> rate)) >> 4;
	hix = (hix - 1) >> 4;
	hiy = (hiy - 1) >> 4;
	if (multisample && !line) { lox--; loy--; hix++; hiy++; }
	s.bb[0] = max(lox, int32_t(prim.bb[0]));
	s.bb[1] = max(loy, int32_t(prim.bb[1]));
	s.bb[2] = min(hix, int32_t(prim.bb[2]));
	s.bb[3] = min(hiy, int32_t(prim.bb[3]));

	if (sprite && !perspective)
	{
		const int32_t u0 = int32_t(a0.uv & 0xffffu) - ax, v0 = int32_t(a0.uv >> 16) - ay;
		const int32_t u1 = int32_t(a1.uv & 0xffffu) - dx, v1 = int32_t(a1.uv >> 16) - dy;
		const int32_t minu = min(u0, u1), minv = min(v0, v1), maxu = max(u0, u1), maxv = max(v0, v1);
		const bool cs_u = (minu & 15) == 8, cs_v = (minv & 15) == 8;
		const int32_t distx = abs(dx - ax), disty = abs(dy - ay);
		const bool ns_u = (maxu & 15) == 9 && (distx & (distx - 1)) != 0;
		const bool ns_v = (maxv & 15) == 9 && (disty & (disty - 1)) != 0;
		const bool no_interp = (maxu - minu) <= 1 && (maxv - minv) <= 1;
		if ((ns_u || cs_u) && (ns_v || cs_v) && no_interp)
			tex &= ~TEX_MAG_LINEAR;
	}

	// Super-sampled textures (triangle_setup.comp:677-789; FORCE_LINEAR_SUPER_SAMPLE is off, it
	// needs field-aware rendering). Sprites sampling per sample keep that for scales near 1:1,
	// fall back to layer 0 for downscales and to the resolved layer for upscales by 2 or more;
	// NEAREST sprites map samples to layers unless a deduced blur kernel blends them
	// (TEX_BLUR_KERNEL, set by the host).
	if (rate != 0 && (args.texture_flags & 1u) &&
	    (tex & (TEX_PER_SAMPLE | TEX_SAMPLE_MAPPING | TEX_FEEDBACK)) == TEX_PER_SAMPLE)
	{
		const int32_t du = abs(int32_t(attr[o1].uv & 0xffffu) - int32_t(attr[o0].uv & 0xffffu));
		const int32_t dv = abs(int32_t(attr[o1].uv >> 16) - int32_t(attr[o0].uv >> 16));
		const int32_t distx = abs(dx - ax), disty = abs(dy - ay);
		if (!(tex & TEX_MAG_LINEAR))
		{
			if (perspective)
				tex |= TEX_SAMPLE_MAPPING;
			else if (du > distx + 1 || dv > disty + 1)
				tex &= ~TEX_PER_SAMPLE;
			else if (du >= distx - 16 || dv >= disty - 16)
				tex = (tex & TEX_BLUR_KERNEL) ? (tex & ~TEX_PER_SAMPLE) : (tex | TEX_SAMPLE_MAPPING);
		}
		else if (!perspective)
		{
			if (du > distx + 1 && dv > disty + 1)
				tex &= ~TEX_PER_SAMPLE;
			else if (distx >= 2 * du && disty >= 2 * dv)
				tex |= TEX_SAMPLE_RESOLVED;
		}
		if ((tex & (TEX_PER_SAMPLE | TEX_SAMPLE_MAPPING)) == TEX_PER_SAMPLE)
		{
			// Per-sample interpolation, shifted so that the samples' mean position interpolates
			// as the pixel's single sample would: 2^3 x the grid's mean sample offset.
			state &= ~ST_SNAP_ATTRIBUTE;
			const float corner = mean_sample_offset8(args.texture_flags >> 8 & 3u, rate);
			const float di_dc = __fmul_rn(float(dbdx + dbdy), corner);
			const float dj_dc = __fmul_rn(float(dcdx + dcdy), corner);
			s.error_i = __fsub_rn(s.error_i, __fmul_rn(di_dc, inv_area_scaled));
			s.error_j = __fsub_rn(s.error_j, __fmul_rn(dj_dc, inv_area_scaled));
		}
	}
	out = s;

	if (prim.fbmsk != args.opaque_fbmask && (state & ST_OPAQUE))
		state &= ~ST_OPAQUE;

	if (rate != 0 && !sprite)
	{
		bool per_sample = multisample;
		if (!per_sample && perspective)
			per_sample = d1[2] != 0.0f || d2[2] != 0.0f;
		if (!per_sample)
			per_sample = za != zb || zb != zc;
		if (!per_sample && !line)
			state |= ST_SNAP_ATTRIBUTE;
		// Flat primitive: its depth bucket gets a heuristic workgroup.
		if (!per_sample && args.flat_buckets)
		{
			const uint32_t bucket = min(255u, za >> args.z_bucket_shift);
			atomicOr(&args.flat_buckets[bucket >> 5], 1u << (bucket & 31u));
		}
	}

	prim.state = state;
	prim.tex = tex;
	if (args.state_copy)
		args.state_copy[index] = state;
}

// First level of the binning hierarchy: for every coarse tile (a square of 2^coarse_log2
// pixels on the binning grid), bit `lane` of word `wg` says whether primitive wg * 64 + lane
// has a bounding box that touches the tile. The bounding-box test is the first test of the
// exact per-tile test, and a fine tile lies inside its coarse tile, so every primitive binned
// to a fine tile has its bit set. Every word is written, so the mask needs no clearing.
// Largest coarse grid side (the launcher checks coarse_x, coarse_y and their product).
constexpr uint32_t kMaxCoarseSide = 256;

// col_bits / row_bits: this wave's kMaxCoarseSide words of LDS each; `word` is the wave's mask
// word (primitives 64 word .. 64 word + 63). Every wave of the workgroup calls it (barrier).
__device__ __forceinline__ void coarse_mask_wave(const pgs_setup_args &args, uint32_t word, uint32_t index,
                                                 uint64_t *col_bits, uint64_t *row_bits)
{
	int32_t cx0 = 1, cy0 = 1, cx1 = 0, cy1 = 0; // empty range
	if (index < args.num_primitives)
	{
		const PrimitiveSetup &s = static_cast<const PrimitiveSetup *>(args.setup)[index];
		if (s.bb[2] >= s.bb[0] && s.bb[3] >= s.bb[1])
		{
			const uint32_t l = args.coarse_log2;
			cx0 = max((s.bb[0] - args.coarse_origin_x) >> l, 0);
			cy0 = max((s.bb[1] - args.coarse_origin_y) >> l, 0);
			cx1 = min((s.bb[2] - args.coarse_origin_x) >> l, int32_t(args.coarse_x) - 1);
			cy1 = min((s.bb[3] - args.coarse_origin_y) >> l, int32_t(args.coarse_y) - 1);
		}
	}
	// A bounding box is a column range times a row range, so the word of coarse tile (cx, cy)
	// is the ballot of the lanes whose columns cover cx AND the ballot of the lanes whose rows
	// cover cy: coarse_x + coarse_y ballots, kept in LDS, then every lane writes the words of
	// tiles lane, lane + 64, ...
	const uint32_t lane = threadIdx.x & 63u;
	for (uint32_t cx = 0; cx < args.coarse_x; cx++)
	{
		const uint64_t bits = __ballot(int32_t(cx) >= cx0 && int32_t(cx) <= cx1);
		if (lane == 0)
			col_bits[cx] = bits;
	}
	for (uint32_t cy = 0; cy < args.coarse_y; cy++)
	{
		const uint64_t bits = __ballot(int32_t(cy) >= cy0 && int32_t(cy) <= cy1);
		if (lane == 0)
			row_bits[cy] = bits;
	}
	__syncthreads();
	if (word >= args.mask_words)
		return; // a wave past the last primitive (workgroups of several waves)
	uint64_t *mask = args.coarse_mask + word;
	const uint32_t tiles = args.coarse_x * args.coarse_y;
	for (uint32_t t = lane; t < tiles; t += 64u)
	{
		const uint32_t cy = t / args.coarse_x, cx = t - cy * args.coarse_x;
		mask[size_t(t) * args.mask_words] = col_bits[cx] & row_bits[cy];
	}
}

// LDS of a setup workgroup of up to kSetupMaxWaves waves: each wave's coarse-mask ballots.
constexpr uint32_t kSetupMaxWaves = 4;
// Per-wave coarse-mask scratch of a setup workgroup of up to W waves.
template <uint32_t W>
struct SetupLdsW
{
	uint64_t col[W][kMaxCoarseSide], row[W][kMaxCoarseSide];
};
using SetupLds = SetupLdsW<kSetupMaxWaves>;

// One thread per primitive, any workgroup size of whole waves (at most the LDS type's W).
template <typename Lds>
__device__ __forceinline__ void triangle_setup_body(const pgs_setup_args &args, uint32_t wg, uint32_t nwg, Lds &lds)
{
	const uint32_t waves = blockDim.x >> 6, wave = threadIdx.x >> 6;
	const uint32_t word = wg * waves + wave;
	const uint32_t index = word * 64u + (threadIdx.x & 63u);
	if (index < args.num_primitives)
		triangle_setup_one(args, index);
	if (args.coarse_mask)
		coarse_mask_wave(args, word, index, lds.col[wave], lds.row[wave]);
	// The pass's sampler-feedback masks start from zero; the feedback launch follows setup.
	if (args.clear_words)
		for (uint32_t w = index; w < args.clear_count; w += nwg * blockDim.x)
			args.clear_words[w] = 0;
}

// ---------------------------------------------------------------------------------------
// Single-sample heuristic (single_sample_heuristic.comp:148-290), one workgroup per depth
// bucket. Masks live in LDS: 65536 bits each for 3D depths, demotion candidates and hashed
// vertices that 3D primitives share with candidates.
constexpr uint32_t kHeurThreads = 512;
constexpr uint32_t kMaskWords = 65536 / 32;

__device__ __forceinline__ uint32_t hash_vertex(const VertexPosition &v)
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

struct HeuristicLds
{
	uint32_t full3d[kMaskWords], cand[kMaskWords], shared_v[kMaskWords];
	uint32_t any3d, min3d, max3d;
};

// Any workgroup size of whole waves (the loops stride by the workgroup size).
__device__ __forceinline__ void ssaa_heuristic_body(const pgs_ssaa_heuristic_args &a, uint32_t wg, uint32_t,
                                                    HeuristicLds &lds)
{
	const uint32_t bucket = wg;
	if (((a.flat_buckets[bucket >> 5] >> (bucket & 31u)) & 1u) == 0)
		return;

	uint32_t *full3d = lds.full3d, *cand = lds.cand, *shared_v = lds.shared_v;
	uint32_t &any3d = lds.any3d, &min3d = lds.min3d, &max3d = lds.max3d;
	const uint32_t threads = blockDim.x;
	for (uint32_t i = threadIdx.x; i < kMaskWords; i += threads)
	{
		full3d[i] = 0;
		cand[i] = 0;
		shared_v[i] = 0;
	}
	if (threadIdx.x == 0)
	{
		any3d = 0;
		min3d = ~0u;
		max3d = 0;
	}
	__syncthreads();

	const VertexPosition *pos = static_cast<const VertexPosition *>(a.positions);
	PrimitiveAttribute *prims = static_cast<PrimitiveAttribute *>(a.prims);
	const uint32_t zlo = bucket << a.z_shift_bucket;
	const uint32_t zhi = ((bucket + 1u) << a.z_shift_bucket) - 1u;
	auto snapped = [&](uint32_t z) { return min(a.z_max, z) >> a.z_shift; };
	auto in_range = [&](uint32_t sz) { return sz >= zlo && sz <= zhi; };
	const uint32_t NO_FULL_3D = ST_SPRITE | ST_LINE | ST_SNAP_ATTRIBUTE;
	const uint32_t SA = ST_SNAP_ATTRIBUTE, SR = ST_SNAP_RASTER;

	// The 3D depth range is reduced in registers and across the wave first: one LDS atomic pair
	// per wave instead of one per vertex on the same two words.
	uint32_t zmin = ~0u, zmax = 0u;
	bool has3d = false;
	for (uint32_t i = threadIdx.x; i < a.num_primitives; i += threads)
	{
		const uint32_t st = a.state_copy[i];
		if ((st & NO_FULL_3D) == 0)
		{
			for (uint32_t k = 0; k < 3; k++)
			{
				const uint32_t z = pos[3u * i + k].z;
				zmin = min(zmin, z);
				zmax = max(zmax, z);
				const uint32_t sz = snapped(z);
				if (in_range(sz))
					atomicOr(&full3d[(sz - zlo) >> 5], 1u << ((sz - zlo) & 31u));
			}
			has3d = true;
		}
		else if ((st & (SA | SR)) == SA || (st & ST_LINE))
		{
			const uint32_t sz = snapped(pos[3u * i].z);
			if (in_range(sz))
				atomicOr(&cand[(sz - zlo) >> 5], 1u << ((sz - zlo) & 31u));
		}
	}
	for (int m = 32; m >= 1; m >>= 1)
	{
		zmin = min(zmin, uint32_t(__shfl_xor(int(zmin), m)));
		zmax = max(zmax, uint32_t(__shfl_xor(int(zmax), m)));
	}
	if (__ballot(has3d) != 0 && (threadIdx.x & 63u) == 0)
	{
		atomicMin(&min3d, zmin);
		atomicMax(&max3d, zmax);
		any3d = 1;
	}
	__syncthreads();

	for (uint32_t i = threadIdx.x; i < a.num_primitives; i += threads)
	{
		if ((a.state_copy[i] & NO_FULL_3D) != 0)
			continue;
		for (uint32_t k = 0; k < 3; k++)
		{
			const VertexPosition v = pos[3u * i + k];
			const uint32_t sz = snapped(v.z);
			if (in_range(sz) && ((cand[(sz - zlo) >> 5] >> ((sz - zlo) & 31u)) & 1u))
			{
				const uint32_t h = hash_vertex(v);
				atomicOr(&shared_v[h >> 5], 1u << (h & 31u));
			}
		}
	}
	__syncthreads();

	auto shared_bit = [&](const VertexPosition &v) {
		const uint32_t h = hash_vertex(v);
		return ((shared_v[h >> 5] >> (h & 31u)) & 1u) != 0;
	};
	for (uint32_t i = threadIdx.x; i < a.num_primitives; i += threads)
	{
		if ((a.state_copy[i] & (SA | SR)) != SA)
			continue;
		const VertexPosition va = pos[3u * i];
		const uint32_t sz = snapped(va.z);
		if (!in_range(sz) || ((full3d[(sz - zlo) >> 5] >> ((sz - zlo) & 31u)) & 1u) == 0)
			continue;
		if (shared_bit(va) || shared_bit(pos[3u * i + 1]) || shared_bit(pos[3u * i + 2]))
			atomicAnd(&cand[(sz - zlo) >> 5], ~(1u << ((sz - zlo) & 31u)));
	}
	__syncthreads();

	const bool force_single = any3d == 0;
	for (uint32_t i = threadIdx.x; i < a.num_primitives; i += threads)
	{
		const uint32_t z = pos[3u * i].z;
		const uint32_t sz = snapped(z);
		if (!in_range(sz))
			continue;
		const uint32_t st = a.state_copy[i];
		const bool tri_cand = (st & (SA | SR)) == SA;
		const bool line_cand = force_single && (st & ST_LINE);
		if (!tri_cand && !line_cand)
			continue;
		if ((cand[(sz - zlo) >> 5] >> ((sz - zlo) & 31u)) & 1u)
			prims[i].state = st | SR;
		else if ((st & (ST_Z_TEST | ST_Z_WRITE)) && z > min3d && z < max3d)
			prims[i].state = st & ~SA;
	}
	__syncthreads();
	// Leave the bucket mask clear for the next pass.
	if (threadIdx.x == 0)
		atomicAnd(&a.flat_buckets[bucket >> 5], ~(1u << (bucket & 31u)));
}

// intersect.h:24-81: conservative test of a primitive against a pixel rectangle.
__device__ __forceinline__ bool intersects_tile(const PrimitiveSetup &s, int32_t tlox, int32_t tloy, int32_t thix,
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
		da = wadd(da, 8 * max(abs(ax), abs(s.a[1])));
		db = wadd(db, 8 * max(abs(s.b[0]), abs(s.b[1])));
		dc = wadd(dc, 8 * max(abs(s.c[0]), abs(s.c[1])));
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

// ---------------------------------------------------------------------------------------
// Binning: one workgroup (4 waves) per tile. Candidates are tested 256 at a time; hits are
// compacted in ascending primitive order with wave ballots and an LDS exchange of wave counts.
// With a coarse mask (written by triangle setup) the candidates are the primitives whose bit
// is set in the tile's coarse tile, taken window by window: each thread owns one 64-bit mask
// word, a block-wide prefix sum of the word popcounts places the set bits' indices in LDS in
// ascending order, and only those primitives are tested. Without one, every primitive is a
// candidate. Both give the same lists: the coarse bit is set for every primitive that passes
// the exact test. A window is 128 mask words (8192 primitives): its 16 KiB of candidates let 8
// workgroups share a CU (32 KiB windows allowed 5), which hides more of the latency of the
// candidates' setup-record loads; threads past the window load no word.
constexpr uint32_t kBinWindowWords = 128;
static_assert(kBinWindowWords <= 256, "a window holds at most one mask word per thread");

struct BinningLds
{
	uint32_t wave_counts[4], wave_flags[4], window_counts[4];
	uint16_t candidates[kBinWindowWords * 64];
};

// Workgroups of 256 threads.
__device__ __forceinline__ void binning_body(const pgs_bin_args &args, uint32_t wg, uint32_t, BinningLds &lds)
{
	const uint32_t tile = wg;
	const uint32_t tx = tile % args.tiles_x, ty = tile / args.tiles_x;
	const int32_t size = 1 << args.tile_log2;
	const int32_t lox = args.origin_x + int32_t(tx) * size, loy = args.origin_y + int32_t(ty) * size;
	const int32_t hix = lox + size - 1, hiy = loy + size - 1;
	const PrimitiveSetup *setup = static_cast<const PrimitiveSetup *>(args.setup);
	const PrimitiveAttribute *prims = static_cast<const PrimitiveAttribute *>(args.prims);
	uint16_t *list = args.lists + size_t(tile) * args.list_stride;

	uint32_t *wave_counts = lds.wave_counts, *wave_flags = lds.wave_flags, *window_counts = lds.window_counts;
	uint16_t *candidates = lds.candidates;
	const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	uint32_t written = 0;
	// binning.comp:87-125: flags of the primitives binned to this tile.
	bool req = false, req_zw = false, has_zw = false;

	// Tests primitive i (when `valid`) and appends the workgroup's hits; ends on a barrier.
	auto test_chunk = [&](bool valid, uint32_t i) {
		bool hit = false;
		if (valid)
		{
			const uint32_t st = prims[i].state;
			const bool ms_tri = (st & ST_MULTISAMPLE) && !(st & ST_LINE);
			hit = intersects_tile(setup[i], lox, loy, hix, hiy, ms_tri);
			if (hit)
			{
				if ((st & (ST_SNAP_RASTER | ST_Z_TEST | ST_OPAQUE)) != (ST_SNAP_RASTER | ST_OPAQUE))
					req_zw = true;
				if ((st & (ST_SNAP_RASTER | ST_OPAQUE)) != (ST_SNAP_RASTER | ST_OPAQUE))
					req = true;
				// Sampling a super-sampled texture per sample needs the samples (binning.comp:123-124).
				if ((args.super_sampled & 2u) && (prims[i].tex & (TEX_PER_SAMPLE | TEX_FEEDBACK)) == TEX_PER_SAMPLE)
					req = true;
				if (st & ST_Z_WRITE)
					has_zw = true;
			}
		}
		const uint64_t ballot = __ballot(hit);
		if (lane == 0)
			wave_counts[wave] = uint32_t(__popcll(ballot));
		__syncthreads();
		uint32_t before = 0;
		for (uint32_t w = 0; w < wave; w++)
			before += wave_counts[w];
		const uint32_t total = wave_counts[0] + wave_counts[1] + wave_counts[2] + wave_counts[3];
		if (hit)
		{
			const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(ballot >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(ballot), 0u));
			list[written + before + rank] = uint16_t(i);
		}
		written += total;
		__syncthreads();
	};

	if (!args.coarse_mask)
	{
		for (uint32_t base = 0; base < args.num_primitives; base += 256u)
			test_chunk(base + threadIdx.x < args.num_primitives, base + threadIdx.x);
	}
	else
	{
		const uint32_t ctile = (ty >> args.coarse_shift) * args.coarse_x + (tx >> args.coarse_shift);
		const uint64_t *cmask = args.coarse_mask + size_t(ctile) * args.mask_words;
		for (uint32_t win = 0; win < args.mask_words; win += kBinWindowWords)
		{
			const uint32_t w = win + threadIdx.x;
			uint64_t m = (threadIdx.x < kBinWindowWords && w < args.mask_words) ? cmask[w] : 0ull;
			const uint32_t c = uint32_t(__popcll(m));
			uint32_t incl = c;
			for (uint32_t d = 1; d < 64u; d <<= 1)
			{
				const uint32_t t = __shfl_up(incl, d, 64);
				if (lane >= d)
					incl += t;
			}
			if (lane == 63)
				window_counts[wave] = incl;
			__syncthreads();
			uint32_t pos = incl - c, total = 0;
			for (uint32_t q = 0; q < 4; q++)
			{
				pos += q < wave ? window_counts[q] : 0u;
				total += window_counts[q];
			}
			while (m)
			{
				candidates[pos++] = uint16_t(w * 64u + uint32_t(__builtin_ctzll(m)));
				m &= m - 1u;
			}
			__syncthreads();
			for (uint32_t base = 0; base < total; base += 256u)
			{
				const uint32_t j = base + threadIdx.x;
				test_chunk(j < total, j < total ? uint32_t(candidates[j]) : 0u);
			}
			__syncthreads(); // window_counts and candidates are rewritten by the next window
		}
	}
	if (args.tile_modes)
	{
		const uint32_t f = (__ballot(req) ? 1u : 0u) | (__ballot(req_zw) ? 2u : 0u) | (__ballot(has_zw) ? 4u : 0u);
		if (lane == 0)
			wave_flags[wave] = f;
		__syncthreads();
		if (threadIdx.x == 0)
		{
			const uint32_t all = wave_flags[0] | wave_flags[1] | wave_flags[2] | wave_flags[3];
			const bool super = args.super_sampled && ((all & 1u) || ((all & 2u) && (all & 4u)));
			args.tile_modes[tile] = written == 0 ? 0u : (super ? 2u : 1u);
		}
	}
	if (threadIdx.x == 0)
		args.counts[tile] = written;
}

// The bodies with their own LDS, for their own launches.
// Setup's own launches are one wave per workgroup: 4 KiB of LDS each (the 4-wave layout's
// 16 KiB held a CU to 10 of them).
__device__ __forceinline__ void setup_own_lds(const pgs_setup_args &a, uint32_t wg, uint32_t nwg)
{
	__shared__ SetupLdsW<1> lds;
	triangle_setup_body(a, wg, nwg, lds);
}
__device__ __forceinline__ void heuristic_own_lds(const pgs_ssaa_heuristic_args &a, uint32_t wg, uint32_t nwg)
{
	__shared__ HeuristicLds lds;
	ssaa_heuristic_body(a, wg, nwg, lds);
}
__device__ __forceinline__ void binning_own_lds(const pgs_bin_args &a, uint32_t wg, uint32_t nwg)
{
	__shared__ BinningLds lds;
	binning_body(a, wg, nwg, lds);
}

__global__ void __launch_bounds__(64) k_triangle_setup(pgs_setup_args a) { setup_own_lds(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(kHeurThreads) k_ssaa_heuristic(pgs_ssaa_heuristic_args a) { heuristic_own_lds(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) k_binning(pgs_bin_args a) { binning_own_lds(a, blockIdx.x, gridDim.x); }
PGS_DEFINE_MULTI(64, k_triangle_setup_multi, setup_own_lds, pgs_setup_args)
PGS_DEFINE_MULTI(kHeurThreads, k_ssaa_heuristic_multi, heuristic_own_lds, pgs_ssaa_heuristic_args)
PGS_DEFINE_MULTI(256, k_binning_multi, binning_own_lds, pgs_bin_args)

// Fused launch: the parts' workgroups side by side, 256 threads each, one LDS area shared by
// the parts' bodies (a workgroup runs one part).
union FusedLds
{
	SetupLds setup;
	HeuristicLds heuristic;
	BinningLds binning;
};
__global__ void __launch_bounds__(256) k_fused(pgs_fused_args f)
{
	__shared__ FusedLds lds;
	uint32_t p = 0;
	while (p + 1u < f.num_parts && blockIdx.x >= uni(f.parts[p + 1u].begin))
		p++;
	const pgs_fused_part &part = f.parts[p];
	const uint32_t wg = blockIdx.x - uni(part.begin), n = uni(part.groups);
	switch (uni(part.kind))
	{
	case PGS_FUSED_VRAM_COPY: xfer::resolve_blocks_body(part.args.copy, wg, n); break;
	case PGS_FUSED_VRAM_SNAPSHOT: xfer::snapshot_blocks_body(part.args.snapshot, wg, n); break;
	case PGS_FUSED_CLUT_WRITE: texb::clut_write_body(part.args.clut, wg, n); break;
	case PGS_FUSED_TEXTURE_DECODE: texb::texture_decode_body(part.args.decode, wg, n); break;
	case PGS_FUSED_TRIANGLE_SETUP: triangle_setup_body(part.args.setup, wg, n, lds.setup); break;
	case PGS_FUSED_SSAA_HEURISTIC: ssaa_heuristic_body(part.args.heuristic, wg, n, lds.heuristic); break;
	case PGS_FUSED_BINNING: binning_body(part.args.bin, wg, n, lds.binning); break;
	default: break;
	}
}
}
}

// ---------------------------------------------------------------------------------------
// Shading lives in kernels_shade.hip; launchers for setup and binning:

extern "C" int pgs_k_triangle_setup(const pgs_setup_args *args, hipStream_t stream)
{
	if (!args || !args->positions || !args->attributes || !args->prims || !args->setup || !args->tattr)
		return -int(hipErrorInvalidValue);
	if (args->coarse_mask && (args->coarse_x == 0 || args->coarse_y == 0 || args->coarse_log2 > 16 ||
	                          args->coarse_x > pgs::kMaxCoarseSide || args->coarse_y > pgs::kMaxCoarseSide ||
	                          args->mask_words != (args->num_primitives + 63u) / 64u))
		return -int(hipErrorInvalidValue);
	if (args->num_primitives == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_triangle_setup, dim3((args->num_primitives + 63u) / 64u), dim3(64), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

extern "C" int pgs_k_ssaa_heuristic(const pgs_ssaa_heuristic_args *args, hipStream_t stream)
{
	if (!args || !args->positions || !args->prims || !args->state_copy || !args->flat_buckets)
		return -int(hipErrorInvalidValue);
	if (args->num_primitives == 0)
		return 0;
	if (args->z_shift_bucket > 16 || args->z_shift > 8)
		return -int(hipErrorInvalidValue);
	hipLaunchKernelGGL(pgs::k_ssaa_heuristic, dim3(256), dim3(pgs::kHeurThreads), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

// The parts live in device memory: their layout is the caller's (the renderer checks it on the
// host before the upload); here only the shape of the launch is checked.
extern "C" int pgs_k_fused(const pgs_fused_args *args, hipStream_t stream)
{
	if (!args || !args->parts || args->num_parts == 0 || args->num_parts > PGS_FUSED_MAX_PARTS)
		return -int(hipErrorInvalidValue);
	if (args->groups == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_fused, dim3(args->groups), dim3(256), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

extern "C" int pgs_k_binning(const pgs_bin_args *args, hipStream_t stream)
{
	if (!args || !args->setup || !args->prims || !args->lists || !args->counts)
		return -int(hipErrorInvalidValue);
	if (args->num_primitives > 65536u || args->list_stride < args->num_primitives)
		return -int(hipErrorInvalidValue);
	if (args->coarse_mask && args->tiles_x && args->tiles_y &&
	    (args->coarse_shift > 16 || args->mask_words != (args->num_primitives + 63u) / 64u ||
	     ((args->tiles_x - 1u) >> args->coarse_shift) >= args->coarse_x ||
	     ((args->tiles_y - 1u) >> args->coarse_shift) >= args->coarse_y))
		return -int(hipErrorInvalidValue);
	const uint32_t tiles = args->tiles_x * args->tiles_y;
	if (tiles == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_binning, dim3(tiles), dim3(256), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

// Uploads the reciprocal seed table and float LUT (gs/gs_renderer.cpp:3600-3640).
extern "C" int pgs_k_init_tables(const uint8_t *seed, const float *lut)
{
	hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(pgs::c_rcp_seed), seed, 256);
	if (e == hipSuccess)
		e = hipMemcpyToSymbol(HIP_SYMBOL(pgs::c_rcp_float_lut), lut, 256 * sizeof(float));
	return e == hipSuccess ? 0 : -int(e);
}

extern "C" int pgs_k_rcp_eval(uint32_t mode, uint32_t first, uint64_t count, uint64_t *hash, hipStream_t stream)
{
	if (!hash || mode > 1 || count > (uint64_t(1) << 32))
		return -int(hipErrorInvalidValue);
	if (count == 0)
		return 0;
	const uint64_t blocks = std::min<uint64_t>(4096, (count + 255u) / 256u);
	hipLaunchKernelGGL(pgs::k_rcp_eval, dim3(uint32_t(blocks)), dim3(256), 0, stream, mode, first, count,
	                   reinterpret_cast<unsigned long long *>(hash));
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

uint32_t pgs_grid_triangle_setup(const pgs_setup_args &a) { return (a.num_primitives + 63u) / 64u; }
uint32_t pgs_grid_ssaa_heuristic(const pgs_ssaa_heuristic_args &a) { return a.num_primitives ? 256u : 0u; }
uint32_t pgs_grid_binning(const pgs_bin_args &a) { return a.tiles_x * a.tiles_y; }

int pgs_multi_triangle_setup(const pgs_setup_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                             hipStream_t stream)
{
	hipLaunchKernelGGL(pgs::k_triangle_setup_multi, dim3(total), dim3(64), 0, stream, args, prefix, n);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

int pgs_multi_ssaa_heuristic(const pgs_ssaa_heuristic_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                             hipStream_t stream)
{
	hipLaunchKernelGGL(pgs::k_ssaa_heuristic_multi, dim3(total), dim3(pgs::kHeurThreads), 0, stream, args, prefix, n);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

int pgs_multi_binning(const pgs_bin_args *args, const uint32_t *prefix, uint32_t n, uint32_t total, hipStream_t stream)
{
	hipLaunchKernelGGL(pgs::k_binning_multi, dim3(total), dim3(256), 0, stream, args, prefix, n);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}
