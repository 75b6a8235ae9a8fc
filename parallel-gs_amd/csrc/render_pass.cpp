// Primitive assembly and render-pass batching.
#include "render_pass.hpp"
#include "gs_interface.hpp"
#include <algorithm>
#include <cstring>
#include <cstdlib>
#include <limits>

namespace pgs
{
namespace
{
inline uint32_t field(uint64_t v, uint32_t off, uint32_t bits) { return uint32_t((v >> off) & ((1ull << bits) - 1ull)); }
inline int32_t wrapmul(int32_t a, int32_t b) { return int32_t(uint32_t(a) * uint32_t(b)); }

// gs/gs_util.cpp:237-310
bool parallelogram_candidate(const VertexPosition *p, const VertexAttribute *a, int32_t lox, int32_t loy,
                             int32_t hix, int32_t hiy, uint64_t prim, int32_t order[3])
{
	const uint32_t type = field(prim, 0, 3);
	if (field(prim, 7, 1) && type != 6 && type != 0)
		return false;
	const bool fst = field(prim, 8, 1), tme = field(prim, 4, 1), fge = field(prim, 5, 1);
	if (!fst && tme && (a[0].q != a[1].q || a[1].q != a[2].q))
		return false;
	if (p[0].z != p[1].z || p[1].z != p[2].z)
		return false;
	if (fge && (a[0].fog != a[1].fog || a[1].fog != a[2].fog))
		return false;
	if (a[0].rgba != a[1].rgba || a[1].rgba != a[2].rgba)
		return false;
	const int32_t abx = p[1].x - p[0].x, aby = p[1].y - p[0].y;
	const int32_t acx = p[2].x - p[0].x, acy = p[2].y - p[0].y;
	const int32_t bcx = p[2].x - p[1].x, bcy = p[2].y - p[1].y;
	const int32_t area = std::abs(int32_t(uint32_t(wrapmul(abx, acy)) - uint32_t(wrapmul(aby, acx))));
	if (area != wrapmul(hix - lox, hiy - loy))
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
	if (p[order[1]].x > p[order[2]].x)
		std::swap(order[1], order[2]);
	return true;
}

// gs/gs_util.cpp:312-377
bool triangles_form_parallelogram(const VertexPosition *p, const VertexAttribute *a, const int32_t o[3],
                                  const VertexPosition *lp, const VertexAttribute *la, const int32_t lo[3],
                                  uint64_t prim)
{
	const int32_t p3x = p[o[1]].x + p[o[2]].x - p[o[0]].x, p3y = p[o[1]].y + p[o[2]].y - p[o[0]].y;
	if (p3x != lp[lo[0]].x || p3y != lp[lo[0]].y || p[o[1]].x != lp[lo[1]].x || p[o[1]].y != lp[lo[1]].y ||
	    p[o[2]].x != lp[lo[2]].x || p[o[2]].y != lp[lo[2]].y)
		return false;
	const bool fst = field(prim, 8, 1), tme = field(prim, 4, 1), fge = field(prim, 5, 1);
	if (p[0].z != lp[0].z)
		return false;
	if (!fst && tme && a[0].q != la[0].q)
		return false;
	if (fge && a[0].fog != la[0].fog)
		return false;
	if (a[0].rgba != la[0].rgba)
		return false;
	if (tme)
	{
		if (fst)
		{
			const uint32_t u3 = ((a[o[1]].uv & 0xffffu) + (a[o[2]].uv & 0xffffu) - (a[o[0]].uv & 0xffffu)) & 0xffffu;
			const uint32_t v3 = ((a[o[1]].uv >> 16) + (a[o[2]].uv >> 16) - (a[o[0]].uv >> 16)) & 0xffffu;
			if (u3 != (la[lo[0]].uv & 0xffffu) || v3 != (la[lo[0]].uv >> 16) || a[o[1]].uv != la[lo[1]].uv ||
			    a[o[2]].uv != la[lo[2]].uv)
				return false;
		}
		else
		{
			const float es = std::fabs((a[o[2]].s + a[o[1]].s - a[o[0]].s - la[lo[0]].s) / a[0].q);
			const float et = std::fabs((a[o[2]].t + a[o[1]].t - a[o[0]].t - la[lo[0]].t) / a[0].q);
			if (a[o[1]].s != la[lo[1]].s || a[o[1]].t != la[lo[1]].t || a[o[2]].s != la[lo[2]].s ||
			    a[o[2]].t != la[lo[2]].t || es > 1e-4f || et > 1e-4f)
				return false;
		}
	}
	return true;
}
}

void RenderPassBuilder::set_sample_log2(uint32_t sample_log2, bool sparse)
{
	sample_log2_ = sample_log2;
	// Samples along y, which sets the raster snap and bounding-box rounding
	// (gs/gs_interface.cpp:65-118).
	uint32_t rate_x;
	grid_rates(sample_log2, sparse, rate_x, rate_log2_);
}

void RenderPassBuilder::init(Renderer *renderer, uint32_t sample_log2)
{
	renderer_ = renderer;
	set_sample_log2(sample_log2);
	pos_.reserve(3u * 65536u);
	attr_.reserve(3u * 65536u);
	prims_.reserve(65536u);
}

void RenderPassBuilder::vertex_kick(const RegisterState &r, uint32_t x, uint32_t y, uint32_t z, float fog, bool adc)
{
	if (vq_count_ == 3)
	{
		vq_[0] = vq_[1];
		vq_[1] = vq_[2];
		vq_count_ = 2;
	}
	const uint32_t ctxt = field(r.prim, 9, 1);
	QueuedVertex &v = vq_[vq_count_++];
	v.pos.x = int32_t(x) - int32_t(field(r.xyoffset[ctxt], 0, 16));
	v.pos.y = int32_t(y) - int32_t(field(r.xyoffset[ctxt], 32, 16));
	v.pos.z = z;
	v.pos.pad = 0;
	uint32_t sbits = uint32_t(r.st), tbits = uint32_t(r.st >> 32), qbits = uint32_t(r.rgbaq >> 32);
	memcpy(&v.attr.s, &sbits, 4);
	memcpy(&v.attr.t, &tbits, 4);
	memcpy(&v.attr.q, &qbits, 4);
	v.attr.rgba = uint32_t(r.rgbaq);
	v.attr.fog = fog;
	v.attr.uv = field(r.uv, 0, 14) | (field(r.uv, 16, 14) << 16);

	// gs/gs_interface.cpp:2838-2928
	const uint32_t type = field(r.prim, 0, 3);
	static const uint32_t nverts[8] = { 1, 2, 2, 3, 3, 3, 2, 0 };
	const uint32_t n = nverts[type];
	if (n == 0)
	{
		vq_count_ = 0;
		return;
	}
	if (vq_count_ < n)
		return;
	if (!adc && !draw_is_degenerate(r))
	{
		QueuedVertex v3[3];
		if (n == 1)
		{
			v3[0] = vq_[vq_count_ - 1];
			v3[0].pos.x -= 7;
			v3[0].pos.y -= 7;
			v3[1] = v3[0];
			v3[1].pos.x += 16;
			v3[1].pos.y += 16;
			v3[2] = v3[0];
		}
		else if (n == 2)
		{
			v3[0] = vq_[vq_count_ - 1];
			v3[1] = vq_[vq_count_ - 2];
			v3[2] = v3[0];
		}
		else
		{
			v3[0] = vq_[2];
			v3[1] = vq_[1];
			v3[2] = vq_[0];
		}
		append(r, v3, n, type == 0 || type == 6);
	}
	if (type == 5)
	{
		vq_[1] = vq_[2];
		vq_count_ = 2;
	}
	else if (type == 0 || type == 1 || type == 3 || type == 6)
		vq_count_ = 0;
}

// gs/gs_interface.cpp:2036-2115
bool RenderPassBuilder::draw_is_degenerate(const RegisterState &r) const
{
	const uint32_t ctx = field(r.prim, 9, 1);
	const uint64_t sc = r.scissor[ctx], test = r.test[ctx], frame = r.frame[ctx], zbuf = r.zbuf[ctx], alpha = r.alpha[ctx];
	if (field(sc, 0, 11) > field(sc, 16, 11) || field(sc, 32, 11) > field(sc, 48, 11))
		return true;
	const uint32_t ate = field(test, 0, 1), atst = field(test, 1, 3), afail = field(test, 12, 2);
	const uint32_t zte = field(test, 16, 1), ztst = field(test, 17, 2);
	if (zte && ztst == 0)
		return true;
	if (ate && atst == 0 && afail == 0)
		return true;
	uint32_t fbmsk = uint32_t(frame >> 32);
	uint32_t zmsk = field(zbuf, 32, 1);
	const uint32_t psm = field(frame, 24, 6);
	if (psm == PSMCT24 || psm == PSMZ24)
		fbmsk |= 0xffu << 24;
	if (ate && atst == 0)
	{
		if (afail == 2) fbmsk = ~0u;
		else if (afail == 0) { fbmsk = ~0u; zmsk = 1; }
		else if (afail == 1) zmsk = 1;
		else fbmsk |= 0xffu << 24;
	}
	if (field(r.prim, 6, 1) && field(alpha, 6, 2) == 1 && field(alpha, 0, 2) == field(alpha, 2, 2))
		fbmsk |= 0xffffffu;
	return fbmsk == ~0u && (zmsk || zte == 0);
}

uint32_t RenderPassBuilder::state_index(const StateVector &sv)
{
	uint64_t key = (uint64_t(sv.blend_mode) << 32) ^ (uint64_t(sv.combiner) << 24) ^
	               (uint64_t(sv.dimx0) * 0x9e3779b97f4a7c15ull) ^ (uint64_t(sv.dimx1) * 0xc2b2ae3d27d4eb4full);
	auto it = state_map_.find(key);
	if (it != state_map_.end())
	{
		const StateVector &e = states_[it->second];
		if (e.combiner == sv.combiner && e.blend_mode == sv.blend_mode && e.dimx0 == sv.dimx0 && e.dimx1 == sv.dimx1)
			return it->second;
		// Hash collision: fall back to a linear search.
		for (uint32_t i = 0; i < states_.size(); i++)
			if (states_[i].combiner == sv.combiner && states_[i].blend_mode == sv.blend_mode &&
			    states_[i].dimx0 == sv.dimx0 && states_[i].dimx1 == sv.dimx1)
				return i;
	}
	uint32_t idx = uint32_t(states_.size());
	states_.push_back(sv);
	state_map_.emplace(key, idx);
	return idx;
}

namespace
{
// Formats sharing a swizzle layout (gs/shaders/swizzle_utils.h:298-325): colour-32 and the
// 8H / 4HL / 4HH / 24 formats living in its words share one key.
uint32_t swizzle_compat_key(uint32_t psm)
{
	switch (psm)
	{
	case PSMZ32: return 2;
	case PSMCT16: return 3;
	case PSMZ16: return 4;
	case PSMCT16S: return 5;
	case PSMZ16S: return 6;
	case PSMT8: return 7;
	case PSMT4: return 8;
	case PSMZ24: return 9;
	default: return 1;
	}
}

// Bits of a 32-bit word a format writes (gs/gs_util.cpp:13-29).
uint32_t psm_word_write_mask(uint32_t psm)
{
	switch (psm)
	{
	case PSMCT24: case PSMZ24: return 0xffffffu;
	case PSMT8H: return 0xffu << 24;
	case PSMT4HL: return 0xfu << 24;
	case PSMT4HH: return 0xfu << 28;
	default: return 0xffffffffu;
	}
}
}

// Whether a textured primitive samples its own frame buffer (colour) or depth buffer pixel for
// pixel (gs/gs_interface.cpp:2141-2307 update_color_feedback_state and 2388-2470
// deduce_color_feedback_mode). Pixel: the texture starts at the frame (or depth) buffer with the
// same layout and stride, the primitive is an affine triangle or a sprite inside the texture
// whose UVs land on the pixels' own texels (offsets in [0, 1) texel with NEAREST, exactly half a
// texel with LINEAR); the shader then reads the pixel's current value, per sample in a
// super-sampled pass. Bypass: the texture aliases the frame buffer with another layout; it is
// read as decoded at the start of the pass (no ordering against the pass's own writes).
// Everything else (None) is an ordinary texture read, ordered by the pass's hazard checks.
RenderPassBuilder::Feedback RenderPassBuilder::feedback_mode(const RegisterState &r, uint32_t ctx,
                                                             const VertexPosition *p, const VertexAttribute *a,
                                                             uint32_t n, bool quad, const int32_t bb[4],
                                                             bool &depth) const
{
	depth = false;
	const uint64_t t0 = r.tex0[ctx], t1 = r.tex1[ctx], cl = r.clamp[ctx], frame = r.frame[ctx], zbuf = r.zbuf[ctx];
	if (mmin_has_mipmap(t1) && field(t1, 2, 3) != 0 && !disable_mipmaps)
		return Feedback::None;
	const uint32_t tex_psm = field(t0, 20, 6), tbw = field(t0, 14, 6), fb_psm = field(frame, 24, 6);
	const uint32_t wms = field(cl, 0, 2), wmt = field(cl, 2, 2);
	const bool palette = tex_psm == PSMT8 || tex_psm == PSMT4 || tex_psm == PSMT8H || tex_psm == PSMT4HL ||
	                     tex_psm == PSMT4HH;
	uint32_t tbp0 = field(t0, 0, 14);
	// REGION_CLAMP can align the texture's effective base with the frame buffer.
	if (!palette)
	{
		const Geometry g = psm_geometry(tex_psm);
		if (wms == 2 && field(cl, 4, 10) != 0)
			tbp0 += 32u * (field(cl, 4, 10) >> g.pw);
		if (wmt == 2 && field(cl, 24, 10) != 0)
			tbp0 += 32u * tbw * (field(cl, 24, 10) >> g.ph);
	}
	const uint64_t test = r.test[ctx];
	bool z_sensitive = pass_z_;
	if (field(test, 16, 1))
	{
		const uint32_t ate = field(test, 0, 1), atst = field(test, 1, 3), afail = field(test, 12, 2);
		const bool implied_z_mask = ate && atst == 0 && afail != 2;
		z_sensitive = z_sensitive || field(test, 17, 2) > 1 || (!implied_z_mask && field(zbuf, 32, 1) == 0);
	}
	const bool equal_color = tbp0 == field(frame, 0, 9) * 32u;
	const bool equal_depth = tbp0 == field(zbuf, 0, 9) * 32u && z_sensitive;
	bool awkward = false;
	if (equal_color && (swizzle_compat_key(tex_psm) != swizzle_compat_key(fb_psm) || tbw != field(frame, 16, 6)))
		awkward = true;
	else
	{
		if (equal_depth &&
		    (swizzle_compat_key(tex_psm) != swizzle_compat_key(field(zbuf, 24, 4) | 0x30u) || tbw != field(frame, 16, 6)))
			return Feedback::None;
		if (wms == 3 || wmt == 3 || (!equal_color && !equal_depth))
			return Feedback::None;
	}
	const bool color_fb = equal_color && (awkward || (psm_word_write_mask(tex_psm) & psm_word_write_mask(fb_psm)) != 0);
	const bool depth_fb = !awkward && equal_depth &&
	                      (psm_word_write_mask(tex_psm) & psm_word_write_mask(field(zbuf, 24, 4))) != 0;
	if (!color_fb && !depth_fb)
		return Feedback::None;

	// Lines and points never read feedback per pixel.
	if (!(n == 3 || (quad && n == 2)))
		return Feedback::None;
	if (awkward)
		return Feedback::Bypass;
	const int32_t width = 1 << int32_t(field(t0, 26, 4)), height = 1 << int32_t(field(t0, 30, 4));
	// Texel positions of the vertices in 1/16 texel (compute_uv_bb, gs/gs_interface.cpp:2309-2380).
	int32_t uv[3][2] = {};
	const uint32_t nv = quad ? 2u : 3u;
	if (field(r.prim, 8, 1))
	{
		for (uint32_t i = 0; i < nv; i++)
		{
			uv[i][0] = int32_t(a[i].uv & 0xffffu);
			uv[i][1] = int32_t(a[i].uv >> 16);
		}
	}
	else
	{
		// Perspective varies across a triangle whose Q differ: no pixel correspondence.
		if (!quad && (a[0].q != a[1].q || a[1].q != a[2].q))
			return Feedback::None;
		const float fw = float(width * 16), fh = float(height * 16);
		for (uint32_t i = 0; i < nv; i++)
		{
			const float inv_q = 1.0f / a[i].q;
			const float u = fw * (a[i].s * inv_q), v = fh * (a[i].t * inv_q);
			// Anything not a modest finite position cannot land on the pixel's own texel.
			if (!(std::fabs(u) < 1073741824.0f) || !(std::fabs(v) < 1073741824.0f))
				return Feedback::None;
			uv[i][0] = int32_t(u);
			uv[i][1] = int32_t(v);
		}
	}
	// A primitive reaching outside the texture (or its clamp region) sees clamping or wrapping.
	if (wms == 2 ? (bb[0] < int32_t(field(cl, 4, 10)) || bb[2] > int32_t(field(cl, 14, 10))) : bb[2] >= width)
		return Feedback::None;
	if (wmt == 2 ? (bb[1] < int32_t(field(cl, 24, 10)) || bb[3] > int32_t(field(cl, 34, 10))) : bb[3] >= height)
		return Feedback::None;
	int32_t lo = INT32_MAX, hi = INT32_MIN;
	for (uint32_t i = 0; i < nv; i++)
	{
		const int32_t dx = uv[i][0] - p[i].x, dy = uv[i][1] - p[i].y;
		lo = std::min(lo, std::min(dx, dy));
		hi = std::max(hi, std::max(dx, dy));
	}
	if (field(t1, 5, 1) ? (lo != 8 || hi != 8) : (lo < 0 || hi >= 16))
		return Feedback::None;
	depth = !color_fb;
	return Feedback::Pixel;
}

// gs/gs_interface.cpp:2473-2836 (one frame-buffer instance per render pass).
void RenderPassBuilder::append(const RegisterState &r, const QueuedVertex *vin, uint32_t n, bool quad)
{
	VertexPosition p[3] = { vin[0].pos, vin[1].pos, vin[2].pos };
	VertexAttribute a[3] = { vin[0].attr, vin[1].attr, vin[2].attr };
	const uint64_t prim = r.prim;
	const uint32_t ctx = field(prim, 9, 1);
	const bool is_line = !quad && n == 2;
	const uint32_t type = field(prim, 0, 3);
	const bool aa1 = field(prim, 7, 1) && type != 6 && type != 0;

	int32_t lox = std::min(p[0].x, p[1].x), loy = std::min(p[0].y, p[1].y);
	int32_t hix = std::max(p[0].x, p[1].x), hiy = std::max(p[0].y, p[1].y);
	if (!quad && !is_line)
	{
		lox = std::min(lox, p[2].x); loy = std::min(loy, p[2].y);
		hix = std::max(hix, p[2].x); hiy = std::max(hiy, p[2].y);
	}
	const int32_t pre_lox = lox, pre_loy = loy, pre_hix = hix, pre_hiy = hiy;
	hix -= 1; hiy -= 1;
	lox += (1 << (4 - int32_t(rate_log2_))) - 1;
	loy += (1 << (4 - int32_t(rate_log2_))) - 1;
	lox >>= 4; loy >>= 4; hix >>= 4; hiy >>= 4;
	if (is_line || aa1)
	{
		lox -= 1; loy -= 1; hix += 1; hiy += 1;
	}
	const uint64_t sc = r.scissor[ctx];
	lox = std::max(lox, int32_t(field(sc, 0, 11)));
	loy = std::max(loy, int32_t(field(sc, 32, 11)));
	hix = std::min(hix, int32_t(field(sc, 16, 11)));
	hiy = std::min(hiy, int32_t(field(sc, 48, 11)));
	if (hix < lox || hiy < loy)
		return;

	if (field(prim, 4, 1) && host)
		host->before_texture_read();

	// Frame / depth buffer change ends the render pass.
	const uint64_t frame = r.frame[ctx], zbuf = r.zbuf[ctx];
	if (!prims_.empty())
	{
		if ((frame_ & 0x3f3f01ffull) != (frame & 0x3f3f01ffull) ||
		    ((zbuf_ & 0x0f0001ffull) != (zbuf & 0x0f0001ffull) && pass_z_))
			flush();
	}

	bool candidate = false;
	int32_t order[3] = { 0, 1, 2 };
	if (n == 3)
	{
		candidate = parallelogram_candidate(p, a, pre_lox, pre_loy, pre_hix, pre_hiy, prim, order);
		if (!state_dirty_ && candidate && last_candidate_ && !prims_.empty() &&
		    triangles_form_parallelogram(p, a, order, &pos_[pos_.size() - 3], &attr_[attr_.size() - 3], last_order_,
		                                 prim))
		{
			prims_.back().state |= ST_PARALLELOGRAM | (uint32_t(last_order_[0]) << ST_PROVOKING_SHIFT);
			last_candidate_ = false;
			return;
		}
		memcpy(last_order_, order, sizeof(order));
	}

	// A texture read of pages the pending pass writes renders the pass first, so the texture
	// sees every earlier primitive (the ordering gs/page_tracker.cpp enforces). Fused
	// parallelograms above skip the check, as in gs/gs_interface.cpp:2570-2599.
	const bool tme = field(prim, 4, 1);
	int32_t slot = -1;
	Feedback feedback = Feedback::None;
	bool feedback_depth = false;
	if (tme)
	{
		const int32_t pbb[4] = { lox, loy, hix, hiy };
		feedback = feedback_mode(r, ctx, p, a, n, quad, pbb, feedback_depth);
	}
	if (feedback == Feedback::Pixel)
	{
		// One feedback kind and format per pass (gs/gs_interface.cpp:909-933).
		const uint64_t t0 = r.tex0[ctx];
		const uint32_t kind = feedback_depth ? 2u : 1u, psm = field(t0, 20, 6);
		const bool palette = psm == PSMT8 || psm == PSMT4 || psm == PSMT8H || psm == PSMT4HL || psm == PSMT4HH;
		const uint32_t cpsm = palette ? field(t0, 51, 4) : 0u;
		if (!prims_.empty() && pass_feedback_ &&
		    (kind != pass_feedback_ || psm != pass_feedback_psm_ || cpsm != pass_feedback_cpsm_))
			flush();
	}
	else if (tme)
	{
		// Bypass reads the texture as it is when the pass starts, whatever the pass writes.
		if (!prims_.empty() && feedback == Feedback::None && texture_hazard(r, ctx))
			flush();
		slot = texture_slot(r, ctx, type == 6);
		if (slot < 0)
		{
			flush();
			slot = texture_slot(r, ctx, type == 6);
		}
	}

	bool expands = prims_.empty();
	if (prims_.empty())
	{
		if (host)
			host->before_first_primitive();
		frame_ = frame;
		zbuf_ = zbuf;
		bb_[0] = lox; bb_[1] = loy; bb_[2] = hix; bb_[3] = hiy;
		color_write_mask_ = 0;
	}
	zbuf_ = (zbuf_ & ~0x0f0001ffull) | (zbuf & 0x0f0001ffull);
	expands = expands || lox < bb_[0] || loy < bb_[1] || hix > bb_[2] || hiy > bb_[3];
	bb_[0] = std::min(bb_[0], lox); bb_[1] = std::min(bb_[1], loy);
	bb_[2] = std::max(bb_[2], hix); bb_[3] = std::max(bb_[3], hiy);

	const uint64_t test = r.test[ctx], alpha = r.alpha[ctx];
	const uint32_t ate = field(test, 0, 1), atst = field(test, 1, 3), aref = field(test, 4, 8);
	const uint32_t afail = field(test, 12, 2), date = field(test, 14, 1), datm = field(test, 15, 1);
	const uint32_t zte = field(test, 16, 1), ztst = field(test, 17, 2);
	const bool abe = field(prim, 6, 1);

	// State vector (gs/gs_interface.cpp:1065-1128).
	StateVector sv{};
	if (field(r.dthe, 0, 1))
	{
		sv.blend_mode |= BLEND_DTHE;
		sv.dimx0 = uint32_t(r.dimx);
		sv.dimx1 = uint32_t(r.dimx >> 32);
	}
	if (ate && atst != 1 && !(atst == 0 && afail == 1))
		sv.blend_mode |= BLEND_ATE | (atst << 2) | (afail << 5);
	sv.blend_mode |= date ? BLEND_DATE : 0;
	sv.blend_mode |= datm ? BLEND_DATM : 0;
	if (abe || aa1)
		sv.blend_mode |= field(alpha, 0, 2) << 11 | field(alpha, 2, 2) << 13 | field(alpha, 4, 2) << 15 |
		                 field(alpha, 6, 2) << 17;
	if (abe)
		sv.blend_mode |= BLEND_ABE | (field(r.pabe, 0, 1) ? BLEND_PABE : 0);
	sv.blend_mode |= field(r.colclamp, 0, 1) ? BLEND_COLCLAMP : 0;
	sv.blend_mode |= field(r.fba[ctx], 0, 1) ? BLEND_FBA : 0;
	if (tme)
	{
		const uint64_t t0 = r.tex0[ctx];
		sv.combiner |= COMB_TME | (field(t0, 34, 1) ? COMB_TCC : 0) | field(t0, 35, 2);
	}
	sv.combiner |= field(prim, 5, 1) ? COMB_FOG : 0;

	PrimitiveAttribute pa{};
	uint32_t state = state_index(sv);
	if (feedback == Feedback::Pixel)
	{
		// The texel is the pixel's own colour or depth (ubershader.comp:955-1040): the index
		// carries the feedback bit and CSA, the word the formats, tex2 the TEXA expansion.
		const uint64_t t0 = r.tex0[ctx];
		const uint32_t psm = field(t0, 20, 6);
		const bool palette = psm == PSMT8 || psm == PSMT4 || psm == PSMT8H || psm == PSMT4HL || psm == PSMT4HH;
		uint32_t cpsm = palette ? field(t0, 51, 4) : 0u;
		if (cpsm == PSMCT24)
			cpsm = PSMCT32;
		pa.tex = TEX_FEEDBACK | field(t0, 56, 5) | psm << TEX_FEEDBACK_PSM_SHIFT | cpsm << TEX_FEEDBACK_CPSM_SHIFT |
		         (feedback_depth ? TEX_FEEDBACK_DEPTH : 0u);
		// The sampler bits stay as for a texture: triangle setup reads the magnification filter.
		const uint64_t t1 = r.tex1[ctx], cl = r.clamp[ctx];
		pa.tex |= field(t1, 5, 1) ? TEX_MAG_LINEAR : 0;
		pa.tex |= (field(cl, 0, 2) == 1 || field(cl, 0, 2) == 2) ? TEX_CLAMP_S : 0;
		pa.tex |= (field(cl, 2, 2) == 1 || field(cl, 2, 2) == 2) ? TEX_CLAMP_T : 0;
		pa.tex2 = TEX2_FIXED_LOD | (field(r.texa, 15, 1) ? TEX2_FEEDBACK_AEM : 0u) |
		          field(r.texa, 0, 8) << TEX2_FEEDBACK_TA0_SHIFT | field(r.texa, 32, 8) << TEX2_FEEDBACK_TA1_SHIFT;
		if (!pass_feedback_)
		{
			pass_feedback_ = feedback_depth ? 2u : 1u;
			pass_feedback_psm_ = psm;
			pass_feedback_cpsm_ = palette ? field(t0, 51, 4) : 0u;
		}
		pass_z_feedback_ = pass_z_feedback_ || feedback_depth;
	}
	else if (tme)
	{
		const uint64_t t1 = r.tex1[ctx], cl = r.clamp[ctx];
		pa.tex = uint32_t(slot);
		pa.tex |= field(t1, 5, 1) ? TEX_MAG_LINEAR : 0;
		const uint32_t wms = field(cl, 0, 2), wmt = field(cl, 2, 2);
		pa.tex |= (wms == 1 || wms == 2) ? TEX_CLAMP_S : 0;
		pa.tex |= (wmt == 1 || wmt == 2) ? TEX_CLAMP_T : 0;
		// Mipmapped sampling (gs/gs_interface.cpp:1834-1863): LCM, L and K go to tex2, MXL and
		// the MMIN filters to the texture word; otherwise the LOD is fixed at 0.
		if (mmin_has_mipmap(t1) && !disable_mipmaps)
		{
			pa.tex2 = field(t1, 0, 1) | field(t1, 19, 2) << TEX2_L_SHIFT | field(t1, 32, 12) << TEX2_K_SHIFT;
			pa.tex |= field(t1, 2, 3) << TEX_MAX_MIP_LEVEL_SHIFT;
			const uint32_t mmin = field(t1, 6, 3);
			pa.tex |= (mmin == 4 || mmin == 5) ? TEX_MIN_LINEAR : 0;
			pa.tex |= (mmin == 3 || mmin == 5) ? TEX_MIP_LINEAR : 0;
		}
		else
			pa.tex2 = TEX2_FIXED_LOD;
		// A super-sampled texture is sampled per sample; palette formats with REGION_REPEAT
		// map each sample to its own layer (gs/gs_interface.cpp:1783-1790, 1824-1831).
		if (slot_keys_[uint32_t(slot)].samples)
		{
			const uint32_t psm = field(r.tex0[ctx], 20, 6);
			const bool palette = psm == PSMT8 || psm == PSMT4 || psm == PSMT8H || psm == PSMT4HL || psm == PSMT4HH;
			pa.tex |= TEX_ARRAYED | TEX_PER_SAMPLE;
			if (palette && (wms == 3 || wmt == 3))
				pa.tex |= TEX_SAMPLE_MAPPING;
			pass_arrays_ = true;
		}
	}
	if (zte)
	{
		if (ztst > 1)
			state |= ST_Z_TEST | (ztst == 3 ? ST_Z_GREATER : 0);
		const bool implied_z_mask = ate && atst == 0 && afail != 2;
		if (!implied_z_mask && field(zbuf, 32, 1) == 0)
			state |= ST_Z_WRITE;
	}
	bool needs_prev = false;
	if (abe || aa1)
		needs_prev = field(alpha, 0, 2) == 1 || field(alpha, 2, 2) == 1 || field(alpha, 4, 2) == 1 ||
		             field(alpha, 6, 2) == 1;
	if (date)
		needs_prev = true;
	else if (ate && atst != 1)
	{
		const bool can_write_z = (state & ST_Z_WRITE) != 0;
		if (afail != 1 || (can_write_z && afail != 2))
			needs_prev = true;
	}
	// Reading the pixel's own value is a kind of blending (gs/gs_interface.cpp:1928-1931).
	if (feedback == Feedback::Pixel)
		needs_prev = true;
	if (!needs_prev)
		state |= ST_OPAQUE;
	if (aa1)
		state |= ST_MULTISAMPLE;
	const uint32_t msk = field(r.scanmsk, 0, 2);
	if (msk >= 2)
	{
		state |= 1u << (23 + msk - 2);
		// A scan-masked primitive taller than 400 lines means the frame depends on field order:
		// the next 4 vsyncs do not force progressive output (gs/gs_interface.cpp:1947-1960).
		if (loy + 400 < hiy)
			complex_scanmsk_timeout = 4;
	}
	if (!field(prim, 8, 1))
		state |= ST_PERSPECTIVE;
	if (field(prim, 3, 1))
		state |= ST_IIP;
	if (field(prim, 10, 1))
		state |= ST_FIX;
	if (quad)
	{
		state |= ST_PARALLELOGRAM | ST_SPRITE | ST_SNAP_RASTER | ST_SNAP_ATTRIBUTE;
		state &= ~ST_MULTISAMPLE;
	}
	else if (is_line)
	{
		state |= ST_PARALLELOGRAM | ST_LINE;
		if (state & ST_MULTISAMPLE)
			state &= ~ST_Z_WRITE;
	}
	if (n == 1)
		state |= ST_FIX | ST_SNAP_RASTER | ST_SNAP_ATTRIBUTE;

	pa.state = state;
	pa.fbmsk = uint32_t(frame >> 32);
	pa.fogcol = uint32_t(r.fogcol);
	pa.alpha = (field(alpha, 32, 8) << 8) | aref;
	pa.bb[0] = int16_t(lox); pa.bb[1] = int16_t(loy); pa.bb[2] = int16_t(hix); pa.bb[3] = int16_t(hiy);
	color_write_mask_ |= ~pa.fbmsk;

	pass_z_ |= (state & (ST_Z_TEST | ST_Z_WRITE)) != 0;
	pass_zwrite_ |= (state & ST_Z_WRITE) != 0;
	prims_.push_back(pa);
	// A draw that grows the pass's box marks the box's frame-buffer pages (and depth pages,
	// once the pass writes depth) as possibly super-sampled (gs/gs_interface.cpp:2786-2816).
	if (expands && bb_[2] >= bb_[0] && bb_[3] >= bb_[1])
	{
		const uint32_t x0 = uint32_t(bb_[0]), y0 = uint32_t(bb_[1]), x1 = uint32_t(bb_[2]), y1 = uint32_t(bb_[3]);
		writes_.mark_super_sample(page_rect(field(frame_, 0, 9) * 32u, field(frame_, 16, 6), field(frame_, 24, 6), x0,
		                                    y0, x1, y1),
		                          true);
		if (pass_zwrite_)
			writes_.mark_super_sample(page_rect(field(zbuf_, 0, 9) * 32u, field(frame_, 16, 6),
			                                    0x30u | field(zbuf_, 24, 4), x0, y0, x1, y1),
			                          true);
	}
	pos_.insert(pos_.end(), p, p + 3);
	attr_.insert(attr_.end(), a, a + 3);
	last_candidate_ = candidate;
	state_dirty_ = false;
	if (prims_.size() >= 65536u || states_.size() >= 4096u)
		flush();
}

void RenderPassBuilder::flush()
{
	if (!prims_.empty())
	{
		RenderPassDesc d{};
		d.positions = pos_.data();
		d.attributes = attr_.data();
		d.prims = prims_.data();
		d.states = states_.data();
		d.num_primitives = uint32_t(prims_.size());
		d.num_states = uint32_t(states_.size());
		d.fb_psm = field(frame_, 24, 6);
		d.fbp = field(frame_, 0, 9);
		d.fbw = field(frame_, 16, 6);
		d.zbp = field(zbuf_, 0, 9);
		uint32_t zpsm = field(zbuf_, 24, 4) | 0x30u;
		// deduce_effectice_z_psm (gs/gs_renderer.cpp:2147-2170)
		if (d.fb_psm == PSMZ32 || d.fb_psm == PSMZ24 || d.fb_psm == PSMZ16S)
		{
			if (zpsm == PSMZ32) zpsm = PSMCT32;
			else if (zpsm == PSMZ24) zpsm = PSMCT24;
			else if (zpsm == PSMZ16S || zpsm == PSMZ16) zpsm = PSMCT16S;
		}
		d.z_psm = pass_z_ ? zpsm : 0xffffffffu;
		memcpy(d.bb, bb_, sizeof(bb_));
		d.opaque_fbmask = ~color_write_mask_;
		d.rate_log2 = rate_log2_;
		d.z_sensitive = pass_z_;
		d.z_write = pass_zwrite_;
		d.z_feedback = pass_z_feedback_;
		d.arrays = pass_arrays_;
		if (pass_arrays_)
			mark_blur_kernels();
		d.zbuf_psm = field(zbuf_, 24, 4) | 0x30u;
		d.decodes = decodes_.data();
		d.tex_descs = tex_descs_.data();
		d.num_textures = uint32_t(decodes_.size());
		d.tex_groups = group_total_;
		const uint32_t mode = renderer_->sampler_feedback_mode();
		d.sparse_decode = mode == 1u || (mode == 0u && texel_total_ >= Renderer::kFeedbackAutoTexels);
		renderer_->submit_render_pass(d);

		// Whole decoded images join the cache (a sparse decode holds only this pass's blocks),
		// stamped before this pass's own frame-buffer and depth writes are logged.
		uint32_t decoded = 0;
		for (SlotImage &im : slot_images_)
			if (im.decoded)
			{
				decoded++;
				if (!d.sparse_decode)
				{
					im.entry.decoded_at = writes_.now();
					cache_.push_back(im.entry);
				}
			}
		renderer_->note_textures(decoded, uint32_t(slot_images_.size()) - decoded);
		if (bb_[2] >= bb_[0] && bb_[3] >= bb_[1])
		{
			const uint32_t x0 = uint32_t(bb_[0]), y0 = uint32_t(bb_[1]), x1 = uint32_t(bb_[2]), y1 = uint32_t(bb_[3]);
			writes_.write(page_span(d.fbp * 32u, d.fbw, d.fb_psm, x0, y0, x1, y1));
			if (d.z_write)
				writes_.write(page_span(d.zbp * 32u, d.fbw, d.zbuf_psm, x0, y0, x1, y1));
		}
		prims_.clear();
		slot_keys_.clear();
		slot_images_.clear();
		decodes_.clear();
		tex_descs_.clear();
		texel_total_ = group_total_ = 0;
		pass_z_ = pass_zwrite_ = false;
		pass_feedback_ = 0;
		pass_feedback_psm_ = pass_feedback_cpsm_ = 0;
		pass_z_feedback_ = false;
		pass_arrays_ = false;
		pos_.clear();
		attr_.clear();
		states_.clear();
		state_map_.clear();
	}
	state_dirty_ = true;
	last_candidate_ = false;
}
}

namespace pgs
{
// A pass of at most 64 primitives that samples one super-sampled texture at two different
// phases (first vertex UV less position) is taken for a blur (gs/gs_renderer.cpp:1848-1885);
// then its primitives that blend Cs * alpha + Cd without a depth test are blur taps, which
// triangle setup keeps off per-sample sampling (triangle_setup.comp:705-757).
void RenderPassBuilder::mark_blur_kernels()
{
	if (prims_.size() > 64u)
		return;
	uint32_t last = ~0u;
	int32_t phase_u = 0, phase_v = 0;
	bool blur = false;
	for (size_t i = 0; i < prims_.size() && !blur; i++)
	{
		const PrimitiveAttribute &p = prims_[i];
		if ((p.tex & (TEX_PER_SAMPLE | TEX_FEEDBACK)) != TEX_PER_SAMPLE)
			continue;
		if (p.state & ST_PERSPECTIVE)
			return;
		const uint32_t index = p.tex & TEX_INDEX_MASK;
		const int32_t pu = int32_t(attr_[3 * i].uv & 0xffffu) - pos_[3 * i].x;
		const int32_t pv = int32_t(attr_[3 * i].uv >> 16) - pos_[3 * i].y;
		if (index == last)
			blur = pu != phase_u || pv != phase_v;
		else
		{
			phase_u = pu;
			phase_v = pv;
		}
		last = index;
	}
	if (!blur)
		return;
	for (PrimitiveAttribute &p : prims_)
	{
		const uint32_t bm = states_[p.state & STATE_INDEX_MASK].blend_mode;
		const bool taps = (bm & BLEND_ABE) && !(p.state & ST_Z_TEST) && ((bm >> 11) & 3u) == 0u &&
		                  ((bm >> 13) & 3u) == 2u && ((bm >> 17) & 3u) == 1u;
		if ((p.tex & (TEX_PER_SAMPLE | TEX_FEEDBACK)) == TEX_PER_SAMPLE && taps)
			p.tex |= TEX_BLUR_KERNEL;
	}
}

RenderPassBuilder::SlotKey RenderPassBuilder::slot_key(const RegisterState &r, uint32_t ctx, bool sprite) const
{
	SlotKey k{};
	k.tex0 = r.tex0[ctx] & ((1ull << 61) - 1ull); // CLD does not change the texture
	k.clamp = r.clamp[ctx];
	k.texa = r.texa;
	k.levels = texture_levels(k.tex0, r.tex1[ctx], disable_mipmaps);
	// Levels 1..3 in MIPTBP1, 4..6 in MIPTBP2, 20 bits each.
	const uint32_t l1 = std::min(k.levels - 1u, 3u), l2 = k.levels > 4u ? k.levels - 4u : 0u;
	k.mip1 = r.miptbp1[ctx] & ((1ull << (20u * l1)) - 1ull);
	k.mip2 = r.miptbp2[ctx] & ((1ull << (20u * l2)) - 1ull);
	// A sprite's single-level texture of at least 8 bits per texel whose first page was last
	// written by a render pass is super-sampled (gs/gs_interface.cpp:1593-1603; the first page of
	// the level-0 span is the page rectangle's base page, gs/page_tracker.cpp:147-153).
	k.samples = 0;
	if (sprite && super_sampled_textures && sample_log2_ != 0 && k.levels == 1u && psm_bpp(field(k.tex0, 20, 6)) >= 8u)
	{
		const TexLayout l = tex_layout(k.tex0, k.clamp, 1u, 0u, 0u);
		if (writes_.may_super_sample(texture_span(l).s[0].first))
			k.samples = 1u << sample_log2_;
	}
	return k;
}

bool RenderPassBuilder::texture_hazard(const RegisterState &r, uint32_t ctx)
{
	const SlotKey k = slot_key(r, ctx);
	if (!(k == span_key_))
	{
		span_key_ = k;
		span_ = texture_span(tex_layout(k.tex0, k.clamp, k.levels, k.mip1, k.mip2));
	}
	const uint32_t x0 = uint32_t(bb_[0]), y0 = uint32_t(bb_[1]), x1 = uint32_t(bb_[2]), y1 = uint32_t(bb_[3]);
	const PageSpan fb = page_span(field(frame_, 0, 9) * 32u, field(frame_, 16, 6), field(frame_, 24, 6), x0, y0, x1, y1);
	if (span_.overlaps(fb))
		return true;
	if (pass_zwrite_)
	{
		const PageSpan z = page_span(field(zbuf_, 0, 9) * 32u, field(frame_, 16, 6), 0x30u | field(zbuf_, 24, 4), x0,
		                             y0, x1, y1);
		if (span_.overlaps(z))
			return true;
	}
	return false;
}

// Image arena as a ring: the next image goes after the last one, or at the start when it
// does not fit before the end. Cached images it overlaps are dropped; an image of the current
// pass must not be overwritten, so then the pass has to end first (false).
bool RenderPassBuilder::allocate_image(uint32_t texels, uint32_t &offset)
{
	uint32_t off = ring_;
	if (uint64_t(off) + texels > kTexImageTexels)
		off = 0;
	auto overlaps = [&](uint32_t o, uint32_t n) { return o < off + texels && off < o + n; };
	for (const SlotImage &im : slot_images_)
		if (overlaps(im.offset, im.texels))
			return false;
	cache_.erase(std::remove_if(cache_.begin(), cache_.end(),
	                            [&](const CachedTexture &c) { return overlaps(c.offset, c.texels); }),
	             cache_.end());
	ring_ = off + texels;
	offset = off;
	return true;
}

int32_t RenderPassBuilder::texture_slot(const RegisterState &r, uint32_t ctx, bool sprite)
{
	const SlotKey key = slot_key(r, ctx, sprite);
	for (uint32_t i = 0; i < slot_keys_.size(); i++)
		if (slot_keys_[i] == key)
			return int32_t(i);
	const uint64_t key0 = key.tex0, keyc = key.clamp, keya = key.texa;

	const TexLayout l = tex_layout(key0, keyc, key.levels, key.mip1, key.mip2);
	// A super-sampled texture holds samples + 1 layers of its single level.
	const uint32_t texels = l.texels() * (key.samples ? key.samples + 1u : 1u);
	if (slot_keys_.size() >= kMaxPassTextures || texels > kTexImageTexels)
		return prims_.empty() ? -2 : -1;

	// A valid cached image of an earlier pass is reused; otherwise the pass decodes one.
	const bool palette = l.psm == PSMT8 || l.psm == PSMT4 || l.psm == PSMT8H || l.psm == PSMT4HL || l.psm == PSMT4HH;
	SlotImage im{};
	bool hit = false;
	for (size_t i = 0; i < cache_.size(); i++)
	{
		const CachedTexture &c = cache_[i];
		if (!(c.key == key))
			continue;
		bool stale = false;
		for (uint32_t lv = 0; lv < c.span.n; lv++)
			stale = stale || writes_.written_since(c.span.s[lv], c.decoded_at);
		if (!stale && (!c.palette || c.clut_gen == writes_.clut_generation()))
		{
			im.offset = c.offset;
			im.texels = c.texels;
			hit = true;
		}
		else
			cache_.erase(cache_.begin() + std::ptrdiff_t(i));
		break;
	}
	if (!hit)
	{
		if (!allocate_image(texels, im.offset))
			return prims_.empty() ? -2 : -1;
		im.texels = texels;
		im.decoded = true;
		im.entry = { key, im.offset, texels, 0, writes_.clut_generation(), palette, texture_span(l) };
	}

	// TexInfo (gs/gs_interface.cpp:1752-1781).
	const uint32_t wms = field(keyc, 0, 2), wmt = field(keyc, 2, 2);
	TextureDesc td{};
	td.image_offset = im.offset;
	td.width = l.width;
	td.height = l.height;
	td.levels = l.levels;
	// Clamp bounds floor(region) + 0.5 (the regions are whole texels, so the sum is exact).
	const float inf = std::numeric_limits<float>::infinity();
	td.region[0] = td.region[1] = -inf;
	td.region[2] = td.region[3] = inf;
	if (wms == 1) { td.region[0] = 0.5f; td.region[2] = float(l.width - 1u) + 0.5f; }
	else if (wms == 2) { td.region[0] = float(field(keyc, 4, 10)) + 0.5f; td.region[2] = float(field(keyc, 14, 10)) + 0.5f; }
	if (wmt == 1) { td.region[1] = 0.5f; td.region[3] = float(l.height - 1u) + 0.5f; }
	else if (wmt == 2) { td.region[1] = float(field(keyc, 24, 10)) + 0.5f; td.region[3] = float(field(keyc, 34, 10)) + 0.5f; }
	td.region[0] = std::min(td.region[0], td.region[2]);
	td.region[1] = std::min(td.region[1], td.region[3]);
	td.fwidth = float(l.width);
	td.fheight = float(l.height);
	td.size[0] = float(1u << l.tw);
	td.size[1] = float(1u << l.th);
	td.size[2] = 1.0f / float(l.width);
	td.size[3] = 1.0f / float(l.height);
	td.bias[0] = -float(l.bx) * td.size[2];
	td.bias[1] = -float(l.by) * td.size[3];

	// Decode work; a reused image gets an empty entry (no workgroups), which keeps decode,
	// descriptor and sampler-feedback indices equal to the slot.
	TexDecode dd{};
	dd.image_offset = im.offset;
	dd.width = l.width;
	dd.height = l.height;
	dd.levels = l.levels;
	dd.texels = im.decoded ? texels : 0u;
	for (uint32_t lv = 1; lv < l.levels; lv++)
		dd.level_buf[lv - 1u] = l.level_tbp[lv] | l.level_tbw[lv] << 14;
	dd.tbp = l.tbp;
	dd.tbw = l.tbw;
	dd.psm = l.psm;
	dd.bx = l.bx; dd.by = l.by;
	dd.umsk = l.umsk; dd.ufix = l.ufix; dd.vmsk = l.vmsk; dd.vfix = l.vfix;
	const uint32_t cpsm = field(key0, 51, 4);
	dd.clut16 = (cpsm == PSMCT16 || cpsm == PSMCT16S) ? 1u : 0u;
	// A 32-bit CLUT addresses the lower 16 banks only (gs/gs_interface.cpp:1306-1311).
	dd.csa = dd.clut16 ? field(key0, 56, 5) : (field(key0, 56, 5) & 15u);
	dd.aem = field(keya, 15, 1);
	dd.ta0 = field(keya, 0, 8);
	dd.ta1 = field(keya, 32, 8);
	dd.group_begin = group_total_;
	dd.samples = key.samples;

	slot_keys_.push_back(key);
	slot_images_.push_back(im);
	decodes_.push_back(dd);
	tex_descs_.push_back(td);
	if (im.decoded)
	{
		texel_total_ += texels;
		group_total_ += (texels + 255u) / 256u;
	}
	return int32_t(slot_keys_.size() - 1);
}
}
