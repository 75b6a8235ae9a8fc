// CLUT loads and texture decode for gfx950 (replaces clut_write.comp and upload.comp).
//
// k_clut_write: one workgroup of 256 threads, thread i handles CLUT entry i of the load. The
// entries a load writes are distinct per thread (a 32-bit palette writes entry k and k + 256,
// both inside disjoint ranges), so every thread reads VRAM and stores directly.
//
// k_texture_decode: one launch decodes every texture of a render pass into RGBA8 images in a
// per-pass arena. Each workgroup of 256 threads decodes 256 consecutive texels of one texture
// (row-major), so image stores are coalesced; the texture of a workgroup is found from the
// per-texture first-workgroup table.
#include "gs_defs.hpp"
#include "gs_raster.hpp"
#include "texture_layout.hpp"
#include "pgs_kernels.h"
#include "multi_launch.hpp"
#include "multi_dispatch.hpp"

#include "texture_bodies.hpp"

namespace pgs
{
namespace
{
using texb::clut_write_body;
using texb::texture_decode_body;

// ---------------------------------------------------------------------------------------
// Sampler feedback (replaces sampler_feedback.comp:97-307; own bound, same purpose: the set of
// 8x8 texel blocks a pass's primitives can sample).
__device__ __forceinline__ int32_t fb_wadd(int32_t a, int32_t b) { return int32_t(uint32_t(a) + uint32_t(b)); }
__device__ __forceinline__ int32_t fb_idot3(int32_t e0, int32_t e1, int32_t e2, int32_t x, int32_t y)
{
	return fb_wadd(fb_wadd(int32_t(uint32_t(e0) * uint32_t(x)), int32_t(uint32_t(e1) * uint32_t(y))), e2);
}

// The shader's texel index for a clamped coordinate u1 (sample_texture: 1/16 snap, normalize,
// scale by the image size): monotone non-decreasing in u1.
__device__ __forceinline__ int32_t feedback_texel(float u1, float inv_size, float bias, float extent)
{
	float n = floorf(__fadd_rn(__fmul_rn(u1, 16.0f), 1.0f / 64.0f));
	n = __fadd_rn(__fmul_rn(__fmul_rn(__fadd_rn(n, 16.0f / 1024.0f), inv_size), 0.0625f), bias);
	return int32_t(floorf(fminf(fmaxf(__fmul_rn(n, extent), -1.0e6f), 1.0e6f)));
}

// Inclusive texel range [lo, hi] of an axis of `size` texels after the sampler's wrap: one or two
// ranges inside [0, size).
__device__ __forceinline__ uint32_t wrap_ranges(int32_t lo, int32_t hi, int32_t size, int32_t r[4])
{
	if (hi - lo + 1 >= size)
	{
		r[0] = 0; r[1] = size - 1;
		return 1;
	}
	int32_t start = lo % size;
	start += start < 0 ? size : 0;
	const int32_t end = start + (hi - lo);
	if (end < size)
	{
		r[0] = start; r[1] = end;
		return 1;
	}
	r[0] = start; r[1] = size - 1; r[2] = 0; r[3] = end - size;
	return 2;
}

__device__ __forceinline__ void sampler_feedback_body(const pgs_feedback_args &a, uint32_t wg, uint32_t)
{
	const uint32_t i = wg * 64u + threadIdx.x;
	if (i >= a.num_primitives)
		return;
	const PrimitiveAttribute &pa = static_cast<const PrimitiveAttribute *>(a.prims)[i];
	const uint32_t st = pa.state;
	const StateVector &sv = static_cast<const StateVector *>(a.states)[st & STATE_INDEX_MASK];
	if (!(sv.combiner & COMB_TME))
		return;
	const uint32_t slot = pa.tex & TEX_INDEX_MASK;
	if (slot >= a.num_textures)
		return;
	const TextureDesc td = static_cast<const TextureDesc *>(a.textures)[slot];
	if (td.width > 1024u || td.height > 1024u || td.width == 0 || td.height == 0 || td.levels > 1u)
		return; // decoded whole
	uint32_t *mask = a.masks + slot * PGS_FEEDBACK_WORDS;
	const int32_t w = int32_t(td.width), h = int32_t(td.height);
	int32_t xr[4], yr[4];
	uint32_t nx = 0, ny = 0;
	const PrimitiveSetup s = static_cast<const PrimitiveSetup *>(a.setup)[i];
	const TransformedAttributes &ta = static_cast<const TransformedAttributes *>(a.tattr)[i];
	const bool persp = (st & ST_PERSPECTIVE) != 0;
	bool all = false;
	float umin = 3.0e38f, umax = -3.0e38f, vmin = 3.0e38f, vmax = -3.0e38f;
	if (s.bb[2] < s.bb[0] || s.bb[3] < s.bb[1])
		return; // never shaded
	// Corners of the pixel box grown by one pixel, in the shader's 1/8-pixel units.
	const int32_t xs[2] = { (s.bb[0] - 1) * 8, (s.bb[2] + 1) * 8 + 7 };
	const int32_t ys[2] = { (s.bb[1] - 1) * 8, (s.bb[3] + 1) * 8 + 7 };
	const bool para = (s.a[0] & 1) != 0;
	const int32_t offb = int32_t((uint32_t(s.a[0]) >> 2) & 3u) - 1;
	const int32_t offc = int32_t((uint32_t(s.a[0]) >> 4) & 3u) - 1;
	for (uint32_t k = 0; k < 4; k++)
	{
		const int32_t sx = xs[k & 1], sy = ys[k >> 1];
		const int32_t A = fb_idot3(s.a[0] >> 6, s.a[1], s.a[2], sx, sy);
		int32_t B = fb_idot3(s.b[0], s.b[1], s.b[2], sx, sy);
		int32_t C = fb_idot3(s.c[0], s.c[1], s.c[2], sx, sy);
		// A parallelogram's second half uses shifted edges; take both forms at every corner.
		for (uint32_t form = 0; form < (para ? 2u : 1u); form++)
		{
			if (form == 1)
			{
				B = fb_wadd(B, fb_wadd(A, offb));
				C = fb_wadd(C, fb_wadd(A, offc));
			}
			const float fi = __fadd_rn(__fmul_rn(float(B), s.inv_area), s.error_i);
			const float fj = __fadd_rn(__fmul_rn(float(C), s.inv_area), s.error_j);
			float q[3];
			for (uint32_t c = 0; c < 3; c++)
				q[c] = __fadd_rn(__fadd_rn(ta.stqf0[c], __fmul_rn(ta.stqf1[c], fi)), __fmul_rn(ta.stqf2[c], fj));
			float u = q[0], v = q[1];
			if (persp)
			{
				// s / q is monotone between corners only while q keeps its sign.
				if (!(q[2] > 1.0e-20f))
					all = true;
				u = __fdiv_rn(u, q[2]);
				v = __fdiv_rn(v, q[2]);
			}
			if (!(fabsf(u) < 1.0e30f) || !(fabsf(v) < 1.0e30f))
				all = true;
			umin = fminf(umin, u); umax = fmaxf(umax, u);
			vmin = fminf(vmin, v); vmax = fmaxf(vmax, v);
		}
	}
	if (!all)
	{
		// The shader's chain: clamp to +-2047, scale (STQ), region clamp, snap, texel index.
		auto axis = [&](float lo, float hi, float size, float rlo, float rhi, float inv, float bias, int32_t extent,
		                int32_t r[4]) -> uint32_t {
			lo = fminf(fmaxf(lo, -2047.0f), 2047.0f);
			hi = fminf(fmaxf(hi, -2047.0f), 2047.0f);
			if (persp)
			{
				lo = __fmul_rn(lo, size);
				hi = __fmul_rn(hi, size);
			}
			// The descriptor's bounds are floor(region) + 0.5, infinite without clamping.
			lo = fminf(fmaxf(lo, rlo), rhi);
			hi = fminf(fmaxf(hi, rlo), rhi);
			// Bilinear reads x0 = floor(n * extent - 0.5) and x0 + 1; one texel more each way
			// absorbs the rounding of the corner evaluation.
			const int32_t tlo = feedback_texel(lo, inv, bias, float(extent)) - 2;
			const int32_t thi = feedback_texel(hi, inv, bias, float(extent)) + 2;
			return wrap_ranges(tlo, thi, extent, r);
		};
		nx = axis(umin, umax, td.size[0], td.region[0], td.region[2], td.size[2], td.bias[0], w, xr);
		ny = axis(vmin, vmax, td.size[1], td.region[1], td.region[3], td.size[3], td.bias[1], h, yr);
	}
	else
	{
		xr[0] = 0; xr[1] = w - 1; yr[0] = 0; yr[1] = h - 1;
		nx = ny = 1;
	}
	for (uint32_t yi = 0; yi < ny; yi++)
		for (int32_t by = yr[2 * yi] >> 3; by <= yr[2 * yi + 1] >> 3; by++)
			for (uint32_t xi = 0; xi < nx; xi++)
			{
				const int32_t bx0 = xr[2 * xi] >> 3, bx1 = xr[2 * xi + 1] >> 3;
				for (int32_t word = bx0 >> 5; word <= bx1 >> 5; word++)
				{
					const int32_t lo = max(bx0, word * 32) - word * 32, hi = min(bx1, word * 32 + 31) - word * 32;
					const uint32_t bits = (hi == 31 ? 0xffffffffu : ((1u << (hi + 1)) - 1u)) & ~((1u << lo) - 1u);
					atomicOr(&mask[by * 4 + word], bits);
				}
			}
}

__global__ void __launch_bounds__(256) k_clut_write(pgs_clut_args a) { clut_write_body(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) k_texture_decode(pgs_tex_decode_args a) { texture_decode_body(a, blockIdx.x, gridDim.x); }
PGS_DEFINE_MULTI(256, k_clut_write_multi, clut_write_body, pgs_clut_args)
PGS_DEFINE_MULTI(256, k_texture_decode_multi, texture_decode_body, pgs_tex_decode_args)
__global__ void __launch_bounds__(64) k_sampler_feedback(pgs_feedback_args a) { sampler_feedback_body(a, blockIdx.x, gridDim.x); }
PGS_DEFINE_MULTI(64, k_sampler_feedback_multi, sampler_feedback_body, pgs_feedback_args)
}
}

extern "C" int pgs_k_clut_write(const pgs_clut_args *args, hipStream_t stream)
{
	if (!args || !args->vram || !args->clut || args->csa > 31u || args->cbp > 0x3fffu)
		return -int(hipErrorInvalidValue);
	hipLaunchKernelGGL(pgs::k_clut_write, dim3(1), dim3(256), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

extern "C" int pgs_k_sampler_feedback(const pgs_feedback_args *args, hipStream_t stream)
{
	if (!args || !args->setup || !args->prims || !args->tattr || !args->states || !args->textures || !args->masks)
		return -int(hipErrorInvalidValue);
	if (args->num_primitives == 0 || args->num_textures == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_sampler_feedback, dim3((args->num_primitives + 63u) / 64u), dim3(64), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

extern "C" int pgs_k_texture_decode(const pgs_tex_decode_args *args, hipStream_t stream)
{
	if (!args || !args->vram || !args->clut || !args->images || !args->decodes)
		return -int(hipErrorInvalidValue);
	if (args->num_textures == 0 || args->num_groups == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_texture_decode, dim3(args->num_groups), dim3(256), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

uint32_t pgs_grid_clut_write(const pgs_clut_args &) { return 1u; }
uint32_t pgs_grid_texture_decode(const pgs_tex_decode_args &a) { return a.num_textures ? a.num_groups : 0u; }
uint32_t pgs_grid_sampler_feedback(const pgs_feedback_args &a)
{
	return a.num_textures ? (a.num_primitives + 63u) / 64u : 0u;
}

int pgs_multi_sampler_feedback(const pgs_feedback_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                               hipStream_t stream)
{
	hipLaunchKernelGGL(pgs::k_sampler_feedback_multi, dim3(total), dim3(64), 0, stream, args, prefix, n);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

int pgs_multi_clut_write(const pgs_clut_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                         hipStream_t stream)
{
	hipLaunchKernelGGL(pgs::k_clut_write_multi, dim3(total), dim3(256), 0, stream, args, prefix, n);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

int pgs_multi_texture_decode(const pgs_tex_decode_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                             hipStream_t stream)
{
	hipLaunchKernelGGL(pgs::k_texture_decode_multi, dim3(total), dim3(256), 0, stream, args, prefix, n);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}
