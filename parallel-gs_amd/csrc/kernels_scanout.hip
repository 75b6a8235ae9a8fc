// CRTC scanout kernels for gfx950: read-circuit sampling, circuit merge, EXTWRITE feedback and
// the FastMAD deinterlacer.
//
// The reference draws these as full-screen Vulkan fragment passes into RGBA8 images
// (gs/gs_renderer.cpp:4055-5099). Here each stage is one compute launch with one lane per
// output pixel: a workgroup covers 64 x 4 pixels, so every wave writes 256 contiguous bytes of
// an image row. The frames are small (a 640 x 448 frame is 1.1 MiB), so the stages are
// latency-bound; they run back to back on the context's stream after the frame's rendering.
//
// Float arithmetic is written with explicit round-to-nearest operations in the order of the
// reference's shader expressions, so a CPU restatement with the same order gives the same bits.
#include "gs_defs.hpp"
#include "pgs_kernels.h"

namespace pgs
{
namespace
{
constexpr uint32_t kTileW = 64, kTileH = 4;

__device__ __forceinline__ float unorm(uint32_t c, uint32_t ch)
{
	return __fdiv_rn(float((c >> (8u * ch)) & 0xffu), 255.0f);
}

// packUnorm: round(clamp(x, 0, 1) * 255), ties to even.
__device__ __forceinline__ uint32_t to_unorm8(float x)
{
	return uint32_t(rintf(__fmul_rn(fminf(fmaxf(x, 0.0f), 1.0f), 255.0f)));
}

__device__ __forceinline__ uint32_t pack4(const float c[4])
{
	return to_unorm8(c[0]) | (to_unorm8(c[1]) << 8) | (to_unorm8(c[2]) << 16) | (to_unorm8(c[3]) << 24);
}

// One pixel of a plane as RGBA8 (sample_circuit.frag:50-71): 16-bit formats expand 5:5:5 with
// alpha 0xff where bit 15 is set and 0 otherwise; 32-bit formats other than PSMCT32 read alpha
// 0x80.
__device__ __forceinline__ uint32_t circuit_texel(const uint8_t *__restrict__ plane, uint32_t addr, bool is16,
                                                  bool ct32)
{
	if (is16)
	{
		const uint32_t h = *reinterpret_cast<const uint16_t *>(plane + addr);
		return ((h & 0x1fu) << 3) | (((h >> 5) & 0x1fu) << 11) | (((h >> 10) & 0x1fu) << 19) |
		       ((h & 0x8000u) ? 0xff000000u : 0u);
	}
	uint32_t w = *reinterpret_cast<const uint32_t *>(plane + addr);
	if (!ct32)
		w = (w & 0x00ffffffu) | 0x80000000u;
	return w;
}

__global__ void __launch_bounds__(256) k_scanout_circuit(pgs_scanout_circuit_args a)
{
	const uint32_t x = blockIdx.x * kTileW + (threadIdx.x & (kTileW - 1u));
	const uint32_t y = blockIdx.y * kTileH + (threadIdx.x / kTileW);
	if (x >= a.width || y >= a.height)
		return;
	uint32_t out = 0;
	if (x < a.valid_width && y < a.valid_height)
	{
		const uint32_t S = a.super_samples;
		const uint32_t sx = S >= 4u ? x >> 1 : x, sy = S >= 4u ? y >> 1 : y;
		const uint32_t cx = sx + a.dbx, cy = sy * a.phase_stride + a.dby + a.phase;
		const bool is16 = psm_is_16(a.psm), ct32 = a.psm == PSMCT32;
		const uint32_t addr = swizzle_addr(cx, cy, a.base_block, a.fbw, a.psm);
		bool valid = false;
		if (S >= 4u)
		{
			const uint8_t *ref = a.vram + kVramSize;
			if (is16)
				valid = *reinterpret_cast<const uint16_t *>(ref + addr) == 0xffffu;
			else
			{
				const uint32_t w = *reinterpret_cast<const uint32_t *>(ref + addr);
				valid = ct32 ? w == 0xffffffffu : (w & 0x00ffffffu) == 0x00ffffffu;
			}
		}
		if (valid)
		{
			// The quad's share of the samples: NUM = S / 4 consecutive planes from
			// 2 + NUM * quad; the 8x checkerboard numbers the quad x-major.
			const uint32_t num = S / 4u;
			const uint32_t quad = S != 8u ? (y & 1u) + (x & 1u) * 2u : (x & 1u) + (y & 1u) * 2u;
			const uint32_t base = 2u + num * quad;
			float f[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
			for (uint32_t i = 0; i < num; i++)
			{
				const uint32_t t = circuit_texel(a.vram + size_t(base + i) * kVramSize, addr, is16, ct32);
				for (uint32_t ch = 0; ch < 4; ch++)
					f[ch] = __fadd_rn(f[ch], unorm(t, ch));
			}
			for (uint32_t ch = 0; ch < 4; ch++)
				f[ch] = __fdiv_rn(f[ch], float(num));
			out = pack4(f);
		}
		else
			out = circuit_texel(a.vram, addr, is16, ct32);
	}
	a.image[size_t(y) * a.width + x] = out;
}

// blit_circuit.frag at merged pixel (px, py) inside viewport (vx, vy, vw, vh): the texture
// coordinate of the pixel centre, sampled linear-clamp with 8-bit sub-texel weights. The
// filtered value of a channel is acc / (255 * 65536) with acc an exact integer.
__device__ __forceinline__ void bilinear(const uint32_t *__restrict__ img, uint32_t w, uint32_t h, int32_t px,
                                         int32_t py, int32_t vx, int32_t vy, uint32_t vw, uint32_t vh, float out[4])
{
	auto axis = [](int32_t p, int32_t v, uint32_t vsize, uint32_t size, int32_t &i0, int32_t &i1, uint32_t &wt) {
		const float u = __fdiv_rn(__fadd_rn(float(p - v), 0.5f), float(vsize));
		const float s = __fsub_rn(__fmul_rn(u, float(size)), 0.5f);
		const int32_t q = int32_t(floorf(__fadd_rn(__fmul_rn(s, 256.0f), 0.5f)));
		const int32_t i = q >> 8;
		wt = uint32_t(q & 255);
		i0 = min(max(i, 0), int32_t(size) - 1);
		i1 = min(max(i + 1, 0), int32_t(size) - 1);
	};
	int32_t x0, x1, y0, y1;
	uint32_t ax, ay;
	axis(px, vx, vw, w, x0, x1, ax);
	axis(py, vy, vh, h, y0, y1, ay);
	const uint32_t t00 = img[size_t(y0) * w + x0], t10 = img[size_t(y0) * w + x1];
	const uint32_t t01 = img[size_t(y1) * w + x0], t11 = img[size_t(y1) * w + x1];
	for (uint32_t ch = 0; ch < 4; ch++)
	{
		auto c = [ch](uint32_t t) { return (t >> (8u * ch)) & 0xffu; };
		const uint32_t top = c(t00) * (256u - ax) + c(t10) * ax;
		const uint32_t bot = c(t01) * (256u - ax) + c(t11) * ax;
		const uint32_t acc = top * (256u - ay) + bot * ay;
		out[ch] = __fdiv_rn(float(acc), 16711680.0f);
	}
}

__device__ __forceinline__ bool in_viewport(int32_t px, int32_t py, int32_t vx, int32_t vy, uint32_t vw, uint32_t vh)
{
	return px >= vx && py >= vy && int64_t(px) < int64_t(vx) + int64_t(vw) && int64_t(py) < int64_t(vy) + int64_t(vh);
}

__global__ void __launch_bounds__(256) k_scanout_merge(pgs_scanout_merge_args a)
{
	const uint32_t x = blockIdx.x * kTileW + (threadIdx.x & (kTileW - 1u));
	const uint32_t y = blockIdx.y * kTileH + (threadIdx.x / kTileW);
	if (x >= a.width || y >= a.height)
		return;
	const int32_t px = int32_t(x), py = int32_t(y);
	// Render-pass clear to BGCOLOR with alpha 0 (gs/gs_renderer.cpp:4798-4800).
	uint32_t dst = a.bgcolor & 0x00ffffffu;
	float f[4];
	if ((a.draw & 2u) && in_viewport(px, py, a.vp_x[1], a.vp_y[1], a.vp_width[1], a.vp_height[1]))
	{
		bilinear(a.circuit[1], a.circuit_width[1], a.circuit_height[1], px, py, a.vp_x[1], a.vp_y[1], a.vp_width[1],
		         a.vp_height[1], f);
		dst = pack4(f);
	}
	if ((a.draw & 1u) && in_viewport(px, py, a.vp_x[0], a.vp_y[0], a.vp_width[0], a.vp_height[0]))
	{
		bilinear(a.circuit[0], a.circuit_width[0], a.circuit_height[0], px, py, a.vp_x[0], a.vp_y[0], a.vp_width[0],
		         a.vp_height[0], f);
		if (a.blend != 0u)
		{
			// Colour: src * k + dst * (1 - k), k = ALP / 255 or the source alpha; alpha: src.
			const float k = a.blend == 1u ? __fmul_rn(float(a.alp), 1.0f / 255.0f) : f[3];
			const float kd = __fsub_rn(1.0f, k);
			for (uint32_t ch = 0; ch < 3; ch++)
				f[ch] = __fadd_rn(__fmul_rn(f[ch], k), __fmul_rn(unorm(dst, ch), kd));
		}
		dst = pack4(f);
	}
	a.merged[size_t(y) * a.width + x] = dst;
}

__global__ void __launch_bounds__(256) k_scanout_extwrite(pgs_scanout_extwrite_args a)
{
	const uint32_t x = blockIdx.x * kTileW + (threadIdx.x & (kTileW - 1u));
	const uint32_t y = blockIdx.y * kTileH + (threadIdx.x / kTileW);
	if (x >= a.width || y >= a.height)
		return;
	const uint32_t addr = swizzle_addr(x + a.wdx, y + a.wdy, a.exbp, a.exbw, PSMCT32);
	const float u = __fadd_rn(a.uv_base[0], __fmul_rn(float(x), a.uv_scale[0]));
	const float v = __fadd_rn(a.uv_base[1], __fmul_rn(float(y), a.uv_scale[1]));
	const int32_t tx = min(max(int32_t(floorf(__fmul_rn(u, float(a.view_width)))), 0), int32_t(a.view_width) - 1);
	const int32_t ty = min(max(int32_t(floorf(__fmul_rn(v, float(a.view_height)))), 0), int32_t(a.view_height) - 1);
	const uint32_t t = a.view[size_t(ty) * a.view_width + tx];
	const float r = unorm(t, 0), g = unorm(t, 1), b = unorm(t, 2), al = unorm(t, 3);

	// extwrite.comp:40-49.
	const float ey = __fadd_rn(__fadd_rn(__fmul_rn(0.587f, g), __fmul_rn(0.114f, b)), __fmul_rn(0.299f, r));
	const float ecb = __fsub_rn(__fadd_rn(__fmul_rn(-0.311f, g), __fmul_rn(0.500f, b)), __fmul_rn(0.169f, r));
	const float ecr = __fadd_rn(__fsub_rn(__fmul_rn(-0.419f, g), __fmul_rn(0.081f, b)), __fmul_rn(0.500f, r));
	const float Y = __fdiv_rn(__fadd_rn(__fmul_rn(219.0f, ey), 16.0f), 255.0f);
	const float Cb = __fdiv_rn(__fadd_rn(__fmul_rn(224.0f, ecb), 128.0f), 255.0f);
	const float Cr = __fdiv_rn(__fadd_rn(__fmul_rn(224.0f, ecr), 128.0f), 255.0f);

	float c[4];
	switch (a.emodc)
	{
	case 0: c[0] = r; c[1] = g; c[2] = b; break;
	case 1: c[0] = Y; c[1] = Y; c[2] = Y; break;
	case 2: c[0] = Y; c[1] = Cb; c[2] = Cr; break;
	default: c[0] = al; c[1] = al; c[2] = al; break;
	}
	switch (a.emoda)
	{
	case 0: c[3] = al; break;
	case 1: c[3] = Y; break;
	case 2: c[3] = __fmul_rn(0.5f, Y); break;
	default: c[3] = 0.0f; break;
	}
	*reinterpret_cast<uint32_t *>(a.vram + addr) = pack4(c);
	if (a.clear_reference)
		*reinterpret_cast<uint32_t *>(a.vram + kVramSize + addr) = 0u;
}

// dot(a - b, (0.299, 0.587, 0.114)), summed left to right.
__device__ __forceinline__ float luma_diff(uint32_t p, uint32_t q)
{
	const float dr = __fsub_rn(unorm(p, 0), unorm(q, 0));
	const float dg = __fsub_rn(unorm(p, 1), unorm(q, 1));
	const float db = __fsub_rn(unorm(p, 2), unorm(q, 2));
	return fabsf(__fadd_rn(__fadd_rn(__fmul_rn(dr, 0.299f), __fmul_rn(dg, 0.587f)), __fmul_rn(db, 0.114f)));
}

__global__ void __launch_bounds__(256) k_scanout_weave(pgs_scanout_weave_args a)
{
	const uint32_t x = blockIdx.x * kTileW + (threadIdx.x & (kTileW - 1u));
	const uint32_t y = blockIdx.y * kTileH + (threadIdx.x / kTileW);
	if (x >= a.width || y >= a.height)
		return;
	const size_t w = a.width;
	uint32_t out;
	if ((y & 1u) == a.phase)
		out = a.field[0][size_t(y >> 1) * w + x];
	else if (y == 0u || y == a.height - 1u)
		out = a.field[1][size_t(y >> 1) * w + x];
	else
	{
		const uint32_t up = (y - 1u) >> 1;
		const uint32_t c0 = a.field[0][size_t(up) * w + x], c1 = a.field[0][size_t(up + 1u) * w + x];
		const uint32_t p0 = a.field[2][size_t(up) * w + x], p1 = a.field[2][size_t(up + 1u) * w + x];
		const uint32_t f1 = a.field[1][size_t(y >> 1) * w + x], f3 = a.field[3][size_t(y >> 1) * w + x];
		const float mh = luma_diff(c0, p0), ml = luma_diff(c1, p1), mc = luma_diff(f1, f3);
		const float diff = fmaxf(fminf(mh, ml), mc);
		// smoothstep(0.04, 0.06, diff) = t * t * (3 - 2 t).
		const float t = fminf(fmaxf(__fdiv_rn(__fsub_rn(diff, 0.04f), __fsub_rn(0.06f, 0.04f)), 0.0f), 1.0f);
		const float k = __fmul_rn(__fmul_rn(t, t), __fsub_rn(3.0f, __fmul_rn(2.0f, t)));
		const float kd = __fsub_rn(1.0f, k);
		float c[4];
		for (uint32_t ch = 0; ch < 3; ch++)
		{
			const float bob = __fmul_rn(0.5f, __fadd_rn(unorm(c0, ch), unorm(c1, ch)));
			c[ch] = __fadd_rn(__fmul_rn(unorm(f1, ch), kd), __fmul_rn(bob, k));
		}
		c[3] = 1.0f;
		out = pack4(c);
	}
	a.out[size_t(y) * w + x] = out;
}

dim3 grid_for(uint32_t w, uint32_t h)
{
	return dim3((w + kTileW - 1u) / kTileW, (h + kTileH - 1u) / kTileH, 1u);
}
} // namespace
} // namespace pgs

using namespace pgs;

extern "C" int pgs_k_scanout_circuit(const pgs_scanout_circuit_args *args, hipStream_t stream)
{
	if (!args->width || !args->height)
		return 0;
	hipLaunchKernelGGL(k_scanout_circuit, grid_for(args->width, args->height), dim3(256), 0, stream, *args);
	return -int(hipGetLastError());
}

extern "C" int pgs_k_scanout_merge(const pgs_scanout_merge_args *args, hipStream_t stream)
{
	if (!args->width || !args->height)
		return 0;
	hipLaunchKernelGGL(k_scanout_merge, grid_for(args->width, args->height), dim3(256), 0, stream, *args);
	return -int(hipGetLastError());
}

extern "C" int pgs_k_scanout_extwrite(const pgs_scanout_extwrite_args *args, hipStream_t stream)
{
	if (!args->width || !args->height || !args->view_width || !args->view_height)
		return 0;
	hipLaunchKernelGGL(k_scanout_extwrite, grid_for(args->width, args->height), dim3(256), 0, stream, *args);
	return -int(hipGetLastError());
}

extern "C" int pgs_k_scanout_weave(const pgs_scanout_weave_args *args, hipStream_t stream)
{
	if (!args->width || !args->height)
		return 0;
	hipLaunchKernelGGL(k_scanout_weave, grid_for(args->width, args->height), dim3(256), 0, stream, *args);
	return -int(hipGetLastError());
}
