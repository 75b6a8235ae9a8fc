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
#include "texture_layout.hpp"
#include "pgs_kernels.h"
#include "multi_launch.hpp"
#include "multi_dispatch.hpp"

namespace pgs
{
namespace
{
__device__ __forceinline__ uint32_t rgba16_expand(uint32_t c16, uint32_t aem, uint32_t ta0, uint32_t ta1)
{
	const uint32_t r = (c16 & 0x1fu) << 3, g = ((c16 >> 5) & 0x1fu) << 3, b = ((c16 >> 10) & 0x1fu) << 3;
	const bool zero_alpha = aem && c16 == 0;
	const uint32_t a = zero_alpha ? 0u : (((c16 >> 15) & 1u) ? ta1 : ta0);
	return r | (g << 8) | (b << 16) | (min(a, 255u) << 24);
}

__device__ __forceinline__ uint32_t rgb24_expand(uint32_t c, uint32_t aem, uint32_t ta0)
{
	const bool zero_alpha = aem && (c & 0xffffffu) == 0;
	return (c & 0xffffffu) | ((zero_alpha ? 0u : ta0) << 24);
}

// clut_write.comp:46-155. CSM1 reads the palette from a 16x16 (8-bit indices) or 8x2 (4-bit)
// rectangle whose entries are arranged in 8x2 blocks; CSM2 reads a row of the buffer at
// (COU * 16, COV) with stride CBW.
__device__ __forceinline__ void clut_write_body(const pgs_clut_args &a, uint32_t, uint32_t)
{
	const uint32_t lane = threadIdx.x;
	const bool ct32 = a.cpsm == PSMCT32;
	uint32_t write_coord = lane + a.csa * 16u;
	uint32_t read_coord = lane;
	bool active;
	if (a.index4)
		active = lane < 16u;
	else if (a.csm == 0)
	{
		const uint32_t banks = ct32 ? 16u : 32u;
		active = lane < (banks - a.csa) * 16u;
		read_coord += a.csa * 16u;
	}
	else
		active = true;
	if (!active)
		return;
	uint32_t x, y;
	if (a.csm)
	{
		x = read_coord + a.cou * 16u;
		y = a.cov;
	}
	else
	{
		x = (read_coord & 7u) | (((read_coord >> 4) & 1u) << 3);
		y = ((read_coord >> 3) & 1u) | (((read_coord >> 5) & 7u) << 1);
	}
	const uint32_t cbw = a.csm ? a.cbw : 0u;
	if (ct32)
	{
		const uint32_t addr = swizzle_addr(x, y, a.cbp, cbw, PSMCT32);
		const uint32_t d = *reinterpret_cast<const uint32_t *>(a.vram + addr);
		if (write_coord < 512u)
			a.clut[write_coord] = uint16_t(d);
		if (write_coord + 256u < 512u)
			a.clut[write_coord + 256u] = uint16_t(d >> 16);
	}
	else
	{
		const uint32_t addr = swizzle_addr(x, y, a.cbp, cbw, a.cpsm);
		if (write_coord < 512u)
			a.clut[write_coord] = *reinterpret_cast<const uint16_t *>(a.vram + addr);
	}
}

__device__ __forceinline__ uint32_t clut_color(const uint16_t *clut, const TexDecode &d, uint32_t index)
{
	uint32_t bi = d.csa * 16u + index;
	if (d.psm == PSMT8 || d.psm == PSMT8H)
	{
		// Indices past the last bank clamp to it (upload.comp:76-86).
		const uint32_t bank = min(bi >> 4, d.clut16 ? 31u : 15u);
		bi = bank * 16u + (bi & 15u);
	}
	if (d.clut16)
		return rgba16_expand(clut[bi & 511u], d.aem, d.ta0, d.ta1);
	return uint32_t(clut[bi & 511u]) | (uint32_t(clut[(bi + 256u) & 511u]) << 16);
}

// upload.comp:108-228 (single-sampled source).
__device__ __forceinline__ void texture_decode_body(const pgs_tex_decode_args &a, uint32_t wg, uint32_t)
{
	const TexDecode *decodes = static_cast<const TexDecode *>(a.decodes);
	uint32_t s = 0;
	while (s + 1u < a.num_textures && decodes[s + 1u].group_begin <= wg)
		s++;
	const TexDecode d = decodes[s];
	const uint32_t t = (wg - d.group_begin) * 256u + threadIdx.x;
	if (t >= d.width * d.height)
		return;
	const uint32_t x = t % d.width, y = t / d.width;
	const uint32_t tx = ((x & d.umsk) + d.bx) | d.ufix, ty = ((y & d.vmsk) + d.by) | d.vfix;
	const uint32_t addr = swizzle_addr(tx, ty, d.tbp, d.tbw, d.psm);
	const uint8_t *v = a.vram;
	uint32_t c;
	switch (d.psm)
	{
	case PSMCT24: case PSMZ24:
		c = rgb24_expand(*reinterpret_cast<const uint32_t *>(v + addr), d.aem, d.ta0);
		break;
	case PSMCT16: case PSMCT16S: case PSMZ16: case PSMZ16S:
		c = rgba16_expand(*reinterpret_cast<const uint16_t *>(v + addr), d.aem, d.ta0, d.ta1);
		break;
	case PSMT8: c = clut_color(a.clut, d, v[addr]); break;
	case PSMT8H: c = clut_color(a.clut, d, v[addr + 3u]); break;
	case PSMT4: c = clut_color(a.clut, d, (v[addr >> 1] >> ((addr & 1u) * 4u)) & 0xfu); break;
	case PSMT4HL: c = clut_color(a.clut, d, (*reinterpret_cast<const uint32_t *>(v + addr) >> 24) & 0xfu); break;
	case PSMT4HH: c = clut_color(a.clut, d, *reinterpret_cast<const uint32_t *>(v + addr) >> 28); break;
	default: c = *reinterpret_cast<const uint32_t *>(v + addr); break;
	}
	a.images[d.image_offset + t] = c;
}

__global__ void __launch_bounds__(256) k_clut_write(pgs_clut_args a) { clut_write_body(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) k_texture_decode(pgs_tex_decode_args a) { texture_decode_body(a, blockIdx.x, gridDim.x); }
PGS_DEFINE_MULTI(256, k_clut_write_multi, clut_write_body, pgs_clut_args)
PGS_DEFINE_MULTI(256, k_texture_decode_multi, texture_decode_body, pgs_tex_decode_args)
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
