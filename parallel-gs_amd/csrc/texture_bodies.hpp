// Device bodies of the CLUT load and texture decode kernels, shared by their own launches
// (kernels_texture.hip) and the fused launch (kernels_raster.hip). Replaces clut_write.comp
// and upload.comp; the launch shapes are described at the top of kernels_texture.hip.
#pragma once
#include "gs_defs.hpp"
#include "gs_raster.hpp"
#include "texture_layout.hpp"
#include "pgs_kernels.h"

namespace pgs
{
namespace texb
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

// upload.comp:108-228 (single-sampled source). A mipmapped texture's levels follow each other
// in the image, level l being max(width >> l, 1) x max(height >> l, 1) texels read from its own
// TBP / TBW with the rectangle origin, MSK and FIX shifted right by l
// (gs/gs_renderer.cpp:3543-3593).
__device__ __forceinline__ void texture_decode_body(const pgs_tex_decode_args &a, uint32_t wg, uint32_t)
{
	const TexDecode *decodes = static_cast<const TexDecode *>(a.decodes);
	uint32_t s = 0;
	while (s + 1u < a.num_textures && decodes[s + 1u].group_begin <= wg)
		s++;
	const TexDecode d = decodes[s];
	const uint32_t t0 = (wg - d.group_begin) * 256u + threadIdx.x;
	if (t0 >= d.texels)
		return;
	uint32_t t = t0, level = 0, w = d.width, h = d.height, tbp = d.tbp, tbw = d.tbw;
	// A super-sampled texture's layers follow each other: layer 0 from VRAM, layer k from the
	// plane of sample k - 1 (upload.comp:133-228 with SUPER_SAMPLED).
	uint32_t layer = 0;
	if (d.samples)
	{
		layer = t / (w * h);
		t -= layer * (w * h);
	}
	while (t >= w * h && level + 1u < d.levels)
	{
		t -= w * h;
		level++;
		w = max(d.width >> level, 1u);
		h = max(d.height >> level, 1u);
		// Read from the record in memory: a variable index into the local copy would place it
		// in scratch.
		const uint32_t buf = decodes[s].level_buf[level - 1u];
		tbp = buf & 0x3fffu;
		tbw = buf >> 14;
	}
	if (t >= w * h)
		return;
	const uint32_t x = t % w, y = t / w;
	if (a.feedback && d.levels == 1u && !d.samples && w <= 1024u && h <= 1024u)
	{
		const uint32_t word = a.feedback[s * PGS_FEEDBACK_WORDS + (y >> 3) * 4u + (x >> 8)];
		if (!((word >> ((x >> 3) & 31u)) & 1u))
			return; // no primitive of the pass can sample this 8x8 block
	}
	const uint32_t tx = ((x & (d.umsk >> level)) + (d.bx >> level)) | (d.ufix >> level);
	const uint32_t ty = ((y & (d.vmsk >> level)) + (d.by >> level)) | (d.vfix >> level);
	const uint32_t addr = swizzle_addr(tx, ty, tbp, tbw, d.psm);
	const uint8_t *v = a.vram;
	// Sample plane and reference plane of a sample layer: the sample's value replaces the
	// single-sampled one where the reference plane marks all its bits (24-bit formats: the
	// marked bits).
	const uint8_t *ref = v + kVramSize, *hi = v + size_t(layer + 1u) * kVramSize;
	uint32_t c;
	switch (d.psm)
	{
	case PSMCT24: case PSMZ24:
	{
		uint32_t w32 = *reinterpret_cast<const uint32_t *>(v + addr);
		if (layer)
		{
			const uint32_t m = *reinterpret_cast<const uint32_t *>(ref + addr);
			if (m)
				w32 = (*reinterpret_cast<const uint32_t *>(hi + addr) & m) | (w32 & ~m);
		}
		c = rgb24_expand(w32, d.aem, d.ta0);
		break;
	}
	case PSMCT16: case PSMCT16S: case PSMZ16: case PSMZ16S:
	{
		uint32_t w16 = *reinterpret_cast<const uint16_t *>(v + addr);
		if (layer && *reinterpret_cast<const uint16_t *>(ref + addr) == 0xffffu)
			w16 = *reinterpret_cast<const uint16_t *>(hi + addr);
		c = rgba16_expand(w16, d.aem, d.ta0, d.ta1);
		break;
	}
	case PSMT8:
		c = clut_color(a.clut, d, (layer && ref[addr] == 0xffu) ? hi[addr] : v[addr]);
		break;
	case PSMT8H:
		c = clut_color(a.clut, d, (layer && ref[addr + 3u] == 0xffu) ? hi[addr + 3u] : v[addr + 3u]);
		break;
	case PSMT4: c = clut_color(a.clut, d, (v[addr >> 1] >> ((addr & 1u) * 4u)) & 0xfu); break;
	case PSMT4HL: c = clut_color(a.clut, d, (*reinterpret_cast<const uint32_t *>(v + addr) >> 24) & 0xfu); break;
	case PSMT4HH: c = clut_color(a.clut, d, *reinterpret_cast<const uint32_t *>(v + addr) >> 28); break;
	default:
		c = *reinterpret_cast<const uint32_t *>(v + addr);
		if (layer && *reinterpret_cast<const uint32_t *>(ref + addr) == ~0u)
			c = *reinterpret_cast<const uint32_t *>(hi + addr);
		break;
	}
	a.images[d.image_offset + t0] = c;
}

}
}
