// Device bodies of the VRAM transfer kernels (block resolve and snapshot), shared by their own
// launches (kernels_transfer.hip) and the fused launch (kernels_raster.hip). Replaces
// gs/shaders/vram_copy.comp; the design notes are at the top of kernels_transfer.hip.
#pragma once
#include "gs_defs.hpp"
#include "pgs_kernels.h"

namespace pgs
{
namespace xfer
{
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Reads one pixel of a local->local source (vram_copy.comp:199-231).
__device__ __forceinline__ uint32_t read_source_local(const uint8_t *plane, uint32_t spsm, uint32_t addr)
{
	switch (spsm)
	{
	case PSMCT16: case PSMCT16S: case PSMZ16: case PSMZ16S:
		return *reinterpret_cast<const uint16_t *>(plane + addr);
	case PSMT8:
		return plane[addr];
	case PSMT8H:
		return plane[addr | 3u];
	case PSMT4:
		return (plane[addr >> 1] >> ((addr & 1u) * 4u)) & 0xfu;
	case PSMT4HL:
		return (*reinterpret_cast<const uint32_t *>(plane + addr) >> 24) & 0xfu;
	case PSMT4HH:
		return *reinterpret_cast<const uint32_t *>(plane + addr) >> 28;
	default:
		return *reinterpret_cast<const uint32_t *>(plane + addr);
	}
}

// Reads one host payload pixel packed at the destination's bits per pixel
// (vram_copy.comp:232-256). Returns false past the end of the available payload.
__device__ __forceinline__ bool read_source_host(const uint8_t *p, uint32_t bytes, uint32_t bpp, uint32_t idx,
                                                 uint32_t &out)
{
	switch (bpp)
	{
	case 24:
		if (idx >= bytes / 3u) return false;
		out = uint32_t(p[idx * 3u]) | (uint32_t(p[idx * 3u + 1u]) << 8) | (uint32_t(p[idx * 3u + 2u]) << 16);
		return true;
	case 16:
		if (idx >= bytes / 2u) return false;
		out = *reinterpret_cast<const uint16_t *>(p + idx * 2u);
		return true;
	case 8:
		if (idx >= bytes) return false;
		out = p[idx];
		return true;
	case 4:
		if (idx >= bytes * 2u) return false;
		out = (p[idx >> 1] >> ((idx & 1u) * 4u)) & 0xfu;
		return true;
	default:
		if (idx >= bytes / 4u) return false;
		out = *reinterpret_cast<const uint32_t *>(p + idx * 4u);
		return true;
	}
}

__device__ __forceinline__ void resolve_blocks_body(const pgs_vram_copy_args &a, uint32_t wg, uint32_t)
{
	const uint32_t wave = (wg * blockDim.x + threadIdx.x) >> 6;
	const uint32_t lane = threadIdx.x & 63u;
	if (wave >= a.num_blocks)
		return;

	const DestBlock *blocks = reinterpret_cast<const DestBlock *>(a.blocks);
	const BlockRecord *recs = reinterpret_cast<const BlockRecord *>(a.records);
	const CopyDesc *descs = reinterpret_cast<const CopyDesc *>(a.descs);
	uint32_t *vram = reinterpret_cast<uint32_t *>(a.vram);

	const uint32_t block = uniform(blocks[wave].block);
	const uint32_t rec_begin = uniform(blocks[wave].rec_begin);
	const uint32_t rec_count = uniform(blocks[wave].rec_count);
	const uint32_t word = block * 64u + lane;

	uint32_t v = vram[word];
	uint32_t touched = 0;

	for (uint32_t r = 0; r < rec_count; r++)
	{
		const uint32_t copy = uniform(recs[rec_begin + r].copy);
		const uint32_t x0 = uniform(recs[rec_begin + r].x0);
		const uint32_t y0 = uniform(recs[rec_begin + r].y0);
		const CopyDesc &c = descs[copy];
		const uint32_t dpsm = uniform(c.dpsm);
		const uint32_t flags = uniform(c.flags);
		const uint32_t w = uniform(c.w), h = uniform(c.h);
		const uint32_t dsax = uniform(c.dsax), dsay = uniform(c.dsay);
		const uint32_t sc = psm_storage_class(dpsm);
		const uint32_t npix = psm_pixels_per_dword(dpsm);

		// Pixels beyond the destination stride are dropped except on the last page row of
		// the transfer (vram_copy.comp:282-300).
		uint32_t eff_stride = uniform(c.dbw);
		if (dpsm == PSMT4 || dpsm == PSMT8)
			eff_stride &= ~1u;
		const uint32_t ph = psm_geometry(dpsm).ph;
		const uint32_t last_page_row = (dsay + h - 1u) >> ph;

		for (uint32_t k = 0; k < npix; k++)
		{
			uint32_t dx, dy;
			block_lane_pixel(sc, lane, k, dx, dy);
			const uint32_t rx = (x0 + dx - dsax) & 2047u;
			const uint32_t ry = (y0 + dy - dsay) & 2047u;
			if (rx >= w || ry >= h)
				continue;
			if (rx >= eff_stride * 64u && ((dsay + ry) >> ph) != last_page_row)
				continue;

			uint32_t data;
			if (flags & COPY_LOCAL)
			{
				const uint32_t sx = (rx + c.ssax) & 2047u;
				const uint32_t sy = (ry + c.ssay) & 2047u;
				const uint32_t spsm = uniform(c.spsm);
				const uint32_t addr = swizzle_addr(sx, sy, uniform(c.sbp), uniform(c.sbw), spsm);
				const uint8_t *plane = (flags & COPY_SOURCE_SHADOW) ? a.shadow : a.vram;
				data = read_source_local(plane, spsm, addr);
			}
			else
			{
				const uint32_t idx = ry * w + rx;
				if (idx < c.first_pixel || idx >= c.end_pixel)
					continue;
				if (!read_source_host(a.payload + c.payload_offset, c.payload_bytes, psm_bpp(dpsm), idx, data))
					continue;
			}

			const uint32_t mask = psm_lane_mask(dpsm, k);
			v = (v & ~mask) | ((data << psm_lane_shift(dpsm, k)) & mask);
			touched |= mask;
		}
	}

	vram[word] = v;
	// Transfers invalidate the super-sampled copy of what they overwrite
	// (vram_copy.comp INVALIDATE_SUPER_SAMPLE).
	if (a.ref_plane && touched)
		reinterpret_cast<uint32_t *>(a.ref_plane)[word] &= ~touched;
}

// Snapshots whole blocks of VRAM into the shadow plane before a batch whose local copies read
// memory the batch also writes (gs/gs_renderer.cpp:3921-3942 shadow-page sync).
__device__ __forceinline__ void snapshot_blocks_body(const pgs_vram_snapshot_args &a, uint32_t wg, uint32_t)
{
	const uint32_t wave = (wg * blockDim.x + threadIdx.x) >> 6;
	const uint32_t lane = threadIdx.x & 63u;
	if (wave >= a.num_blocks)
		return;
	const uint32_t block = __builtin_amdgcn_readfirstlane(a.block_list[wave]);
	const uint32_t word = block * 64u + lane;
	reinterpret_cast<uint32_t *>(a.shadow)[word] = reinterpret_cast<const uint32_t *>(a.vram)[word];
}

}
}
