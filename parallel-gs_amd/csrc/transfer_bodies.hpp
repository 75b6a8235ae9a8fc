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

// Per-pixel read rules of vram_copy.comp:199-256: a local->local source pixel by the source
// PSM (whole dword, halfword, byte, the top byte, a nibble, or the top nibbles of a dword), and
// a host payload pixel packed at the destination's bits per pixel. Every read except the
// 24-bit host one is one aligned dword load and a bit-field extract: the byte offset, shift and
// width come from the element's address.
struct PixelRead
{
	uint32_t width;     // field width in bits
	uint32_t fixed;     // shift of a field at a fixed place in a dword (top byte / top nibbles)
	uint32_t log2_bits; // log2 of the address unit in bits (8: byte addresses, 4: nibble
	                    // addresses, 16 / 32: halfword / dword indices), 0 for fixed fields
};

__device__ __forceinline__ PixelRead local_read(uint32_t spsm)
{
	switch (spsm)
	{
	case PSMCT16: case PSMCT16S: case PSMZ16: case PSMZ16S: return { 16u, 0u, 3u };
	case PSMT8: return { 8u, 0u, 3u };
	case PSMT8H: return { 8u, 24u, 0u };
	case PSMT4: return { 4u, 0u, 2u };
	case PSMT4HL: return { 4u, 24u, 0u };
	case PSMT4HH: return { 4u, 28u, 0u };
	default: return { 32u, 0u, 0u };
	}
}

// The field at address `e` of `base` (in units of 2^log2_bits bits; byte addresses of whole
// dwords for fixed fields): the aligned dword holding it, shifted and masked.
__device__ __forceinline__ uint32_t element_load(const uint8_t *base, uint32_t e, const PixelRead &r)
{
	const uint32_t bit = r.log2_bits ? e << r.log2_bits : (e << 3);
	const uint32_t d = *reinterpret_cast<const uint32_t *>(base + ((bit >> 3) & ~3u));
	const uint32_t shift = r.log2_bits ? (bit & 31u) : r.fixed;
	return r.width == 32u ? d : (d >> shift) & ((1u << r.width) - 1u);
}

// Applies one record (one copy's pixels in this block) to the lane's dword v: every pixel's
// source address and validity are computed first and every load is issued before any result
// is used, so a dword's NPIX pixels cost one memory round trip; pixels that are not written
// load from a valid dummy address. NPIX is the destination's pixels per dword.
template <uint32_t NPIX>
__device__ __forceinline__ void apply_record(const pgs_vram_copy_args &a, const CopyDesc &c, uint32_t lane,
                                             uint32_t x0, uint32_t y0, uint32_t &v, uint32_t &touched)
{
	constexpr uint32_t sc = NPIX == 1 ? SC_32 : NPIX == 2 ? SC_16 : NPIX == 4 ? SC_8 : SC_4;
	const uint32_t dpsm = c.dpsm;
	const uint32_t w = c.w, h = c.h;
	const uint32_t dsax = c.dsax, dsay = c.dsay;
	// Pixels beyond the destination stride are dropped except on the last page row of the
	// transfer (vram_copy.comp:282-300).
	uint32_t eff_stride = c.dbw;
	if (dpsm == PSMT4 || dpsm == PSMT8)
		eff_stride &= ~1u;
	const uint32_t ph = psm_geometry(dpsm).ph;
	const uint32_t last_page_row = (dsay + h - 1u) >> ph;
	const bool local = (c.flags & COPY_LOCAL) != 0;

	// Element index of each pixel's source: a VRAM byte / nibble address for local copies, the
	// raster index into the payload for host copies (payload addresses of 4-bit pixels are
	// nibbles, of 8/16/32-bit ones whole elements).
	uint32_t src[NPIX];
	uint32_t ok = 0;
#pragma unroll
	for (uint32_t k = 0; k < NPIX; k++)
	{
		uint32_t dx, dy;
		block_lane_pixel(sc, lane, k, dx, dy);
		const uint32_t rx = (x0 + dx - dsax) & 2047u;
		const uint32_t ry = (y0 + dy - dsay) & 2047u;
		bool valid = rx < w && ry < h && !(rx >= eff_stride * 64u && ((dsay + ry) >> ph) != last_page_row);
		if (local)
			src[k] = swizzle_addr((rx + c.ssax) & 2047u, (ry + c.ssay) & 2047u, c.sbp, c.sbw, c.spsm);
		else
		{
			src[k] = ry * w + rx;
			valid = valid && src[k] >= c.first_pixel && src[k] < c.end_pixel;
		}
		ok |= valid ? 1u << k : 0u;
	}

	uint32_t data[NPIX];
	if (local)
	{
		// VRAM addresses are masked to VRAM (nibble addresses for 4-bit sources), so every load is
		// in bounds; pixels that are not written read element 0.
		const uint8_t *plane = (c.flags & COPY_SOURCE_SHADOW) ? a.shadow : a.vram;
		const PixelRead r = local_read(c.spsm);
#pragma unroll
		for (uint32_t k = 0; k < NPIX; k++)
			data[k] = element_load(plane, ((ok >> k) & 1u) ? src[k] : 0u, r);
	}
	else
	{
		// Host payload packed at the destination's bits per pixel; pixels past the available
		// payload are not written. Loads of pixels that are not written read element 0.
		const uint8_t *p = a.payload + c.payload_offset;
		const uint32_t bytes = c.payload_bytes;
		const uint32_t bpp = psm_bpp(dpsm);
		const uint32_t avail = bpp == 24u ? bytes / 3u : (bytes * 8u) >> (31u - __builtin_clz(bpp));
#pragma unroll
		for (uint32_t k = 0; k < NPIX; k++)
			ok &= src[k] < avail ? ~0u : ~(1u << k);
		if (bpp == 24u)
		{
#pragma unroll
			for (uint32_t k = 0; k < NPIX; k++)
			{
				const uint8_t *q = p + (((ok >> k) & 1u) ? src[k] * 3u : 0u);
				data[k] = uint32_t(q[0]) | (uint32_t(q[1]) << 8) | (uint32_t(q[2]) << 16);
			}
		}
		else
		{
			const PixelRead r = { bpp, 0u, uint32_t(31 - __builtin_clz(bpp)) };
#pragma unroll
			for (uint32_t k = 0; k < NPIX; k++)
				data[k] = element_load(p, ((ok >> k) & 1u) ? src[k] : 0u, r);
		}
	}

#pragma unroll
	for (uint32_t k = 0; k < NPIX; k++)
		if ((ok >> k) & 1u)
		{
			const uint32_t mask = psm_lane_mask(dpsm, k);
			v = (v & ~mask) | ((data[k] << psm_lane_shift(dpsm, k)) & mask);
			touched |= mask;
		}
}

typedef uint32_t xfer_u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t xfer_u32x8 __attribute__((ext_vector_type(8)));
typedef uint32_t xfer_u32x16 __attribute__((ext_vector_type(16)));

// A wave-uniform pointer in scalar registers (the operand of a scalar load).
template <typename T>
__device__ __forceinline__ const T *uniform_ptr(const T *p)
{
	const uint64_t v = reinterpret_cast<uint64_t>(p);
	return reinterpret_cast<const T *>(uint64_t(uniform(uint32_t(v >> 32))) << 32 | uniform(uint32_t(v)));
}

// A work item (32 bytes) and a copy descriptor (72 bytes) through the scalar cache, each load
// with its wait in one asm statement: the metadata is never written by a launch, and left to
// itself the compiler reads fields past the first store with vector loads, one round trip each.
__device__ __forceinline__ CopyItem load_item(const CopyItem *p)
{
	xfer_u32x8 v;
	asm volatile("s_load_dwordx8 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=&s"(v) : "s"(uniform_ptr(p)));
	CopyItem it;
	__builtin_memcpy(&it, &v, sizeof(it));
	return it;
}

__device__ __forceinline__ CopyDesc load_desc(const CopyDesc *p)
{
	xfer_u32x16 lo;
	xfer_u32x2 hi;
	asm volatile("s_load_dwordx16 %0, %2, 0x0\n\t"
	             "s_load_dwordx2 %1, %2, 0x40\n\t"
	             "s_waitcnt lgkmcnt(0)"
	             : "=&s"(lo), "=&s"(hi)
	             : "s"(uniform_ptr(p)));
	CopyDesc c;
	__builtin_memcpy(&c, &lo, 64);
	__builtin_memcpy(reinterpret_cast<uint8_t *>(&c) + 64, &hi, 8);
	return c;
}

__device__ __forceinline__ BlockRecord load_record(const BlockRecord *p)
{
	xfer_u32x2 v;
	asm volatile("s_load_dwordx2 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=&s"(v) : "s"(uniform_ptr(p)));
	BlockRecord r;
	__builtin_memcpy(&r, &v, sizeof(r));
	return r;
}
static_assert(sizeof(CopyDesc) == 72 && sizeof(BlockRecord) == 8, "scalar load sizes");

__device__ __forceinline__ void apply_any(const pgs_vram_copy_args &a, const CopyDesc &c, uint32_t lane, uint32_t xy,
                                          uint32_t &v, uint32_t &touched)
{
	const uint32_t x0 = xy & 0xffffu, y0 = xy >> 16;
	switch (psm_storage_class(c.dpsm))
	{
	case SC_16: apply_record<2>(a, c, lane, x0, y0, v, touched); break;
	case SC_8: apply_record<4>(a, c, lane, x0, y0, v, touched); break;
	case SC_4: apply_record<8>(a, c, lane, x0, y0, v, touched); break;
	default: apply_record<1>(a, c, lane, x0, y0, v, touched); break;
	}
}

// One wave per item; lane L owns dword L of the item's block. A destination block applies its
// records in submission order, the first one inline in the item; a block the next level
// snapshots stores its result there too, and a snapshot item copies a block the level does not
// write into the next snapshot. The metadata is read through the scalar cache (it is never
// written by a launch).
__device__ __forceinline__ void resolve_blocks_body(const pgs_vram_copy_args &a, uint32_t wg, uint32_t)
{
	const uint32_t wave = uniform((wg * blockDim.x + threadIdx.x) >> 6);
	const uint32_t lane = threadIdx.x & 63u;
	if (wave >= a.num_items)
		return;
	const CopyItem *items = reinterpret_cast<const CopyItem *>(a.items);
	const BlockRecord *recs = reinterpret_cast<const BlockRecord *>(a.records);
	const CopyDesc *descs = reinterpret_cast<const CopyDesc *>(a.descs);
	uint32_t *vram = reinterpret_cast<uint32_t *>(a.vram);

	const CopyItem it = load_item(items + wave);
	const uint32_t word = it.block * 64u + lane;
	uint32_t v = vram[word];
	if (it.flags & COPY_ITEM_SNAPSHOT)
	{
		reinterpret_cast<uint32_t *>(a.shadow_next)[word] = v;
		return;
	}
	uint32_t touched = 0;
	apply_any(a, load_desc(descs + it.copy), lane, it.xy, v, touched);
	for (uint32_t r = 0; r < it.rec_count; r++)
	{
		const BlockRecord rec = load_record(recs + it.rec_begin + r);
		apply_any(a, load_desc(descs + rec.copy), lane, uint32_t(rec.x0) | (uint32_t(rec.y0) << 16), v, touched);
	}
	vram[word] = v;
	if (it.flags & COPY_ITEM_TO_NEXT_SNAPSHOT)
		reinterpret_cast<uint32_t *>(a.shadow_next)[word] = v;
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
