// Shared host/device definitions for the MI355X GS backend.
//
// Register encodings follow the GS register map (gs/gs_registers.hpp:42-532 and
// gs/gs_register_addr.hpp:8-62 in the reference); they are decoded here with explicit
// shifts and masks instead of bitfield structs so the same code compiles for host and
// device. Struct layouts of the primitive records mirror the reference's shared structs
// (gs/shaders/data_structures.h:32-80) because the front end and the kernels exchange them.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define PGS_HD __host__ __device__ inline

namespace pgs
{
// ---------------------------------------------------------------------------------------
// Pixel storage modes (gs/shaders/data_structures.h:311-323).
enum : uint32_t
{
	PSMCT32 = 0x00, PSMCT24 = 0x01, PSMCT16 = 0x02, PSMCT16S = 0x0a,
	PSMT8 = 0x13, PSMT4 = 0x14, PSMT8H = 0x1b, PSMT4HL = 0x24, PSMT4HH = 0x2c,
	PSMZ32 = 0x30, PSMZ24 = 0x31, PSMZ16 = 0x32, PSMZ16S = 0x3a
};

constexpr uint32_t kVramSize = 4u * 1024u * 1024u;
constexpr uint32_t kVramMask = kVramSize - 1u;
constexpr uint32_t kBlockBytes = 256u;
constexpr uint32_t kPageBytes = 8192u;
constexpr uint32_t kNumBlocks = kVramSize / kBlockBytes; // 16384
constexpr uint32_t kNumPages = kVramSize / kPageBytes;   // 512

PGS_HD uint32_t psm_bpp(uint32_t psm)
{
	switch (psm)
	{
	case PSMCT24: case PSMZ24: return 24;
	case PSMCT16: case PSMCT16S: case PSMZ16: case PSMZ16S: return 16;
	case PSMT8: case PSMT8H: return 8;
	case PSMT4: case PSMT4HL: case PSMT4HH: return 4;
	default: return 32;
	}
}

// Samples per pixel along x and y (log2) of a sampling grid (gs/gs_interface.cpp:65-117):
// ordered 2x = 1 x 2, 4x = 2 x 2, 8x = 2 x 4, 16x = 4 x 4; sparse 4x = 1 x 4, 16x = 2 x 8.
PGS_HD void grid_rates(uint32_t sample_log2, bool sparse, uint32_t &rate_x, uint32_t &rate_y)
{
	rate_x = sample_log2 / 2u;
	if (sparse && (sample_log2 == 2u || sample_log2 == 4u))
		rate_x--;
	rate_y = sample_log2 - rate_x;
}

PGS_HD bool psm_is_16(uint32_t psm) { return psm == PSMCT16 || psm == PSMCT16S || psm == PSMZ16 || psm == PSMZ16S; }
PGS_HD bool psm_is_z(uint32_t psm) { return psm == PSMZ32 || psm == PSMZ24 || psm == PSMZ16 || psm == PSMZ16S; }

// Storage class of a PSM inside a 256-byte block: how many pixels share one dword.
enum StorageClass : uint32_t { SC_32 = 0, SC_16 = 1, SC_8 = 2, SC_4 = 3 };
PGS_HD uint32_t psm_storage_class(uint32_t psm)
{
	if (psm_is_16(psm)) return SC_16;
	if (psm == PSMT8) return SC_8;
	if (psm == PSMT4) return SC_4;
	return SC_32;
}

// Page / block geometry (log2 of page width/height, block width/height, column height).
struct Geometry { uint32_t pw, ph, bw, bh, ch; };
PGS_HD Geometry psm_geometry(uint32_t psm)
{
	switch (psm_storage_class(psm))
	{
	case SC_16: return { 6, 6, 4, 3, 1 };
	case SC_8: return { 7, 6, 4, 4, 2 };
	case SC_4: return { 7, 7, 5, 4, 2 };
	default: return { 6, 5, 3, 3, 1 };
	}
}

// Block index of a block coordinate inside a page: a 5-bit interleave of block x/y that
// differs per storage layout (the GS block arrangement tables).
PGS_HD uint32_t block_interleave(uint32_t psm, uint32_t bx, uint32_t by)
{
	if (psm == PSMCT16 || psm == PSMZ16 || psm == PSMT4)
		return (by & 1u) | ((bx & 1u) << 1) | ((by & 2u) << 1) | ((bx & 2u) << 2) | ((by & 4u) << 2);
	if (psm == PSMCT16S || psm == PSMZ16S)
		return (by & 1u) | ((bx & 1u) << 1) | (by & 4u) | ((by & 2u) << 2) | ((bx & 2u) << 3);
	return (bx & 1u) | ((by & 1u) << 1) | ((bx & 2u) << 1) | ((by & 2u) << 2) | ((bx & 4u) << 2);
}

// Byte address (4-bit formats: nibble address) of pixel (x, y) in a buffer at block
// `base` with `stride` in units of 64 pixels. Same mapping as swizzle_PS2
// (gs/shaders/swizzle_utils.h:329-522) except that the result is always a byte (nibble)
// address; callers shift down to the element they access.
PGS_HD uint32_t swizzle_addr(uint32_t x, uint32_t y, uint32_t base, uint32_t stride, uint32_t psm)
{
	const Geometry g = psm_geometry(psm);
	const uint32_t sc = psm_storage_class(psm);
	if (sc == SC_8 || sc == SC_4)
		stride >>= 1;

	uint32_t page = (y >> g.ph) * stride + (x >> g.pw);
	uint32_t bx = (x & ((1u << g.pw) - 1u)) >> g.bw;
	uint32_t by = (y & ((1u << g.ph) - 1u)) >> g.bh;
	uint32_t col = (y & ((1u << g.bh) - 1u)) >> g.ch;
	uint32_t px = x & ((1u << g.bw) - 1u);
	uint32_t py = y & ((1u << g.ch) - 1u);

	uint32_t block = block_interleave(psm, bx, by) + base;
	if (psm_is_z(psm))
		block ^= 0x18u;
	uint32_t colbase = page * kPageBytes + block * kBlockBytes + col * 64u;

	switch (sc)
	{
	case SC_16:
	{
		uint32_t pi = ((px >> 3) & 1u) | ((px & 1u) << 1) | ((py & 1u) << 2) | ((px & 2u) << 2) | ((px & 4u) << 2);
		return (colbase + pi * 2u) & kVramMask;
	}
	case SC_8:
	{
		uint32_t hi = ((px >> 2) ^ (py >> 1) ^ col) & 1u;
		uint32_t pi = ((py >> 1) & 1u) | ((px & 8u) >> 2) | ((px & 1u) << 2) | ((py & 1u) << 3) |
		              ((px & 2u) << 3) | (hi << 5);
		return (colbase + pi) & kVramMask;
	}
	case SC_4:
	{
		uint32_t hi = ((px >> 2) ^ (py >> 1) ^ col) & 1u;
		uint32_t pi = ((py >> 1) & 1u) | ((px & 8u) >> 2) | ((px & 16u) >> 2) | ((px & 1u) << 3) |
		              ((py & 1u) << 4) | ((px & 2u) << 4) | (hi << 6);
		return ((colbase << 1) + pi) & ((kVramMask << 1) | 1u);
	}
	default:
	{
		uint32_t pi = (px & 1u) | ((py & 1u) << 1) | ((px & 2u) << 1) | ((px & 4u) << 1);
		return (colbase + pi * 4u) & kVramMask;
	}
	}
}

// Inverse of the in-block layout: the (dx, dy) pixel offset inside a block for sub-pixel k
// of dword `lane` (0..63). Sub-pixel k is the k-th element of the dword in ascending bit
// order (halfword, byte or nibble index).
PGS_HD void block_lane_pixel(uint32_t sc, uint32_t lane, uint32_t k, uint32_t &dx, uint32_t &dy)
{
	uint32_t col = lane >> 4;
	switch (sc)
	{
	case SC_16:
	{
		uint32_t pi = ((lane & 15u) << 1) | k;
		dx = ((pi >> 1) & 1u) | (((pi >> 3) & 1u) << 1) | (((pi >> 4) & 1u) << 2) | ((pi & 1u) << 3);
		dy = col * 2u + ((pi >> 2) & 1u);
		break;
	}
	case SC_8:
	{
		uint32_t pi = ((lane & 15u) << 2) | k;
		uint32_t py1 = pi & 1u;
		uint32_t px2 = ((pi >> 5) ^ py1 ^ col) & 1u;
		dx = ((pi >> 2) & 1u) | (((pi >> 4) & 1u) << 1) | (px2 << 2) | (((pi >> 1) & 1u) << 3);
		dy = col * 4u + ((py1 << 1) | ((pi >> 3) & 1u));
		break;
	}
	case SC_4:
	{
		uint32_t pi = ((lane & 15u) << 3) | k;
		uint32_t py1 = pi & 1u;
		uint32_t px2 = ((pi >> 6) ^ py1 ^ col) & 1u;
		dx = ((pi >> 3) & 1u) | (((pi >> 5) & 1u) << 1) | (px2 << 2) | (((pi >> 1) & 1u) << 3) |
		     (((pi >> 2) & 1u) << 4);
		dy = col * 4u + ((py1 << 1) | ((pi >> 4) & 1u));
		break;
	}
	default:
	{
		uint32_t pi = lane & 15u;
		dx = (pi & 1u) | (((pi >> 2) & 1u) << 1) | (((pi >> 3) & 1u) << 2);
		dy = col * 2u + ((pi >> 1) & 1u);
		break;
	}
	}
}

// Bits of a dword a pixel of this PSM owns, for sub-pixel k (vram_copy.comp write masks).
PGS_HD uint32_t psm_lane_mask(uint32_t psm, uint32_t k)
{
	switch (psm)
	{
	case PSMCT24: case PSMZ24: return 0x00ffffffu;
	case PSMCT16: case PSMCT16S: case PSMZ16: case PSMZ16S: return 0xffffu << (16u * k);
	case PSMT8: return 0xffu << (8u * k);
	case PSMT8H: return 0xff000000u;
	case PSMT4: return 0xfu << (4u * k);
	case PSMT4HL: return 0x0f000000u;
	case PSMT4HH: return 0xf0000000u;
	default: return 0xffffffffu;
	}
}

// Shift of the pixel value inside the dword for sub-pixel k.
PGS_HD uint32_t psm_lane_shift(uint32_t psm, uint32_t k)
{
	switch (psm)
	{
	case PSMCT16: case PSMCT16S: case PSMZ16: case PSMZ16S: return 16u * k;
	case PSMT8: return 8u * k;
	case PSMT8H: return 24u;
	case PSMT4: return 4u * k;
	case PSMT4HL: return 24u;
	case PSMT4HH: return 28u;
	default: return 0u;
	}
}

PGS_HD uint32_t psm_pixels_per_dword(uint32_t psm)
{
	switch (psm_storage_class(psm))
	{
	case SC_16: return 2;
	case SC_8: return 4;
	case SC_4: return 8;
	default: return 1;
	}
}

// ---------------------------------------------------------------------------------------
// Transfer work description shared by the front end and the copy kernels.

enum : uint32_t
{
	COPY_LOCAL = 1u << 0,          // local->local (source is VRAM), else host->local
	COPY_SOURCE_SHADOW = 1u << 1,  // read the source from the batch-start snapshot
};

struct CopyDesc
{
	uint32_t dpsm, dbp, dbw, dsax, dsay;
	uint32_t spsm, sbp, sbw, ssax, ssay;
	uint32_t w, h;
	uint32_t first_pixel, end_pixel; // host->local: raster pixel range written by this flush
	uint32_t flags;
	uint32_t payload_bytes;          // bytes of host payload available (source_size)
	uint64_t payload_offset;         // byte offset of this transfer's payload in the payload arena
};
static_assert(sizeof(CopyDesc) == 72, "CopyDesc layout");

// One destination block touched by one copy: the block-aligned VRAM-space origin of the
// pixels that land in it.
struct BlockRecord
{
	uint32_t copy;
	uint16_t x0, y0;
};

// One wave item of a transfer level: a destination block with its first record inline and the
// ordered slice of further records that write it, or a block the next level's snapshot copies.
struct CopyItem
{
	uint32_t block;
	uint32_t flags;     // COPY_ITEM_*
	uint32_t copy;      // first record: copy index
	uint32_t xy;        // first record: block origin x | y << 16
	uint32_t rec_begin; // further records
	uint32_t rec_count;
	uint32_t pad[2];
};
static_assert(sizeof(CopyItem) == 32, "CopyItem layout");

enum : uint32_t
{
	COPY_ITEM_TO_NEXT_SNAPSHOT = 1u << 0, // the next level snapshots this block: store the result there too
	COPY_ITEM_SNAPSHOT = 1u << 1,         // copy the block from VRAM into the next snapshot only
};
}
