// Render-pass records shared by the host front end and the raster kernels.
//
// Field layouts follow the reference's shared structs (gs/shaders/data_structures.h:32-80)
// because the front end and the kernels exchange them; state/blend/combiner bit positions are
// the reference's (data_structures.h:206-307).
#pragma once
#include "gs_defs.hpp"

namespace pgs
{
enum : uint32_t
{
	STATE_INDEX_MASK = 0xfffu,
	ST_Z_TEST = 1u << 12, ST_Z_WRITE = 1u << 13, ST_OPAQUE = 1u << 14, ST_PARALLELOGRAM = 1u << 15,
	ST_MULTISAMPLE = 1u << 16, ST_Z_GREATER = 1u << 17, ST_PERSPECTIVE = 1u << 18, ST_IIP = 1u << 19,
	ST_FIX = 1u << 20, ST_SPRITE = 1u << 21, ST_LINE = 1u << 22, ST_SCANMSK_EVEN = 1u << 23,
	ST_SCANMSK_ODD = 1u << 24, ST_SNAP_RASTER = 1u << 25, ST_SNAP_ATTRIBUTE = 1u << 26,
	ST_PROVOKING_SHIFT = 27,
};

enum : uint32_t
{
	TEX_INDEX_MASK = 0xffffu,
	TEX_MAG_LINEAR = 1u << 16, TEX_MIN_LINEAR = 1u << 17, TEX_MIP_LINEAR = 1u << 18,
	TEX_CLAMP_S = 1u << 19, TEX_CLAMP_T = 1u << 20,
	TEX_MAX_MIP_LEVEL_SHIFT = 21, // TEX1.MXL of a mipmapped primitive, 3 bits
	// Super-sampled textures (data_structures.h:251-253): the texture has one layer per sample
	// after the single-sampled layer 0 (TEX_ARRAYED, set for every primitive that samples it);
	// sample through the phase filter (PER_SAMPLE), each sample from its own layer
	// (SAMPLE_MAPPING), or layer 0 (SAMPLE_RESOLVED, or PER_SAMPLE cleared by setup).
	TEX_PER_SAMPLE = 1u << 24, TEX_SAMPLE_MAPPING = 1u << 25, TEX_SAMPLE_RESOLVED = 1u << 26,
	TEX_ARRAYED = 1u << 27,
	// Host hint for triangle setup: the pass is a deduced blur (gs/gs_renderer.cpp:1848-1885) and
	// the primitive blends Cs * alpha + Cd without a depth test (triangle_setup.comp:713-722).
	TEX_BLUR_KERNEL = 1u << 28,
	TEX2_FIXED_LOD = 1u << 0,     // LCM; L in bits 1..2, K (signed, 1/16 units) in bits 3..14
	TEX2_L_SHIFT = 1, TEX2_K_SHIFT = 3,
	// On-tile frame-buffer feedback (the texel is the pixel's own colour or depth): bit 15 of the
	// index marks it and bits 0..4 hold CSA; the texture's PSM, CPSM and the depth source sit in
	// bits 21..31 (a feedback primitive is never mipmapped), TEXA in tex2.
	TEX_FEEDBACK = 1u << 15, TEX_FEEDBACK_PSM_SHIFT = 21, TEX_FEEDBACK_CPSM_SHIFT = 27,
	TEX_FEEDBACK_DEPTH = 1u << 31,
	TEX2_FEEDBACK_AEM = 1u << 1, TEX2_FEEDBACK_TA0_SHIFT = 2, TEX2_FEEDBACK_TA1_SHIFT = 10,
};

enum : uint32_t
{
	COMB_MODE_MASK = 3u, COMB_TME = 1u << 2, COMB_FOG = 1u << 3, COMB_TCC = 1u << 4,
	BLEND_ATE = 1u << 1, BLEND_DATE = 1u << 7, BLEND_DATM = 1u << 8, BLEND_ABE = 1u << 9,
	BLEND_PABE = 1u << 10, BLEND_COLCLAMP = 1u << 19, BLEND_FBA = 1u << 20, BLEND_DTHE = 1u << 21,
};

struct VertexPosition
{
	int32_t x, y;
	uint32_t z;
	int32_t pad;
};

struct VertexAttribute
{
	float s, t, q;
	uint32_t rgba;
	float fog;
	uint32_t uv; // u | v << 16, 10.4 fixed point
};

struct PrimitiveAttribute
{
	int16_t bb[4];
	uint32_t state, tex, tex2, alpha, fbmsk, fogcol;
};
static_assert(sizeof(PrimitiveAttribute) == 32, "PrimitiveAttribute layout");

struct StateVector
{
	uint32_t combiner, blend_mode, dimx0, dimx1;
};

struct PrimitiveSetup
{
	int32_t a[3];
	float inv_area;
	int32_t b[3];
	float error_i;
	int32_t c[3];
	float error_j;
	int32_t bb[4];
	uint32_t z[4];
};
static_assert(sizeof(PrimitiveSetup) == 80, "PrimitiveSetup layout");

struct TransformedAttributes
{
	float stqf0[4], stqf1[4], stqf2[4];
	uint32_t rgba[3];
	uint32_t pad;
	// The vertex colours as unorm floats, written once by setup for the shader's interpolation:
	// c0 = rgba[0] / 255 and the differences c1 - c0, c2 - c0, per channel, each rounded as the
	// shader's own single-precision operations would round it.
	float col0[4], dcol1[4], dcol2[4];
};
static_assert(sizeof(TransformedAttributes) == 112, "TransformedAttributes layout");

// One cached texture of a render pass: a decoded RGBA8 image plus sampling parameters
// (gs/shaders/data_structures.h:98-105 TexInfo content, gs/gs_interface.cpp:1754-1781).
struct TextureDesc
{
	uint32_t image_offset; // offset in dwords into the texture image arena
	uint32_t width, height; // effective rectangle size (the decoded image's level 0)
	uint32_t levels;        // mip levels, stored one after another from image_offset
	// Clamp bounds of the texel coordinate (min u, min v, max u, max v): the sampler's
	// floor(region) + 0.5 on an axis with CLAMP or REGION_CLAMP, -inf / +inf on the others, so
	// that the clamp is unconditional; min <= max (a region with min > max clamps everything to
	// max, so it is stored as [max, max]).
	float region[4];
	float size[4];         // 2^TW, 2^TH, 1 / width, 1 / height
	float bias[2];         // -rectangle origin / rectangle size
	float fwidth, fheight; // width and height as floats
};
static_assert(sizeof(TextureDesc) == 64, "TextureDesc layout");
}
