// Lookup tables of the raster reciprocals.
//
// kRcpSeed is the 8-bit seed of the fixed-point Newton reciprocal used for a primitive's
// inverse area. It is the reference's hand-tuned table (gs/gs_renderer.cpp:3602-3636): the
// two integer Newton steps only give the reference's inverse areas, and therefore its
// barycentrics, depths and colours, with exactly these seeds. The float reciprocal seed LUT is
// generated from its formula 1 / (1 + (i + 0.5) / 256) (gs/gs_renderer.cpp:3638-3640).
#pragma once
#include <cstdint>

namespace pgs
{
static const uint8_t kRcpSeed[256] = {
	255, 253, 251, 249, 247, 245, 243, 241, 240, 238, 236, 234, 232, 230, 229, 227,
	225, 223, 221, 220, 218, 216, 215, 213, 211, 210, 208, 206, 205, 203, 201, 200,
	198, 197, 195, 194, 192, 191, 189, 188, 186, 185, 183, 182, 180, 179, 177, 176,
	174, 173, 172, 170, 169, 167, 166, 165, 164, 162, 161, 159, 158, 157, 156, 154,
	153, 152, 150, 149, 148, 147, 145, 144, 143, 142, 140, 139, 138, 137, 136, 135,
	134, 132, 131, 130, 129, 128, 127, 126, 124, 123, 122, 121, 120, 119, 118, 116,
	116, 115, 113, 113, 112, 111, 110, 109, 107, 107, 106, 104, 103, 103, 102, 101,
	100,  99,  98,  97,  96,  95,  94,  93,  92,  91,  91,  89,  88,  88,  86,  86,
	 85,  84,  83,  82,  81,  81,  80,  79,  78,  78,  76,  75,  75,  74,  72,  72,
	 72,  70,  70,  69,  68,  67,  67,  66,  65,  64,  63,  63,  62,  61,  60,  59,
	 59,  58,  57,  56,  56,  55,  54,  54,  53,  52,  52,  51,  50,  49,  48,  48,
	 47,  46,  46,  45,  45,  44,  43,  42,  42,  41,  40,  39,  39,  38,  38,  37,
	 36,  35,  35,  34,  33,  33,  33,  32,  31,  31,  30,  29,  29,  28,  28,  27,
	 26,  25,  25,  25,  24,  23,  23,  22,  22,  21,  20,  19,  19,  18,  18,  17,
	 17,  16,  15,  15,  14,  14,  14,  13,  12,  12,  11,  11,  10,  10,   9,   8,
	  8,   7,   7,   7,   6,   5,   5,   4,   4,   3,   3,   2,   2,   1,   1,   1,
};

inline void make_rcp_float_lut(float lut[256])
{
	for (int i = 0; i < 256; i++)
		lut[i] = 1.0f / (1.0f + (float(i) + 0.5f) / 256.0f);
}
}
