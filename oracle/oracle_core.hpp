// CPU oracle for the GS rendering hot path.
//
// TEST INFRASTRUCTURE ONLY. This code is the checker that the HIP path is compared
// against. Nothing in the product library may include, link or call it; only tests,
// the smoke check and the CPU-baseline leg of the benchmark load it.
//
// It is a single-threaded, immediate-mode restatement of the reference's semantics:
// every transfer and every primitive is applied to VRAM in submission order, one pixel
// at a time. Each function names the reference file:line it follows.
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>

namespace orc
{
// PSM ids (gs/shaders/data_structures.h:311-323).
enum : uint32_t
{
	CT32 = 0x00, CT24 = 0x01, CT16 = 0x02, CT16S = 0x0a,
	T8 = 0x13, T4 = 0x14, T8H = 0x1b, T4HL = 0x24, T4HH = 0x2c,
	Z32 = 0x30, Z24 = 0x31, Z16 = 0x32, Z16S = 0x3a
};

constexpr uint32_t VRAM_SIZE = 4u * 1024u * 1024u;
constexpr uint32_t VRAM_MASK = VRAM_SIZE - 1u;

uint32_t bits_per_pixel(uint32_t psm);
// Byte address for 32/16-bit formats is returned in units of the element (word/halfword),
// for PSMT8 in bytes and PSMT4 in nibbles, as gs/shaders/swizzle_utils.h:329-522 does.
uint32_t swizzle(uint32_t x, uint32_t y, uint32_t base_block, uint32_t page_stride, uint32_t psm);

struct Vec2 { int32_t x, y; };
}
