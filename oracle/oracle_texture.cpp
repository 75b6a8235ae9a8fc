// CPU oracle: CLUT loads, texture cache decode and texture sampling (TEST INFRASTRUCTURE ONLY).
#include "oracle_render.hpp"
#include <cstring>

namespace orc
{
void GS::load_clut(uint32_t) {}

uint32_t GS::texture_slot(uint32_t ctx)
{
	TextureSlot t{};
	t.tex0 = regs.tex0[ctx];
	rp_textures.push_back(t);
	return uint32_t(rp_textures.size() - 1);
}

void sample_texture(const TextureSlot &, uint32_t, float, float, float, bool, float out[4])
{
	out[0] = out[1] = out[2] = out[3] = 128.0f;
}
}
