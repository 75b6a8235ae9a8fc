// CPU oracle: primitive assembly and per-pixel rendering (TEST INFRASTRUCTURE ONLY).
#include "oracle_render.hpp"
#include <cstring>

namespace orc
{
void GS::flush()
{
	partial_flush_transfer();
}

uint64_t GS::stat(uint32_t which) const
{
	switch (which)
	{
	case 0: return covered_samples;
	case 1: return primitives_drawn;
	default: return 0;
	}
}

void GS::write_tex0(uint32_t ctx, uint64_t v) { regs.tex0[ctx] = v; }
void GS::write_tex2(uint32_t ctx, uint64_t v) { regs.tex0[ctx] = v; }
void GS::kick_xyz(uint64_t, bool) {}
void GS::kick_xyzf(uint64_t, bool) {}
void GS::push_vertex(int32_t, int32_t, uint32_t, float) {}
void GS::drawing_kick(bool) {}
void GS::draw(const Vertex *, uint32_t, uint32_t) {}
void GS::load_clut(uint32_t) {}
}
