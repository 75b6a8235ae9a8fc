// CPU oracle: GS state machine declaration (TEST INFRASTRUCTURE ONLY, see oracle_core.hpp).
#pragma once
#include "oracle_core.hpp"
#include <cstdint>
#include <vector>

namespace orc
{
uint32_t read_local_pixel(const uint8_t *v, uint32_t psm, uint32_t x, uint32_t y, uint32_t base, uint32_t stride);
void write_local_pixel(uint8_t *v, uint32_t psm, uint32_t x, uint32_t y, uint32_t base, uint32_t stride, uint32_t d);

struct Copy
{
	uint32_t sbp, sbw, spsm, dbp, dbw, dpsm;
	uint32_t ssax, ssay, dsax, dsay, w, h;
};

// Per-vertex data as latched at the vertex kick (gs/gs_interface.cpp:858-899).
struct Vertex
{
	int32_t x, y;   // 12.4 fixed point, window relative (XYOFFSET subtracted)
	uint32_t z;
	float s, t, q;
	uint32_t rgba;
	float fog;
	uint32_t u, v;  // 10.4 fixed point texel coordinates
};

struct Registers
{
	uint64_t prim;
	uint32_t rgbaq;
	float rgbaq_q;
	float s, t;
	uint64_t uv;
	uint32_t fog;
	uint64_t tex0[2], clamp[2], tex1[2], xyoffset[2], scissor[2], alpha[2], test[2], fba[2], frame[2], zbuf[2];
	uint64_t miptbp1[2], miptbp2[2];
	uint64_t prmodecont, prmode, texclut, scanmsk, texa, fogcol, dimx, dthe, colclamp, pabe;
	uint64_t bitbltbuf, trxpos, trxreg, trxdir;
};

class GS
{
public:
	explicit GS(uint32_t ssaa_log2);
	void gif_transfer(uint32_t path, const void *data, size_t size);
	void write_register(uint32_t addr, uint64_t v);
	void flush();
	uint64_t stat(uint32_t which) const;

	std::vector<uint8_t> vram;   // plane 0 = VRAM; at SSAA: plane 1 = reference, 2.. = samples
	std::vector<uint16_t> clut;  // 512 x 16-bit CLUT entries (gs/shaders/clut_write.comp)
	uint32_t sample_log2 = 0;

	void (*xfer_hook)(void *, uint32_t dir, uint32_t w, uint32_t h, uint32_t psm) = nullptr;
	void *xfer_hook_user = nullptr;

private:
	struct Path { uint32_t nloop = 0, loop = 0, flg = 0, nreg = 0, reg = 0; uint64_t regs = 0; };
	struct Transfer
	{
		Copy copy{};
		bool active = false;
		uint32_t required_qwords = 0;
		uint32_t written_pixels = 0;
		std::vector<uint64_t> payload;
	};

	Registers regs{};
	float internal_q = 1.0f;
	Path paths[4];
	Transfer xfer;
	Vertex vq[3];
	uint32_t vq_count = 0;

	// Render statistics for the covered-sample count.
	uint64_t covered_samples = 0;
	uint64_t primitives_drawn = 0;

	void reset_registers();
	void packed_write(uint32_t addr, const uint64_t *w);
	void init_transfer();
	void local_copy(const Copy &c);
	void hwreg(const uint64_t *data, size_t count);
	void write_host_pixels();
	void finish_host_transfer();
	void partial_flush_transfer();
	void invalidate_ssaa(uint32_t psm, uint32_t x, uint32_t y, uint32_t base, uint32_t stride);

	// Drawing (oracle_render.cpp).
	void write_tex0(uint32_t ctx, uint64_t v);
	void write_tex2(uint32_t ctx, uint64_t v);
	void kick_xyz(uint64_t v, bool adc);
	void kick_xyzf(uint64_t v, bool adc);
	void push_vertex(int32_t x, int32_t y, uint32_t z, float fog);
	void drawing_kick(bool adc);
	void draw(const Vertex *v, uint32_t nverts, uint32_t prim_type);
	void load_clut(uint32_t ctx);
};
}
