// CPU oracle: GS state machine declaration (TEST INFRASTRUCTURE ONLY, see oracle_core.hpp).
#pragma once
#include "oracle_core.hpp"
#include <cstdint>
#include <vector>
#include <memory>

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

// One assembled primitive of a render pass (gs/shaders/data_structures.h:55-80 content).
struct Prim
{
	Vertex v[3];
	int32_t bb[4];       // scissor-clipped pixel bounding box
	uint32_t state, tex, tex2, alpha, fbmsk, fogcol;
	uint32_t combiner, blend_mode, dimx[2];
	uint32_t tex_slot;    // texture slot in the render pass
	// On-tile feedback: 0 none, 1 the pixel's colour, 2 its depth is the texel; with the
	// texture's PSM, CLUT format and CSA.
	uint32_t feedback, fb_tex_psm, fb_tex_cpsm, fb_csa;
	uint64_t fb_texa;     // TEXA at the draw
};

// Result of triangle setup (gs/shaders/data_structures.h:32-52 content).
struct Setup
{
	int32_t a[3], b[3], c[3];
	float inv_area, error_i, error_j;
	int32_t bb[4];
	uint32_t z[3];
	float stqf0[4], stqf1[4], stqf2[4];
	uint32_t rgba[3];
	bool flat = false;     // super-sampled pass: flat primitive entered into a depth bucket
	uint32_t flat_z = 0;   // its depth (triangle_setup.comp:820-872)
};

// A decoded texture of a render pass (gs/gs_interface.cpp:1754-1781 TexInfo content).
struct TextureSlot
{
	uint64_t tex0, clamp, texa, mip1, mip2;
	uint32_t levels;            // mip levels (1 without mipmapping)
	uint32_t width, height;     // effective rectangle size (level 0)
	std::vector<uint32_t> rgba; // decoded RGBA8 levels, one after another
	uint32_t level_offset[7], level_width[7], level_height[7];
	float size[4];              // 2^TW, 2^TH, 1 / width, 1 / height
	float bias[2];              // -rect origin / rectangle size
	float region[4];            // clamp region: min u, min v, max u, max v
	// Super-sampled texture (gs/gs_interface.cpp:1593-1603): samples per pixel S, else 0; rgba
	// then holds S + 1 images: the single-sampled one, then one per sample.
	uint32_t samples = 0;
};

// Inclusive interval of 8 KiB pages touched by a rectangle of a buffer; `last` may run past
// the end of VRAM (it wraps).
struct PageSpan { uint32_t first, last; };
PageSpan page_span(uint32_t base_block, uint32_t stride, uint32_t psm, uint32_t x0, uint32_t y0, uint32_t x1,
                   uint32_t y1);
bool spans_overlap(PageSpan a, PageSpan b);

// Phase filter table of super-sampled textures (oracle_texture.cpp).
void phase_lut(uint32_t rate_x, uint32_t rate_y, std::vector<uint32_t> &out);

// Q-derived mipmap level of detail in 1/16 units (oracle_texture.cpp); true when above 0.
bool lod_from_q(float q, int32_t k, uint32_t l, int32_t &n);

// VSyncInfo (gs/gs_interface.hpp:155-198) and ScanoutResult (gs/gs_renderer.hpp:23-52) as
// flat words for ctypes.
struct VsyncInfo
{
	uint32_t phase, force_progressive, overscan, anti_blur, crtc_offsets, adapt_to_internal_horizontal_resolution,
	    raw_circuit_scanout, high_resolution_scanout, skip_deinterlace;
};
struct ScanoutOut
{
	uint32_t has_image, width, height, internal_width, internal_height, mode_width, mode_height;
	uint32_t high_resolution_scanout, interlaced, double_strike, interlace_phase;
	float sampling_rate_mhz;
};
struct ScanImage
{
	uint32_t w = 0, h = 0;
	std::vector<uint32_t> px; // RGBA8, row-major
	ScanImage(uint32_t w_, uint32_t h_) : w(w_), h(h_), px(size_t(w_) * h_, 0u) {}
};

class GS
{
public:
	explicit GS(uint32_t ssaa_log2);
	void gif_transfer(uint32_t path, const void *data, size_t size);
	void write_register(uint32_t addr, uint64_t v);
	void flush();
	void reset_context_state();
	uint64_t stat(uint32_t which) const;

	std::vector<uint8_t> vram;   // plane 0 = VRAM; at SSAA: plane 1 = reference, 2.. = samples
	std::vector<uint16_t> clut;  // current CLUT: 512 x 16-bit entries (gs/shaders/clut_write.comp)
	uint32_t sample_log2 = 0;
	// Sparse sampling grids at 4x (1 x 4) and 16x (2 x 8) instead of the ordered 2 x 2 / 4 x 4
	// (gs/gs_interface.cpp:84-113).
	bool sparse_grid = false;
	// log2 of the samples per pixel along x and y (gs/gs_interface.cpp:65-117).
	uint32_t rate_x_log2() const
	{
		if (!sample_log2)
			return 0;
		return sparse_grid && (sample_log2 == 2 || sample_log2 == 4) ? sample_log2 / 2 - 1 : sample_log2 / 2;
	}
	uint32_t rate_y_log2() const { return sample_log2 - rate_x_log2(); }

	void (*xfer_hook)(void *, uint32_t dir, uint32_t w, uint32_t h, uint32_t psm) = nullptr;
	void *xfer_hook_user = nullptr;

	// LOCAL->HOST transfer FIFO (gs/gs_interface.cpp:3150-3166, 3213-3275).
	std::vector<uint8_t> fifo;
	uint32_t fifo_offset_128 = 0;
	void read_transfer_fifo(void *data, uint32_t num_128b_words);

	// CRTC scanout (oracle_vsync.cpp). priv: the 8 KiB privileged register block.
	std::vector<uint8_t> priv = std::vector<uint8_t>(8192, 0);
	uint32_t complex_scanmsk_timeout = 0; // gs/gs_interface.cpp:1947-1960
	// Super-sampled textures (gs/gs_interface.cpp:65-78; off at 1x) and the page tracker's
	// "may super-sample" page bits (gs/page_tracker.cpp:147-176, 255-265, 650-660, 700-711).
	bool super_sampled_textures = false;
	std::vector<uint8_t> may_super_sample = std::vector<uint8_t>(512, 0);
	void mark_pages(uint32_t base_block, uint32_t stride, uint32_t psm, uint32_t x0, uint32_t y0, uint32_t x1,
	                uint32_t y1, bool fb);
	// Returns 0, or -1000 where the reference would switch to field-aware rendering.
	int vsync(const VsyncInfo &info, ScanoutOut &out, std::vector<uint32_t> &image);
	bool vsync_can_skip(const VsyncInfo &info) const;
	std::shared_ptr<ScanImage> last_fields[4];

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

	// Render pass state.
	std::vector<Prim> rp_prims;
	std::vector<TextureSlot> rp_textures;
	uint32_t cached_cbp[2] = { 0, 0 }; // CBP0 / CBP1 of CLD 2..5
	int32_t rp_bb[4] = { 0, 0, 0, 0 };  // union of the pass's primitive boxes
	bool rp_zwrite = false;            // some primitive of the pass writes depth
	bool rp_arrays = false;            // some texture of the pass is super-sampled
	uint64_t rp_frame = 0, rp_zbuf = 0;
	bool state_dirty = true;          // a draw-state register changed since the last append
	bool last_parallelogram_candidate = false;
	int32_t last_order[3] = { 0, 1, 2 };

	// Render statistics for the covered-sample count.
	uint64_t covered_samples = 0;
	uint64_t primitives_drawn = 0;
	uint64_t render_passes = 0;

	void reset_registers();
	void packed_write(uint32_t addr, const uint64_t *w);
	void init_transfer();
	void local_copy(const Copy &c);
	void local_to_host(const Copy &c);
	void hwreg(const uint64_t *data, size_t count);
	void write_host_pixels();
	void finish_host_transfer();
	void partial_flush_transfer();
	void invalidate_ssaa(uint32_t psm, uint32_t x, uint32_t y, uint32_t base, uint32_t stride);
	void set_reg(uint64_t &reg, uint64_t v, bool marks_dirty);

	// Drawing (oracle_render.cpp).
	void write_tex0(uint32_t ctx, uint64_t v);
	void write_tex2(uint32_t ctx, uint64_t v);
	void kick_xyz(uint64_t v, bool adc);
	void kick_xyzf(uint64_t v, bool adc);
	void push_vertex(int32_t x, int32_t y, uint32_t z, float fog);
	void drawing_kick(bool adc);
	void append(const Vertex *v, uint32_t nverts, bool quad);
	bool draw_is_degenerate() const;
	void flush_render_pass();
	void render(const std::vector<Prim> &prims);
	void load_clut(uint32_t ctx);
	uint32_t texture_slot(uint32_t ctx);
	std::vector<PageSpan> texture_span(uint32_t ctx) const;
	bool texture_hazard(uint32_t ctx) const;
	// 0: ordinary texture read; 1: on-tile feedback (depth set for the depth buffer); 2: the
	// texture aliases the frame buffer with another layout and is read as the pass starts.
	uint32_t feedback_kind(uint32_t ctx, const Vertex *v, uint32_t n, bool quad, const int32_t bb[4],
	                       bool &depth) const;
	void feedback_texture(const Prim &p, uint32_t in, float out[4]) const;
};

// Reciprocal helpers (gs/shaders/math_utils.h:27-101) and their tables
// (gs/gs_renderer.cpp:3600-3640).
float rcp_fixed(uint32_t b);
float rcp_float(float v);
const uint8_t *rcp_table();
}
