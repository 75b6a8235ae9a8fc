// Primitive assembly and render-pass batching (host side of drawing).
//
// Counterpart of the drawing half of GSInterface (gs/gs_interface.cpp:845-2943): vertex
// kicks, primitive assembly with bounding box / scissor, degenerate-draw rejection,
// parallelogram fusion, state-vector dedupe and primitive state bits. Primitives accumulate
// in host arrays until the pass ends; Renderer::submit_render_pass uploads them and runs the
// setup / binning / shading kernels.
#pragma once
#include <algorithm>
#include <array>
#include "gs_defs.hpp"
#include "gs_raster.hpp"
#include "renderer.hpp"
#include "texture_layout.hpp"
#include <vector>
#include <unordered_map>

namespace pgs
{
struct RegisterState;

struct RenderPassDesc
{
	const VertexPosition *positions;
	const VertexAttribute *attributes;
	const PrimitiveAttribute *prims;
	const StateVector *states;
	uint32_t num_primitives;
	uint32_t num_states;
	uint32_t fb_psm, fbp, fbw, z_psm, zbp;
	int32_t bb[4];            // union of the primitives' scissor-clipped boxes (pixels)
	uint32_t opaque_fbmask;
	uint32_t rate_log2;
	bool z_sensitive;         // some primitive tests or writes depth
	bool z_write;             // some primitive writes depth
	bool z_feedback;          // some primitive reads its own depth as a texture (renders single-sampled)
	bool arrays;              // some texture of the pass is super-sampled (TEX_ARRAYED)
	uint32_t zbuf_psm;        // ZBUF.PSM as written (0x30 | low bits), before format deduction
	// Textures of the pass, decoded from VRAM before shading.
	const TexDecode *decodes;
	const TextureDesc *tex_descs;
	uint32_t num_textures;
	uint32_t tex_groups;      // 256-texel decode workgroups over all textures
	bool sparse_decode;       // decode only the 8x8 blocks primitives can sample (sampler feedback)
};

// Write generations of the VRAM pages and of the CLUT, for the texture cache: a decoded image
// stays valid while no page it was decoded from has been written since, and (palette formats)
// the CLUT has not been reloaded. Writes are logged when they are queued; the GPU executes
// them in the same order as the passes that read the cache.
class VramWriteLog
{
public:
	void write(PageSpan s)
	{
		clock_++;
		if (s.last - s.first + 1u >= kNumPages)
		{
			page_.fill(clock_);
			return;
		}
		for (uint32_t p = s.first; p <= s.last; p++)
			page_[p % kNumPages] = clock_;
	}
	void write_all()
	{
		clock_++;
		page_.fill(clock_);
		may_ss_.fill(false);
	}
	// Pages whose last write was a render pass's frame-buffer or depth write (`fb`), after which a
	// texture based there may be super-sampled; transfers and host writes make them
	// single-sampled (gs/page_tracker.cpp:147-176, 255-265, 650-660, 700-711).
	void mark_super_sample(const PageRect &r, bool fb)
	{
		for (uint32_t y = 0; y < r.height; y++)
			for (uint32_t x = 0; x < r.width; x++)
				may_ss_[(r.base + y * r.stride + x) % kNumPages] = fb;
	}
	bool may_super_sample(uint32_t page) const { return may_ss_[page % kNumPages]; }
	void clut_loaded() { clut_gen_++; }
	uint64_t now() const { return clock_; }
	uint64_t clut_generation() const { return clut_gen_; }
	// Pages written after time t (now() at that point), ascending.
	std::vector<uint32_t> pages_written_since(uint64_t t) const
	{
		std::vector<uint32_t> pages;
		for (uint32_t p = 0; p < kNumPages; p++)
			if (page_[p] > t)
				pages.push_back(p);
		return pages;
	}
	bool written_since(PageSpan s, uint64_t t) const
	{
		if (s.last - s.first + 1u >= kNumPages)
			return std::any_of(page_.begin(), page_.end(), [&](uint64_t g) { return g > t; });
		for (uint32_t p = s.first; p <= s.last; p++)
			if (page_[p % kNumPages] > t)
				return true;
		return false;
	}

private:
	uint64_t clock_ = 1, clut_gen_ = 0;
	std::array<uint64_t, kNumPages> page_{};
	std::array<bool, kNumPages> may_ss_{};
};

// Per-pass texture limits: slot index width of PrimitiveAttribute.tex and the decoded image
// arena (dwords).
constexpr uint32_t kMaxPassTextures = 256;
constexpr uint32_t kTexImageTexels = 16u << 20;

class RenderPassBuilder
{
public:
	void init(Renderer *renderer, uint32_t sample_log2);
	// A new sample count; the caller has flushed the pass.
	void set_sample_log2(uint32_t sample_log2, bool sparse = false);
	bool has_primitives() const { return !prims_.empty(); }
	void reset_vertex_queue() { vq_count_ = 0; }

	// A register that feeds primitive state changed value (blocks parallelogram fusion).
	void mark_state_dirty() { state_dirty_ = true; }

	// XYZ2/XYZF2/XYZ3/XYZF3 kick (gs/gs_interface.cpp:858-899, 2883-2928).
	void vertex_kick(const RegisterState &r, uint32_t x, uint32_t y, uint32_t z, float fog, bool adc);

	// Ends the render pass: submits all pending primitives.
	void flush();

	// VRAM and CLUT writes the texture cache must see (transfers, CLUT loads, host writes).
	VramWriteLog &writes() { return writes_; }
	// Forgets every cached texture image (a recording must decode everything it samples).
	void clear_texture_cache() { cache_.clear(); }

	// Callbacks into the front end: before_first_primitive runs when a pass starts (pending
	// transfers must execute first); before_texture_read runs before a textured primitive is
	// appended (a partially received upload is written first, gs/gs_interface.cpp:2640-2650).
	struct Host
	{
		virtual ~Host() = default;
		virtual void before_first_primitive() = 0;
		virtual void before_texture_read() = 0;
	};
	Host *host = nullptr;

	// Vsyncs left during which progressive output is not forced (set by tall scan-masked
	// primitives, counted down by GSInterface::vsync).
	uint32_t complex_scanmsk_timeout = 0;

	// Hacks::disable_mipmaps: textures decode level 0 only and sample it at LOD 0.
	bool disable_mipmaps = false;

	// Super-sampled textures (gs/gs_interface.cpp:65-77, 1593-1603): at 2x or more, sprites may
	// sample frame-buffer pages through a texture with one layer per sample.
	bool super_sampled_textures = false;
	// Does the pending pass sample a super-sampled texture?
	bool pass_has_arrays() const { return pass_arrays_; }

private:
	struct QueuedVertex
	{
		VertexPosition pos;
		VertexAttribute attr;
	};

	Renderer *renderer_ = nullptr;
	uint32_t sample_log2_ = 0;
	uint32_t rate_log2_ = 0;
	QueuedVertex vq_[3];
	uint32_t vq_count_ = 0;

	std::vector<VertexPosition> pos_;
	std::vector<VertexAttribute> attr_;
	std::vector<PrimitiveAttribute> prims_;
	std::vector<StateVector> states_;
	std::unordered_map<uint64_t, uint32_t> state_map_;
	uint64_t frame_ = 0, zbuf_ = 0;
	int32_t bb_[4] = { 0, 0, -1, -1 };
	uint32_t color_write_mask_ = 0;
	bool pass_z_ = false;      // some primitive tests or writes depth
	bool pass_zwrite_ = false; // some primitive writes depth
	// On-tile texture feedback of the pass (gs/gs_interface.cpp:909-933): one kind (colour or
	// depth) and one texture format per pass; depth feedback renders the pass single-sampled.
	uint32_t pass_feedback_ = 0; // 0 none, 1 colour, 2 depth
	uint32_t pass_feedback_psm_ = 0, pass_feedback_cpsm_ = 0;
	bool pass_z_feedback_ = false;
	bool pass_arrays_ = false; // some texture of the pass is super-sampled

	// The image depends on TEX0 (without CLD), CLAMP, TEXA, the level count and the level
	// addresses in MIPTBP1 / MIPTBP2 (unused levels' fields cleared).
	struct SlotKey
	{
		uint64_t tex0, clamp, texa, mip1, mip2;
		uint32_t levels;
		uint32_t samples; // super-sampled texture: samples per pixel, else 0
		bool operator==(const SlotKey &o) const
		{
			return tex0 == o.tex0 && clamp == o.clamp && texa == o.texa && mip1 == o.mip1 && mip2 == o.mip2 &&
			       levels == o.levels && samples == o.samples;
		}
	};
	// `sprite`: the primitive being drawn is a sprite (only sprites sample super-sampled textures).
	SlotKey slot_key(const RegisterState &r, uint32_t ctx, bool sprite = false) const;
	std::vector<SlotKey> slot_keys_;
	// Texture cache across passes (gs/gs_renderer.cpp:1317-1386 create_cached_texture, with
	// page_tracker.cpp:842-933 invalidation): fully decoded images of earlier passes, kept in
	// the image arena used as a ring.
	struct CachedTexture
	{
		SlotKey key;
		uint32_t offset, texels;
		uint64_t decoded_at, clut_gen;
		bool palette;
		TexSpans span;
	};
	struct SlotImage
	{
		uint32_t offset, texels;
		bool decoded; // decoded by this pass (else reused from the cache)
		CachedTexture entry;
	};
	std::vector<CachedTexture> cache_;
	std::vector<SlotImage> slot_images_;
	uint32_t ring_ = 0; // next free texel of the image arena
	VramWriteLog writes_;
	bool allocate_image(uint32_t texels, uint32_t &offset);
	std::vector<TexDecode> decodes_;
	std::vector<TextureDesc> tex_descs_;
	uint32_t texel_total_ = 0, group_total_ = 0;
	SlotKey span_key_{ ~0ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0u, 0u }; // last texture whose page spans are cached
	TexSpans span_{};
	bool state_dirty_ = true;
	bool last_candidate_ = false;
	int32_t last_order_[3] = { 0, 1, 2 };

	bool draw_is_degenerate(const RegisterState &r) const;
	void mark_blur_kernels();
	void append(const RegisterState &r, const QueuedVertex *v, uint32_t n, bool quad);
	uint32_t state_index(const StateVector &sv);
	// Texture slot of the current TEX0 / CLAMP / TEXA; -1 when the pass has no room for it.
	int32_t texture_slot(const RegisterState &r, uint32_t ctx, bool sprite);
	// How a textured primitive reads the frame or depth buffer it draws to.
	enum class Feedback { None, Pixel, Bypass };
	Feedback feedback_mode(const RegisterState &r, uint32_t ctx, const VertexPosition *p, const VertexAttribute *a,
	                       uint32_t n, bool quad, const int32_t bb[4], bool &depth) const;
	// Does the texture of `ctx` read pages the pending pass writes?
	bool texture_hazard(const RegisterState &r, uint32_t ctx);
};
}
