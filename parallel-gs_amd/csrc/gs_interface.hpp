// Host front end of the MI355X GS backend: a drop-in for the reference's GSInterface
// (gs/gs_interface.hpp:251-305). GIF packets and register writes are decoded on the host;
// transfers and render passes are handed to the Renderer, which runs HIP kernels.
#pragma once
#include "gs_defs.hpp"
#include "renderer.hpp"
#include "render_pass.hpp"
#include "scanout.hpp"
#include <array>
#include <vector>
#include <cstdint>

namespace pgs
{
enum class SuperSampling : uint32_t { X1 = 1, X2 = 2, X4 = 4, X8 = 8, X16 = 16 };

struct GSOptions
{
	SuperSampling super_sampling = SuperSampling::X1;
	uint32_t vram_size = kVramSize;
	bool dynamic_super_sampling = false;
	bool ordered_super_sampling = true;
	bool super_sampled_textures = false;
	size_t upload_arena_bytes = 256u << 20;
	uint32_t sampler_feedback = 0; // 0 automatic, 1 always, 2 never
	bool texture_stream = false;   // copies, CLUT loads and decodes beside setup and binning
};

// Register file (the GS privileged-less register set the GIF can write).
struct RegisterState
{
	uint64_t prim = 0;
	uint64_t rgbaq = 0; // RGBA in the low word, Q (float bits) in the high word
	uint64_t st = 0;
	uint64_t uv = 0;
	uint64_t fog = 0;
	uint64_t tex0[2] = {}, clamp[2] = {}, tex1[2] = {}, xyoffset[2] = {}, miptbp1[2] = {}, miptbp2[2] = {};
	uint64_t scissor[2] = {}, alpha[2] = {}, test[2] = {}, fba[2] = {}, frame[2] = {}, zbuf[2] = {};
	uint64_t prmodecont = 0, prmode = 0, texclut = 0, scanmsk = 0, texa = 0, fogcol = 0;
	uint64_t dimx = 0, dthe = 0, colclamp = 0, pabe = 0;
	uint64_t bitbltbuf = 0, trxpos = 0, trxreg = 0, trxdir = 0;
	float internal_q = 1.0f;
};

// SignalInterface (gs/gs_interface.hpp:242-250) as C callbacks.
struct SignalCallbacks
{
	int (*on_signal)(void *, uint64_t) = nullptr;
	int (*on_finish)(void *, uint64_t) = nullptr;
	int (*on_label)(void *, uint64_t) = nullptr;
	void *user = nullptr;
};

// Hacks (gs/gs_interface.hpp:211-240) and DebugMode (gs/gs_interface.hpp:137-153), the fields
// this backend acts on.
struct Hacks
{
	bool disable_mipmaps = false;
	bool unsynced_readbacks = false;
};

// TimestampType (gs/gs_renderer.hpp:67-79).
enum class TimestampType : uint32_t
{
	SyncHostToVRAM, CopyVRAM, PaletteUpdate, TextureUpload, TriangleSetup, Binning, Shading, Readback, VSync, Count
};

struct GIFPath
{
	uint32_t nloop = 0, loop = 0, flg = 0, nreg = 0, reg = 0;
	uint64_t regs = 0;
	uint64_t tag_lo = 0;
};

class GSInterface : private RenderPassBuilder::Host
{
public:
	GSInterface() = default;
	bool init(const GSOptions &options);
	void reset_context_state();

	void gif_transfer(uint32_t path, const void *data, size_t size);
	void write_register(uint32_t addr, uint64_t payload);

	void *map_vram_write(size_t offset, size_t size);
	void end_vram_write(size_t offset, size_t size);
	const void *map_vram_read(size_t offset, size_t size);
	// Reads every VRAM plane (plane 0 = VRAM; at SSAA: 1 = reference, 2.. = samples).
	void read_all_planes(void *dst);

	void flush();
	void clobber_register_state();
	void read_transfer_fifo(void *data, uint32_t num_128b_words);

	// gs/gs_interface.cpp:65-118. False (nothing changes) for a rate or grid this backend does
	// not build or that exceeds the allocated planes.
	bool set_super_sampling_rate(SuperSampling rate, bool ordered_grid, bool super_sampled_textures);
	uint32_t super_sampling_rate() const { return 1u << renderer_.sample_log2(); }
	void set_hacks(const Hacks &h)
	{
		hacks_ = h;
		rp_.disable_mipmaps = h.disable_mipmaps;
	}
	void set_signal_interface(const SignalCallbacks &cb) { signal_ = cb; }
	void set_timestamps(bool enable);
	void set_sampler_feedback(uint32_t mode) { renderer_.set_sampler_feedback(mode); }
	double get_accumulated_timestamps(TimestampType type);

	// CRTC scanout (gs/gs_interface.cpp:4626-4695). Flushes pending work first. Returns 0, a
	// negative hipError_t, or kVsyncUnsupported when the request would switch the renderer to
	// field-aware super-sampling (high-resolution scanout of a field-rendered interlaced mode at
	// 2x or more), which this backend does not build.
	static constexpr int kVsyncUnsupported = -1000;
	int vsync(const VSyncInfo &info, ScanoutResult &result);
	bool vsync_can_skip(const VSyncInfo &info) const;

	RegisterState &get_register_state() { return regs_; }
	GIFPath &get_gif_path(uint32_t path) { return paths_[path & 3u]; }
	uint8_t *get_priv_register_state() { return priv_.data(); }
	const uint8_t *get_priv_register_state() const { return priv_.data(); }
	const RegisterState &get_register_state() const { return regs_; }
	const GIFPath &get_gif_path(uint32_t path) const { return paths_[path & 3u]; }
	uint32_t *cached_cbp() { return cached_cbp_; }
	FlushStats consume_flush_stats() { return renderer_.consume_stats(); }
	Renderer &renderer() { return renderer_; }
	RenderPassBuilder &render_pass() { return rp_; }

private:
	struct TransferState
	{
		CopyDesc copy{};
		bool host_to_local_active = false;
		uint32_t required_qwords = 0;
		uint32_t flushed_pixels = 0;
		std::vector<uint64_t> payload;
		std::vector<uint8_t> fifo;
		uint32_t fifo_offset_128 = 0;
		uint32_t fifo_size_128 = 0;
	};

	GSOptions options_;
	Hacks hacks_;
	SignalCallbacks signal_;
	std::array<uint8_t, 8192> priv_{};
	bool timestamps_ = false;
	double timestamp_ms_[uint32_t(TimestampType::Count)] = {};
	void poll_timestamps(bool wait);
	Renderer renderer_;
	RenderPassBuilder rp_;
	Scanout scanout_;
	RegisterState regs_;
	GIFPath paths_[4];
	TransferState xfer_;
	std::vector<uint8_t> host_vram_;
	uint32_t cached_cbp_[2] = { 0, 0 }; // CBP0 / CBP1 of CLD 2..5

	void packed_write(uint32_t addr, const uint64_t *qword);
	void write_tex0(uint32_t ctx, uint64_t v);
	void handle_clut_upload(uint32_t ctx);
	void init_transfer();
	void hwreg(const uint64_t *data, size_t count);
	void flush_pending_transfer(bool keep_alive);
	void begin_transfer_work();
	void flush_render_pass();
	void set_state_reg(uint64_t &reg, uint64_t v);
	void reset_registers();
	void write_prim(uint64_t v);
	void write_prmode(uint64_t v);
	void before_first_primitive() override;
	void before_texture_read() override;
};
}
