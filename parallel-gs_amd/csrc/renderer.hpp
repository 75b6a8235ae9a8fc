// Device-side renderer of the MI355X GS backend: owns HBM buffers and issues kernels.
//
// Counterpart of GSRenderer (gs/gs_renderer.hpp:244-324). The Vulkan command-buffer,
// descriptor and timeline machinery is replaced by one in-order HIP stream, a pinned
// upload arena mirrored in device memory, and the kernel-level C ABI (pgs_kernels.h).
#pragma once
#include "gs_defs.hpp"
#include "gs_raster.hpp"
#include "pgs_kernels.h"
#include <vector>
#include <cstring>
#include <cstddef>

namespace pgs
{
struct RenderPassDesc;

struct FlushStats
{
	uint64_t transfer_batches = 0;
	uint64_t copies = 0;
	uint64_t copy_blocks = 0;
	uint64_t copy_records = 0;
	uint64_t snapshot_blocks = 0;
	uint64_t copy_bytes = 0;       // algorithmic bytes: pixels x bpp/8, read + write
	uint64_t render_passes = 0;
	uint64_t primitives = 0;
	uint64_t syncs = 0;
	uint64_t textures_decoded = 0;   // textures decoded by render passes
	uint64_t texture_cache_hits = 0; // textures a pass reused from an earlier pass
};

class Renderer
{
public:
	Renderer() = default;
	~Renderer();
	Renderer(const Renderer &) = delete;
	Renderer &operator=(const Renderer &) = delete;

	// sample_log2: 0 = 1x, else log2 of the sample count; allocates 2 + samples planes.
	// sampler_feedback: 0 decodes sparsely when a pass decodes at least kFeedbackAutoTexels
	// texels, 1 always, 2 never.
	// texture_stream: run VRAM copies, CLUT loads and texture decodes on a second stream,
	// overlapping triangle setup, the sampling heuristic and binning.
	// capacity_log2 (>= sample_log2): the largest sample count set_sample_log2 may switch to;
	// its planes are allocated up front.
	bool init(uint32_t sample_log2, size_t arena_bytes, uint32_t sampler_feedback = 0, bool texture_stream = false,
	          uint32_t capacity_log2 = 0);
	// Switches the sample count (GSRenderer::invalidate_super_sampling_state,
	// gs/gs_renderer.cpp:315-342): waits for the GPU, then clears every plane above plane 0.
	// False when the rate exceeds the allocated capacity or a recording is open.
	bool set_sample_log2(uint32_t sample_log2);
	// Sparse sampling grids for 4x and 16x (set before the sample count).
	void set_sparse_grid(bool sparse) { sparse_ = sparse; }
	void rates(uint32_t &rate_x, uint32_t &rate_y) const { grid_rates(sample_log2_, sparse_, rate_x, rate_y); }
	uint32_t capacity_log2() const { return capacity_log2_; }
	void set_sampler_feedback(uint32_t mode) { sampler_feedback_ = mode; }
	// Adds the elapsed time of every timed launch whose events have completed to ms[kind]
	// (without waiting), in launch order; with wait, first waits for the GPU and drains all.
	void poll_timings(double *ms, uint32_t kinds, bool wait);
	bool profiling() const { return profiling_; }

	// ---- transfers -------------------------------------------------------------------
	// Queue a host->local upload. The payload bytes are copied into the upload arena.
	void queue_host_copy(const CopyDesc &desc, const void *payload, size_t bytes);
	// Queue a local->local copy.
	void queue_local_copy(const CopyDesc &desc);
	// Close the current transfer batch: build the per-block work lists and launch.
	void flush_transfers();
	bool has_pending_transfers() const { return !batch_.empty(); }
	// No submitted work is still running.
	bool idle() const
	{
		return hipStreamQuery(stream_) == hipSuccess && (!side_ || hipStreamQuery(side_) == hipSuccess);
	}

	// ---- render passes ----------------------------------------------------------------
	// Uploads the pass's primitive arrays and runs triangle setup, binning and shading.
	void submit_render_pass(const RenderPassDesc &desc);
	// Coarse tile size (log2 pixels) used at 1x; super-sampled passes use the reference's
	// binning-cost heuristic because the tile grid decides super- vs single-sampled shading.
	void set_tile_log2(uint32_t v) { tile_log2_1x_ = v; }
	uint32_t sampler_feedback_mode() const { return sampler_feedback_; }
	void note_textures(uint32_t decoded, uint32_t hits)
	{
		stats_.textures_decoded += decoded;
		stats_.texture_cache_hits += hits;
	}
	// Runs a CLUT load (the caller has ordered it after every VRAM write it must see).
	void clut_load(const pgs_clut_args &args);

	// ---- host access ------------------------------------------------------------------
	void sync();
	// Copies `planes` * 4 MiB of device VRAM into `dst` (plane 0 first).
	void read_vram(void *dst, uint32_t first_plane, uint32_t planes);
	void write_vram(const void *src, size_t offset, size_t size);
	uint32_t num_planes() const { return num_planes_; }
	uint32_t sample_log2() const { return sample_log2_; }

	// ---- launch recording for graph replay ----------------------------------------------
	// While recording, uploads are kept alive and every launch is appended to a list that
	// replay() re-issues (optionally captured into one hipGraph).
	// Device work issued by the front end itself (the CRTC scanout): returns the stream on which
	// it runs after every queued transfer and launch; end_external_work() records that it wrote
	// VRAM, so later texture work on the side stream waits for it.
	hipStream_t begin_external_work()
	{
		flush_transfers();
		arena_upload();
		join_texture_stream();
		return stream_;
	}
	void end_external_work(bool wrote_vram)
	{
		if (wrote_vram)
			main_pending_ = true;
	}
	bool recording() const { return recording_; }

	void begin_recording();
	void end_recording();
	int replay(bool use_graph);
	size_t recorded_launches() const { return recorded_.size(); }
	bool texture_stream() const { return side_ != nullptr; }

	// Per-launch HIP event timing: while enabled every launch is bracketed by two events;
	// profile_collect() synchronizes and returns the summed milliseconds per kernel kind.
	void profile_enable(bool enable);
	// kind: 0 = VRAM copy, 1 = snapshot, further kinds are the render kernels.
	void profile_collect(double *ms, uint64_t *launches, uint32_t kinds);

	hipStream_t stream() const { return stream_; }
	FlushStats consume_stats() { FlushStats s = stats_; stats_ = {}; return s; }
	const FlushStats &stats() const { return stats_; }
	int last_error() const { return last_error_; }

	uint8_t *vram_device() { return vram_; }
	size_t lists_capacity() const { return lists_cap_; }
	size_t counts_capacity() const { return counts_cap_; }

	// A kernel launch as recorded for replay (kind + argument block). Fork and Join are the
	// ordering points between the two streams, recorded in place so that replays (and graph
	// captures) keep the same dependencies: Fork makes the texture stream wait for everything
	// issued on the main stream so far, Join makes the main stream wait for the texture stream.
	enum class LaunchKind : uint32_t
	{
		VramCopy = 0, VramSnapshot = 1, Setup = 2, Binning = 3, Shade = 4, ClutWrite = 5, TexDecode = 6,
		SsaaHeuristic = 7, ShadeSuper = 8, SamplerFeedback = 9, Fork = 10, Join = 11,
		// Replays only: independent launches of one dependency level run as one (pgs_k_fused).
		Fused = 12
	};
	// Profile counter of a replay's VRAM restore (after the launch kinds).
	static constexpr uint32_t kProfileRestore = 13;
	struct Launch
	{
		LaunchKind kind;
		union
		{
			pgs_vram_copy_args copy;
			pgs_vram_snapshot_args snapshot;
			pgs_setup_args setup;
			pgs_bin_args bin;
			pgs_shade_args shade;
			pgs_clut_args clut;
			pgs_tex_decode_args decode;
			pgs_ssaa_heuristic_args heuristic;
			pgs_feedback_args feedback;
			pgs_fused_args fused;
		};
	};
	// The recorded kernel launches, in issue order (without the stream ordering points).
	const std::vector<Launch> &recorded() const { return recorded_; }
	// Restores VRAM planes and the CLUT to their state at begin_recording() (on `stream`).
	void restore_snapshot(hipStream_t stream);
	// After end_recording(): the 8 KiB pages (of every plane) the recording wrote; replays then
	// restore only those and the CLUT, with one copy launch.
	void set_restore_pages(const std::vector<uint32_t> &pages);
	const std::vector<pgs_copy_range> &restore_ranges() const { return restore_ranges_; }

private:
	struct Arena
	{
		uint8_t *host = nullptr;
		uint8_t *dev = nullptr;
		size_t cap = 0;
		size_t used = 0;      // bytes handed out
		size_t uploaded = 0;  // bytes already copied to the device
	};

	// Device scratch of a render pass, sized for the largest pass (65536 primitives).
	uint8_t *setup_buf_ = nullptr;    // PrimitiveSetup[65536]
	uint8_t *tattr_buf_ = nullptr;    // TransformedAttributes[65536]
	uint16_t *lists_buf_ = nullptr;   // coarse tile lists
	uint32_t *counts_buf_ = nullptr;  // per coarse tile
	uint32_t *dummy_buf_ = nullptr;   // bound where a pass has no textures
	uint16_t *clut_buf_ = nullptr;    // the 512-entry CLUT
	uint32_t *tex_images_ = nullptr;  // decoded RGBA8 textures of the current pass
	uint32_t *state_copy_ = nullptr;  // super-sampled passes: post-setup states
	uint32_t *flat_buckets_ = nullptr; // super-sampled passes: 256-bit bucket mask
	uint32_t *tile_modes_ = nullptr;  // super-sampled passes: per coarse tile
	uint64_t *coarse_mask_ = nullptr; // first binning level: kMaxCoarseTiles x kMaxMaskWords
	uint32_t *feedback_masks_ = nullptr; // sampler feedback: kMaxPassTextures x PGS_FEEDBACK_WORDS
	// Phase filter tables of super-sampled textures, one slot per sampling grid (rate_y * 3 +
	// rate_x), allocated and filled on first use.
	static constexpr uint32_t kPhaseLutSlots = 12;
	uint32_t *phase_lut_ = nullptr;
	uint32_t phase_lut_loaded_ = 0;
	const uint32_t *phase_lut(uint32_t rate_x, uint32_t rate_y);
	uint32_t sampler_feedback_ = 0;

public:
	// Sparse decode where a pass decodes at least this many texels (sampler_feedback 0).
	static constexpr uint64_t kFeedbackAutoTexels = 1u << 20;

private:
	static constexpr uint32_t kMaxCoarseTiles = 256, kMaxMaskWords = 1024;
	uint8_t *snapshot_ = nullptr;     // VRAM planes + CLUT at begin_recording()
	std::vector<pgs_copy_range> restore_ranges_; // the pages a recording writes, and the CLUT
	pgs_copy_range *restore_dev_ = nullptr;
	size_t lists_cap_ = 0;            // entries
	size_t counts_cap_ = 0;
	uint32_t tile_log2_1x_ = 5;

	struct PendingCopy
	{
		CopyDesc desc;
	};
	// One dependency level of a transfer flush: its copies (hazardous local copies flagged to
	// read the level-start snapshot), the blocks it writes and the blocks its snapshot holds.
	struct TransferLevel
	{
		std::vector<CopyDesc> descs;
		std::vector<uint64_t> written;   // bitset over blocks
		std::vector<uint32_t> snapshot;  // blocks the level's hazardous local copies read
	};
	void plan_transfer_level(const std::vector<uint32_t> &members, TransferLevel &lv);
	void emit_transfer_level(const TransferLevel &lv, uint32_t index, const TransferLevel *next);
	void shade_aliased_pass(const RenderPassDesc &d, const Launch &base, const uint8_t *cache_blocks,
	                        uint32_t num_cache_blocks);

	uint32_t sample_log2_ = 0;
	bool sparse_ = false;
	uint32_t capacity_log2_ = 0;
	uint32_t num_planes_ = 1;
	size_t snapshot_bytes_ = 0;
	uint8_t *vram_ = nullptr;    // num_planes_ * 4 MiB
	uint8_t *shadow_ = nullptr;  // 4 MiB
	// Copy of VRAM (every plane of the context's capacity) read by passes whose depth buffer
	// aliases their colour buffer; allocated by the first such pass.
	uint8_t *alias_cache_ = nullptr;
	hipStream_t stream_ = nullptr;
	// Texture stream: VRAM copies, snapshots, CLUT loads and (without sampler feedback) texture
	// decodes. Before one of them runs, the texture stream waits for the main stream when the
	// main stream has had work since the last such wait; the shading kernels (and a decode on the
	// main stream) wait for the texture stream when it has had work since the last join. Setup,
	// binning, the heuristic and the feedback marking touch no VRAM, CLUT or decoded image, so
	// they run beside it. Every host synchronization joins first.
	hipStream_t side_ = nullptr;
	hipEvent_t fork_event_ = nullptr, join_event_ = nullptr;
	bool main_pending_ = true;  // main stream work the texture stream has not waited for
	bool side_pending_ = false; // texture stream work the main stream has not waited for
	int last_error_ = 0;

	Arena arena_;
	size_t arena_bytes_ = 0;            // configured arena size
	std::vector<Arena> retired_arenas_; // arenas of the current recording (read by replays)
	bool recording_ = false;
	std::vector<Launch> recorded_;   // kernels
	std::vector<Launch> replay_;     // kernels and stream ordering points
	hipGraph_t graph_ = nullptr;
	hipGraphExec_t graph_exec_ = nullptr;

	struct TimedLaunch { uint32_t kind; hipEvent_t start, stop; };
	bool profiling_ = false;
	std::vector<TimedLaunch> timed_;
	std::vector<hipEvent_t> event_pool_;
	hipEvent_t get_event();

	// Current transfer batch.
	std::vector<PendingCopy> batch_;
	std::vector<int32_t> level_written_;    // per block: level of the last queued copy writing it (-1: none)
	std::vector<int32_t> level_read_;       // per block: highest level of a queued copy reading it
	size_t batch_meta_ = 0;                 // arena bytes the queued copies' metadata needs at flush
	std::vector<uint32_t> block_count_;     // per-block record counts (scratch)
	std::vector<uint32_t> touched_blocks_;  // blocks with nonzero count

	FlushStats stats_;

	void check(hipError_t err);
	// Issues a launch with the stream ordering points it needs; issue() runs one entry as is.
	void launch(const Launch &l);
	void issue(const Launch &l);
	bool on_texture_stream(const Launch &l) const;
	// Makes the main stream wait for the texture stream's outstanding work.
	void join_texture_stream();
	void sync_streams();
	// Regroups the replay list by dependency level: launches of one level that read nothing
	// another writes become one fused launch (one stream only).
	void fuse_replay();
	pgs_fused_part *fused_dev_ = nullptr; // the fused launches' parts of the current recording
	// Bump-allocates `bytes` in the arena; returns the host pointer and device offset.
	uint8_t *arena_alloc(size_t bytes, size_t &dev_offset);
	void arena_make_room(size_t bytes);
	void arena_upload();
	void arena_reset_after_sync();
	void destroy_graph();
};

// Calls f(block, x0, y0) for every 256-byte block a w x h rectangle at (sax, say) of a
// buffer (base, stride, psm) touches, in raster order of block origins. Coordinates wrap
// at 2048 as the transmission area does (vram_copy.comp:302-306).
template <typename F>
inline void for_each_rect_block(uint32_t psm, uint32_t base, uint32_t stride, uint32_t sax, uint32_t say,
                                uint32_t w, uint32_t h, F &&f)
{
	const Geometry g = psm_geometry(psm);
	const uint32_t bwid = 1u << g.bw, bhei = 1u << g.bh;
	const bool nibble = psm_storage_class(psm) == SC_4;
	for (uint32_t ry = 0; ry < h;)
	{
		uint32_t y = (say + ry) & 2047u;
		uint32_t y0 = y & ~(bhei - 1u);
		for (uint32_t rx = 0; rx < w;)
		{
			uint32_t x = (sax + rx) & 2047u;
			uint32_t x0 = x & ~(bwid - 1u);
			uint32_t addr = swizzle_addr(x0, y0, base, stride, psm);
			f(nibble ? (addr >> 9) : (addr >> 8), x0, y0);
			rx += x0 + bwid - x;
		}
		ry += y0 + bhei - y;
	}
}
}
