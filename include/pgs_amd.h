/* Context-level C ABI of the MI355X GS backend.
 *
 * Mirrors the public API of the reference's GSInterface (gs/gs_interface.hpp:251-305)
 * one call per method, with the Vulkan device argument replaced by a HIP device index.
 * All calls for one context must come from a single thread, as in the reference.
 * Functions returning int give 0 on success and a negative value on error.
 */
#ifndef PGS_AMD_H
#define PGS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgs_context pgs_context;

/* GSOptions (gs/gs_interface.hpp:200-207). super_sampling is the sample count 1/2/4/8/16. */
typedef struct pgs_options
{
	uint32_t super_sampling;
	uint32_t ordered_super_sampling;
	uint32_t super_sampled_textures;
	uint32_t device_index;
	uint64_t upload_arena_bytes; /* 0 selects the default (256 MiB) */
	/* Sampler feedback (sparse texture decode, gs_renderer.cpp:1388-1422): 0 where a render pass
	 * decodes at least 2^20 texels, 1 always, 2 never. Never changes a pixel. */
	uint32_t sampler_feedback;
	/* 1: VRAM copies, CLUT loads and texture decodes run on a second HIP stream, beside triangle
	 * setup, the sampling heuristic and binning, with event waits between the streams; 0: one
	 * stream. Never changes a pixel. */
	uint32_t texture_stream;
	/* GSOptions::dynamic_super_sampling (gs/gs_interface.hpp:204): 1 allocates planes for 16
	 * samples so that pgs_set_super_sampling_rate can raise the rate later; 0 allocates for
	 * `super_sampling` only. */
	uint32_t dynamic_super_sampling;
} pgs_options;

/* Error codes beyond negative hipError_t values. */
#define PGS_ERROR_INVALID_ARGUMENT (-1)
/* A request the reference accepts but this backend does not implement (the call changes
 * nothing). */
#define PGS_ERROR_UNSUPPORTED (-1000)

/* FlushStats (gs/gs_renderer.hpp:55-65), extended with the backend's own counters. */
typedef struct pgs_flush_stats
{
	uint64_t transfer_batches;
	uint64_t copies;
	uint64_t copy_blocks;
	uint64_t copy_records;
	uint64_t snapshot_blocks;
	uint64_t copy_bytes;
	uint64_t render_passes;
	uint64_t primitives;
	uint64_t syncs;
	uint64_t textures_decoded;   /* textures render passes decoded */
	uint64_t texture_cache_hits; /* textures a pass reused from an earlier pass's decode */
} pgs_flush_stats;

/* GSInterface::GSInterface + GSInterface::init (gs/gs_interface.cpp:15-63). NULL on failure. */
pgs_context *pgs_create(const pgs_options *options);
void pgs_destroy(pgs_context *ctx);

/* GSInterface::reset_context_state (gs/gs_interface.cpp:26-39). */
int pgs_reset_context_state(pgs_context *ctx);

/* GSInterface::gif_transfer (gs/gs_interface.cpp:4290-4397). size is a multiple of 16, path < 4. */
int pgs_gif_transfer(pgs_context *ctx, uint32_t path, const void *data, size_t size);

/* GSInterface::write_register (gs/gs_interface.cpp:4276-4279). */
int pgs_write_register(pgs_context *ctx, uint32_t addr, uint64_t payload);

/* GSInterface::flush (gs/gs_interface.cpp:4257-4264). Submits all pending work. */
int pgs_flush(pgs_context *ctx);

/* GSInterface::map_vram_read (gs/gs_interface.cpp:4230-4255). Blocks until the GPU is idle. */
const void *pgs_map_vram_read(pgs_context *ctx, size_t offset, size_t size);

/* GSInterface::map_vram_write / end_vram_write (gs/gs_interface.cpp:4184-4228). */
void *pgs_map_vram_write(pgs_context *ctx, size_t offset, size_t size);
int pgs_end_vram_write(pgs_context *ctx, size_t offset, size_t size);

/* GSInterface::read_transfer_fifo (gs/gs_interface.cpp:3150-3166). */
int pgs_read_transfer_fifo(pgs_context *ctx, void *data, uint32_t num_128b_words);

/* GSInterface::clobber_register_state (gs/gs_interface.cpp:4266-4274). */
int pgs_clobber_register_state(pgs_context *ctx);

/* GSInterface::consume_flush_stats (gs/gs_interface.cpp:4697-4700). */
int pgs_consume_flush_stats(pgs_context *ctx, pgs_flush_stats *stats);

/* ---- Register state, GIF paths, options (gs/gs_interface.hpp:251-305). ---- */

/* ContextState (gs/gs_interface.hpp:20-34): the per-context registers as 64-bit GS words. */
typedef struct pgs_context_registers
{
	uint64_t tex0, tex1, clamp, xyoffset, miptbp1, miptbp2, scissor, alpha, test, fba, frame, zbuf;
} pgs_context_registers;

/* RegisterState (gs/gs_interface.hpp:36-61) with every register as its 64-bit GS word. */
typedef struct pgs_register_state
{
	pgs_context_registers ctx[2];
	uint64_t prim, rgbaq, st, uv, fog, prmodecont, texclut, texa, fogcol, dimx, dthe, colclamp, pabe;
	uint64_t bitbltbuf, trxpos, trxreg, trxdir, scanmsk;
	float internal_q;
	uint32_t cached_cbp[2];
	uint32_t pad;
} pgs_register_state;

/* GSInterface::get_register_state (gs/gs_interface.hpp:289-290). Writing the state back
 * (a save-state restore) takes effect like the reference's assignment through the returned
 * reference: no register side effects run; call pgs_clobber_register_state afterwards, as the
 * reference's callers do. */
int pgs_get_register_state(pgs_context *ctx, pgs_register_state *out);
int pgs_set_register_state(pgs_context *ctx, const pgs_register_state *in);

/* GIFPath (gs/gs_interface.hpp:63-68): the current GIFtag as two 64-bit words and the
 * position inside it. */
typedef struct pgs_gif_path
{
	uint64_t tag[2];
	uint32_t reg;
	uint32_t loop;
} pgs_gif_path;

/* GSInterface::get_gif_path (gs/gs_interface.hpp:295-296). path < 4. */
int pgs_get_gif_path(pgs_context *ctx, uint32_t path, pgs_gif_path *out);
int pgs_set_gif_path(pgs_context *ctx, uint32_t path, const pgs_gif_path *in);

/* PrivRegisterState (gs/gs_interface.hpp:70-111): the privileged (CRTC) registers, as the
 * 8 KiB block the EE sees them in: PMODE at byte 0x0000, SMODE1 0x0010, SMODE2 0x0020, SRFSH
 * 0x0030, SYNCH1 0x0040, SYNCH2 0x0050, SYNCV 0x0060, DISPFB1 0x0070, DISPLAY1 0x0080, DISPFB2
 * 0x0090, DISPLAY2 0x00a0, EXTBUF 0x00b0, EXTDATA 0x00c0, EXTWRITE 0x00d0, BGCOLOR 0x00e0, CSR
 * 0x1000, IMR 0x1010, BUSDIR 0x1040, SIGLBLID 0x1080 (GSInterface::get_priv_register_state,
 * gs/gs_interface.hpp:292-293). */
#define PGS_PRIV_REGISTER_BYTES 8192u
int pgs_get_priv_register_state(pgs_context *ctx, void *out, size_t size);
int pgs_set_priv_register_state(pgs_context *ctx, const void *in, size_t size);

/* GSInterface::set_super_sampling_rate (gs/gs_interface.hpp:258, gs/gs_interface.cpp:65-118).
 * Submits pending work, then switches the sample count (1, 2, 4, 8, 16) and grid, clearing
 * every super-sample plane. The rate may not exceed what the context allocated (its
 * `super_sampling`, or 16 with `dynamic_super_sampling`). ordered_grid = 0 selects the sparse
 * 1x4 (4x) and 2x8 (16x) grids (gs/gs_interface.cpp:84-113). super_sampled_textures lets sprites
 * sample frame-buffer pages through textures with one layer per sample (off at 1x,
 * gs/gs_interface.cpp:1593-1603). Returns PGS_ERROR_INVALID_ARGUMENT for other rates; the rate
 * is then unchanged. Not allowed while recording. */
int pgs_set_super_sampling_rate(pgs_context *ctx, uint32_t super_sampling, uint32_t ordered_grid,
                                uint32_t super_sampled_textures);
/* The current sample count (1, 2, 4, 8 or 16). */
uint32_t pgs_get_super_sampling_rate(pgs_context *ctx);

/* Hacks (gs/gs_interface.hpp:211-240). disable_mipmaps decodes level 0 only and samples it at
 * LOD 0 with the magnification filter (gs/gs_interface.cpp:1539-1540, 1834); unsynced_readbacks makes LOCAL->HOST transfers return zeros without waiting
 * for the GPU (the reference does so when the data is not ready); backbuffer_promotion and
 * allow_blend_demote are not built: asking for them returns PGS_ERROR_UNSUPPORTED and the
 * other fields still apply. */
typedef struct pgs_hacks
{
	uint32_t disable_mipmaps;
	uint32_t unsynced_readbacks;
	uint32_t backbuffer_promotion;
	uint32_t allow_blend_demote;
} pgs_hacks;
int pgs_set_hacks(pgs_context *ctx, const pgs_hacks *hacks);

/* DebugMode (gs/gs_interface.hpp:137-153). timestamps enables the per-kernel timers read by
 * pgs_get_accumulated_timestamps; deterministic_timeline_query is accepted (every host access
 * here waits for the GPU, which is what it asks for); disable_sampler_feedback turns sparse
 * texture decoding off. feedback_render_target and a draw_mode other than 0 (debug captures
 * of every draw) are not built: PGS_ERROR_UNSUPPORTED, the other fields still apply. */
typedef struct pgs_debug_mode
{
	uint32_t feedback_render_target;
	uint32_t timestamps;
	uint32_t deterministic_timeline_query;
	uint32_t disable_sampler_feedback;
	uint32_t draw_mode;
} pgs_debug_mode;
int pgs_set_debug_mode(pgs_context *ctx, const pgs_debug_mode *mode);

/* SignalInterface (gs/gs_interface.hpp:242-250): called on SIGNAL, FINISH and LABEL register
 * writes with the written value; a nonzero return submits pending work (GSInterface::flush).
 * NULL callbacks are skipped. The pointer is copied; `user` is passed back. */
typedef struct pgs_signal_interface
{
	int (*on_signal)(void *user, uint64_t payload);
	int (*on_finish)(void *user, uint64_t payload);
	int (*on_label)(void *user, uint64_t payload);
	void *user;
} pgs_signal_interface;
int pgs_set_signal_interface(pgs_context *ctx, const pgs_signal_interface *iface);

/* TimestampType (gs/gs_renderer.hpp:67-79). */
#define PGS_TIMESTAMP_SYNC_HOST_TO_VRAM 0u
#define PGS_TIMESTAMP_COPY_VRAM 1u
#define PGS_TIMESTAMP_PALETTE_UPDATE 2u
#define PGS_TIMESTAMP_TEXTURE_UPLOAD 3u
#define PGS_TIMESTAMP_TRIANGLE_SETUP 4u
#define PGS_TIMESTAMP_BINNING 5u
#define PGS_TIMESTAMP_SHADING 6u
#define PGS_TIMESTAMP_READBACK 7u
#define PGS_TIMESTAMP_VSYNC 8u
#define PGS_TIMESTAMP_COUNT 9u
/* GSInterface::get_accumulated_timestamps (gs/gs_interface.cpp:4702-4705): seconds of GPU time
 * spent in the kernels of `type` since timestamps were enabled (pgs_set_debug_mode), measured
 * with HIP events around each launch. Negative on error. */
double pgs_get_accumulated_timestamps(pgs_context *ctx, uint32_t type);

/* VSyncInfo (gs/gs_interface.hpp:155-198) without the Vulkan image layout / stage / access
 * fields: the scanout is a device buffer ordered on pgs_stream(ctx). Booleans are 0 / 1. */
typedef struct pgs_vsync_info
{
	uint32_t phase;                 /* field: 0 top, 1 bottom */
	uint32_t force_progressive;
	uint32_t overscan;
	uint32_t anti_blur;
	uint32_t crtc_offsets;
	uint32_t adapt_to_internal_horizontal_resolution;
	uint32_t raw_circuit_scanout;
	uint32_t high_resolution_scanout;
	uint32_t skip_deinterlace;
} pgs_vsync_info;

/* ScanoutResult (gs/gs_renderer.hpp:23-52). `image` is a device pointer to width x height
 * RGBA8 pixels (R in the low byte, row pitch width * 4 bytes), owned by the context and valid
 * until its next pgs_vsync or pgs_destroy; it is NULL (and everything else 0) when the video
 * mode is unknown. */
typedef struct pgs_scanout
{
	const void *image;
	uint32_t width, height;
	uint32_t internal_width, internal_height; /* single-sampled domain */
	uint32_t mode_width, mode_height;
	uint32_t high_resolution_scanout;
	uint32_t interlaced;
	uint32_t double_strike;
	uint32_t interlace_phase;
	float sampling_rate_mhz;
} pgs_scanout;

/* GSInterface::vsync (gs/gs_interface.cpp:4626-4690, gs/gs_renderer.cpp:4189-5055): scans out
 * the privileged registers' display circuits (pgs_set_priv_register_state) into a frame:
 * read circuits (super-sample aware at high resolution), merge over BGCOLOR, EXTWRITE feedback
 * into VRAM, and FastMAD deinterlacing with a four-field history. Pending transfers and
 * primitives are flushed first. The kernels run asynchronously on pgs_stream(ctx).
 * Returns 0, a negative hipError_t, or PGS_ERROR_UNSUPPORTED for a request that would turn
 * on the reference's field-aware super-sampled rendering (high_resolution_scanout of an
 * interlaced FFMD = 1 mode at 2x or more), for an enabled read circuit whose DISPFB.PSM is not
 * PSMCT32/24/16/16S, and while recording. */
int pgs_vsync(pgs_context *ctx, const pgs_vsync_info *info, pgs_scanout *result);

/* GSInterface::vsync_can_skip (gs/gs_interface.cpp:4692-4695): 1 when a frame with no new
 * work may reuse the previous scanout (the output is not interlaced and no scan-masked
 * primitive asked for field order), else 0; negative on error. */
int pgs_vsync_can_skip(pgs_context *ctx, const pgs_vsync_info *info);

/* Copies the last scanout image to host memory (width * height * 4 bytes). Blocks until the
 * scanout kernels have finished. Backend extension: the reference hands out a Vulkan image. */
int pgs_read_scanout(pgs_context *ctx, void *dst, size_t size);

/* ---- GS dump files: the PCSX2 GS-dump format, state versions 8 and 9 ----------------------
 * A dump is a header, the register state, 4 MiB of VRAM, the four GIF paths, the internal Q
 * and the privileged registers, then Transfer / Vsync / ReadFIFO / PrivRegisters packets. */
typedef struct pgs_dump_reader pgs_dump_reader;

/* GSDumpParser::open / open_raw (dump/gs_dump_parser.cpp:63-82): opens `path` and, unless
 * `raw` (packets only), loads its start state into `ctx` (VRAM, registers, GIF paths, privileged
 * registers). NULL on failure. The context must outlive the reader. */
pgs_dump_reader *pgs_dump_open(pgs_context *ctx, const char *path, uint32_t raw);
/* GSDumpParser::restart (dump/gs_dump_parser.cpp:6-60): rewinds and loads the start state again
 * (a save-state restore). 0 on success. */
int pgs_dump_restart(pgs_dump_reader *reader);
/* GSDumpParser::iterate_until_vsync (dump/gs_dump_parser.cpp:137-206): replays packets until a
 * vsync that follows GIF traffic (every vsync with `conservative_crtc`) or a ReadFIFO packet.
 * Vsyncs run pgs_vsync with the parser's VSyncInfo. Returns 1 when it stopped there, 0 at the end
 * of the file, negative on error. */
int pgs_dump_iterate_until_vsync(pgs_dump_reader *reader, uint32_t high_res_scanout, uint32_t conservative_crtc);
/* GSDumpParser::consume_vsync_result (dump/gs_dump_parser.cpp:208-213): the last vsync's frame
 * (then cleared); it also becomes the frame pgs_read_scanout copies. */
int pgs_dump_consume_vsync_result(pgs_dump_reader *reader, pgs_scanout *result);
void pgs_dump_close(pgs_dump_reader *reader);

typedef struct pgs_dump_writer pgs_dump_writer;
/* GSDumpGenerator::init (dump/gs_dump_generator.cpp:199-237): writes the header and the
 * context's current state (flushes it first). NULL on failure. */
pgs_dump_writer *pgs_dump_writer_create(pgs_context *ctx, const char *path);
/* GSDumpGenerator::write_register / write_packed / write_vsync / write_image_upload
 * (dump/gs_dump_generator.cpp:107-197). The writer only records: it does not send the packets
 * to the context. write_vsync stores the context's privileged registers before the vsync
 * packet. Image uploads use IMAGE packets of at most 32767 qwords. */
int pgs_dump_write_register(pgs_dump_writer *writer, uint32_t addr, uint64_t payload);
int pgs_dump_write_packed(pgs_dump_writer *writer, uint64_t prim, const uint32_t *registers,
                          uint32_t num_registers, uint32_t num_loops, const void *data);
int pgs_dump_write_vsync(pgs_dump_writer *writer, uint32_t field);
int pgs_dump_write_image_upload(pgs_dump_writer *writer, uint32_t addr, uint32_t psm, uint32_t width,
                                uint32_t height, const void *data, size_t size);
/* Flushes and closes the file; 0 when every write succeeded. */
int pgs_dump_writer_close(pgs_dump_writer *writer);

/* Number of 4 MiB VRAM planes: 1 at 1x, 2 + samples with super-sampling
 * (plane layout of GSRenderer::init_vram, gs/gs_renderer.cpp:360-427). */
uint32_t pgs_num_planes(pgs_context *ctx);

/* Copies every plane to dst (pgs_num_planes * 4 MiB). Blocks until the GPU is idle.
 * Backend extension used by parity tests: the reference exposes plane 0 only. */
int pgs_read_vram_planes(pgs_context *ctx, void *dst);

/* Launch recording: work flushed between begin/end is kept resident in device memory and
 * can be re-issued with pgs_replay (optionally as one hipGraph). Backend extension used to
 * time device execution separately from host parsing. */
int pgs_begin_recording(pgs_context *ctx);
int pgs_end_recording(pgs_context *ctx);
int pgs_replay(pgs_context *ctx, int use_graph);
/* Per-launch HIP event timing. While enabled, every kernel launch of the context is
 * bracketed by events on its stream; pgs_profile_collect synchronizes and writes the summed
 * milliseconds and launch counts per kernel kind (0 = VRAM copy, 1 = snapshot, 2 = triangle
 * setup, 3 = binning, 4 = shading, 5 = CLUT write, 6 = texture decode, 7 = heuristic,
 * 8 = super-sampled shading, 9 = sampler feedback, 12 = fused launch of a replay, 13 = a replay's
 * VRAM restore, timed in replays without a graph). */
int pgs_profile_enable(pgs_context *ctx, int enable);
int pgs_profile_collect(pgs_context *ctx, double *ms, uint64_t *launches, uint32_t kinds);
/* Blocks until all submitted work has finished. */
int pgs_sync(pgs_context *ctx);
/* Returns the HIP stream (hipStream_t) the context submits on (its main stream: a replay ends
 * with the texture stream's work joined to it). */
void *pgs_stream(pgs_context *ctx);

/* Host execution of the swizzle the kernels use: byte address (nibble address for PSMT4) of
 * pixel (x, y). Same mapping as swizzle_PS2 (gs/shaders/swizzle_utils.h:329-522). */
uint32_t pgs_host_swizzle_addr(uint32_t x, uint32_t y, uint32_t base, uint32_t stride, uint32_t psm);

/* Host execution of the in-block inverse layout the copy kernel uses: pixel offset (dx, dy)
 * inside a 256-byte block of element k of dword `lane`. */
void pgs_host_block_lane_pixel(uint32_t psm, uint32_t lane, uint32_t k, uint32_t *dx, uint32_t *dy);

/* ---- Batched replay of many contexts (no counterpart in the reference, which renders one
 * context per device queue; for callers that run many independent GS command streams).
 * Every context must have finished a recording (pgs_begin_recording / gif_transfer / ... /
 * pgs_end_recording) and must not record again while the batch exists. A replay restores each
 * context's VRAM as it was when its recording began, then runs all recorded work with
 * launches of the same kernel merged across contexts (each context keeps its own order). */
typedef struct pgs_batch pgs_batch;

typedef struct pgs_batch_tuning
{
	uint32_t tile_log2_1x;    /* coarse tile of single-sampled passes: 0 = as recorded, else 3..6 */
	uint32_t waves_per_group; /* waves per shading workgroup: 1, 2 or 4 */
} pgs_batch_tuning;

typedef struct pgs_batch_tuning_result
{
	pgs_batch_tuning tuning;
	float ms_per_replay;
} pgs_batch_tuning_result;

/* flags: PGS_BATCH_SEQUENTIAL issues every context's launches one context after another on
 * the batch stream without merging (the single-stream baseline the merged batch is measured
 * against). */
#define PGS_BATCH_SEQUENTIAL 1u
pgs_batch *pgs_batch_create(pgs_context *const *contexts, uint32_t count, uint32_t flags);
void pgs_batch_destroy(pgs_batch *batch);
int pgs_batch_apply_tuning(pgs_batch *batch, const pgs_batch_tuning *tuning);
/* Times every candidate tuning over `reps` graph replays, applies the fastest and writes up to
 * `max_results` timings (returns the number of candidates timed, or a negative error). */
int pgs_batch_autotune(pgs_batch *batch, uint32_t reps, pgs_batch_tuning_result *results, uint32_t max_results);
int pgs_batch_replay(pgs_batch *batch, int use_graph);
int pgs_batch_sync(pgs_batch *batch);
void *pgs_batch_stream(pgs_batch *batch);
uint32_t pgs_batch_num_launches(pgs_batch *batch);
/* One replay with HIP events around every merged launch: milliseconds and launch counts per
 * kernel kind (same numbering as pgs_profile_collect). */
int pgs_batch_profile(pgs_batch *batch, double *ms, uint64_t *launches, uint32_t kinds);

/* Shading schedule of a batch: PGS_SCHEDULE_HARDWARE (one multi-launch per merged shading
 * step), PGS_SCHEDULE_QUEUE_STATIC (a work queue of resident workgroups in item order),
 * PGS_SCHEDULE_QUEUE_LEARNED (the queue ordered by a cost model's predictions, largest first,
 * heavy tiles split), PGS_SCHEDULE_DISPATCH_STATIC (one workgroup per chunk of a tile's waves,
 * started by the hardware dispatcher, each XCD taking an eighth of the tiles in tile order) or
 * PGS_SCHEDULE_DISPATCH_LEARNED (the same with the tiles placed on XCDs by predicted cost:
 * heavy tiles first, largest first, to the least loaded XCD; the rest in tile order filling each
 * XCD to the mean). The order is built once, by the first replay after this call (or after a
 * tuning change or a cost collection), which runs eagerly; later replays only shade. Learned
 * schedules need `model`, which must outlive the batch; set the schedule again after changing
 * the model. Pixels are identical under every schedule. */
#define PGS_SCHEDULE_HARDWARE 0u
#define PGS_SCHEDULE_QUEUE_STATIC 1u
#define PGS_SCHEDULE_QUEUE_LEARNED 2u
#define PGS_SCHEDULE_DISPATCH_STATIC 3u
#define PGS_SCHEDULE_DISPATCH_LEARNED 4u
typedef struct pgs_cost_model pgs_cost_model;
int pgs_batch_set_schedule(pgs_batch *batch, uint32_t schedule, const pgs_cost_model *model);
/* What the current schedule's order holds, after its planning replay (the first replay after
 * pgs_batch_set_schedule): shading launches, tiles with work, entries (queue items or dispatch
 * workgroups), tiles the learned queue split beyond their base items (heavy tiles), and tiles
 * the learned dispatch order starts first on their XCD. Returns PGS_ERROR_INVALID_ARGUMENT under
 * the hardware schedule or before the planning replay. */
typedef struct pgs_schedule_stats
{
	uint32_t shading_launches;
	uint32_t tiles;
	uint32_t entries;
	uint32_t split_tiles;
	uint32_t heavy_first;
} pgs_schedule_stats;
int pgs_batch_schedule_stats(pgs_batch *batch, pgs_schedule_stats *stats);

/* One replay with the static work queue measuring every shading workgroup: writes up to
 * `max_items` feature rows (PGS_MLP_INPUTS floats) and log2 shader-clock costs; returns the
 * number of items written (or a negative error). */
int pgs_batch_collect_costs(pgs_batch *batch, float *features, float *log2_costs, uint32_t max_items);

/* ---- Learned tile-cost model (no counterpart in the reference): the 24 -> 64 -> 64 -> 1 BF16
 * MLP of pgs_kernels.h, trained on the GPU, that orders a batch's persistent shading queue.
 * Parameters are PGS_MLP_PARAMS floats in the layout of pgs_kernels.h. All pointers are host
 * pointers. */
pgs_cost_model *pgs_cost_model_create(uint64_t seed);
void pgs_cost_model_destroy(pgs_cost_model *model);
int pgs_cost_model_set_params(pgs_cost_model *model, const float *params);
int pgs_cost_model_get_params(pgs_cost_model *model, float *params);
int pgs_cost_model_predict(pgs_cost_model *model, const float *features, uint32_t count, float *out);
int pgs_cost_model_gradients(pgs_cost_model *model, const float *features, const float *targets, uint32_t count,
                             float *grads);
int pgs_cost_model_train(pgs_cost_model *model, const float *features, const float *targets, uint32_t count,
                         uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *mse);

#ifdef __cplusplus
}
#endif

#endif
