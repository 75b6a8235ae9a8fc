/* Kernel-level C ABI of the MI355X GS backend.
 *
 * One launcher per dispatch site of the reference renderer. Every launcher takes device
 * pointers, a POD argument block and a hipStream_t, never throws, and returns 0 on success
 * or a negative hipError_t. The renderer (parallel-gs_amd/csrc/renderer.cpp) is the only
 * caller inside the library; the declarations are public so that a host integration can
 * drive the kernels directly.
 */
#ifndef PGS_KERNELS_H
#define PGS_KERNELS_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Host->local and local->local VRAM copies of one dependency level of a transfer batch.
 * Replaces the PREPARE + EXECUTE dispatches of vram_copy.comp issued by
 * GSRenderer::flush_transfer / emit_copy_vram (gs/gs_renderer.cpp:3866-4047, 2721-2827).
 * `items` lists every destination 256-byte block of the level (CopyItem: the first record
 * inline, further records in `records`, copy index + block origin) and the blocks the next
 * level's snapshot copies from VRAM; `descs` holds the copies. Each wave takes two items. */
typedef struct pgs_vram_copy_args
{
	uint8_t *vram;             /* plane 0, 4 MiB */
	uint8_t *ref_plane;        /* super-sample reference plane, or NULL at 1x */
	const uint8_t *shadow;     /* level-start snapshot for hazardous local copies */
	uint8_t *shadow_next;      /* the next level's snapshot, or NULL */
	const uint8_t *payload;    /* host->local payload arena */
	const void *descs;         /* CopyDesc[] */
	const void *records;       /* BlockRecord[] */
	const void *items;         /* CopyItem[num_items] */
	uint32_t num_items;
	uint32_t pad;
} pgs_vram_copy_args;

int pgs_k_vram_copy(const pgs_vram_copy_args *args, hipStream_t stream);

/* Copies listed 256-byte blocks of VRAM into the shadow plane.
 * Replaces the shadow-page buffer copies of GSRenderer::flush_transfer
 * (gs/gs_renderer.cpp:3921-3942). */
typedef struct pgs_vram_snapshot_args
{
	const uint8_t *vram;
	uint8_t *shadow;
	const uint32_t *block_list;
	uint32_t num_blocks;
	uint32_t pad;
} pgs_vram_snapshot_args;

int pgs_k_vram_snapshot(const pgs_vram_snapshot_args *args, hipStream_t stream);

/* Copies a list of byte ranges (one workgroup per range; sizes a multiple of 16 bytes at
 * 16-byte aligned addresses). Used to restore the VRAM pages and the CLUT a
 * recording changes before each replay, for one context or for every context of a batch in
 * one launch. No counterpart in the reference (backend extension for replays). */
typedef struct pgs_copy_range
{
	const uint8_t *src;
	uint8_t *dst;
	uint32_t bytes;
	uint32_t pad;
} pgs_copy_range;

typedef struct pgs_restore_args
{
	const pgs_copy_range *ranges; /* device array */
	uint32_t num_ranges;
	uint32_t pad;
} pgs_restore_args;

int pgs_k_restore(const pgs_restore_args *args, hipStream_t stream);

/* Uploads the reciprocal lookup tables the raster kernels read (the 256-entry fixed-point
 * seed table and the 256-entry float seed LUT). Replaces GSRenderer::init_luts
 * (gs/gs_renderer.cpp:3600-3660). Must run once per device before any setup launch. */
int pgs_k_init_tables(const uint8_t *rcp_seed, const float *rcp_float_lut);

/* Test hook (no counterpart in the reference): evaluates the device reciprocals triangle setup
 * uses — rcp_fixed (mode 0, the fixed-point Newton reciprocal of math_utils.h:27-61) or rcp_float
 * (mode 1, inputs are float bits; math_utils.h:65-101) — over `count` consecutive inputs from
 * `first` (at most 2^32) and adds the wrapping 64-bit sum over inputs of
 * ((output bits << 32) | input) * 0x9E3779B97F4A7C15 to *hash (device memory, cleared by the
 * caller). Needs pgs_k_init_tables. Lets a test compare the device arithmetic with a CPU sweep. */
int pgs_k_rcp_eval(uint32_t mode, uint32_t first, uint64_t count, uint64_t *hash, hipStream_t stream);

/* Test hook: the shading kernel's Q-derived mipmap level of detail, ceil(K - 2^(L+4) log2|q|)
 * in 1/16 units (gs/shaders/ubershader.comp:1076-1087, evaluated exactly), for L = l (0..3) and
 * K = k (-2048..2047) over `count` consecutive float bit patterns from `first`; adds the wrapping
 * 64-bit sum of ((lod bits << 32) | input) * 0x9E3779B97F4A7C15 to *hash (device memory, cleared
 * by the caller). */
int pgs_k_lod_eval(uint32_t l, int32_t k, uint32_t first, uint64_t count, uint64_t *hash, hipStream_t stream);

/* Triangle / sprite / line / point setup: one thread per primitive.
 * Replaces triangle_setup.comp dispatched by GSRenderer::dispatch_triangle_setup
 * (gs/gs_renderer.cpp:1898-1977). Writes PrimitiveSetup (80 B) and TransformedAttributes
 * (64 B) per primitive and updates the state / tex words of the primitive records in place. */
typedef struct pgs_setup_args
{
	const void *positions;  /* VertexPosition[3 * num_primitives] */
	const void *attributes; /* VertexAttribute[3 * num_primitives] */
	void *prims;            /* PrimitiveAttribute[num_primitives], updated in place */
	void *setup;            /* PrimitiveSetup[num_primitives] */
	void *tattr;            /* TransformedAttributes[num_primitives] */
	uint32_t *state_copy;   /* super-sampled passes: final state word of every primitive (or NULL) */
	uint32_t *flat_buckets; /* super-sampled passes: 256-bit mask of depth buckets that hold flat
	                           primitives (triangle_setup.comp:820-872); the heuristic clears it */
	uint32_t num_primitives;
	uint32_t rate_log2;      /* log2 of the super-sampling grid per axis (0 at 1x) */
	uint32_t opaque_fbmask;  /* FBMSK every primitive of the pass shares */
	uint32_t z_bucket_shift; /* depth >> shift = bucket: 8, 16, 24 for 16, 24, 32-bit depth */
	/* Optional first level of binning (NULL: none). coarse_mask[ct * mask_words + w] bit b is set
	 * when primitive 64 w + b has a bounding box touching coarse tile ct, a square of
	 * 2^coarse_log2 pixels of the grid at (coarse_origin_x, coarse_origin_y), row-major
	 * coarse_x by coarse_y (each at most 256); mask_words = ceil(num_primitives / 64). Has no
	 * counterpart in the reference, whose binning tests every primitive against every tile. */
	uint64_t *coarse_mask;
	int32_t coarse_origin_x, coarse_origin_y;
	uint32_t coarse_log2;
	uint32_t coarse_x, coarse_y;
	uint32_t mask_words;
	/* Optional: words zeroed by the launch (the sampler-feedback masks of the pass). */
	uint32_t *clear_words;
	uint32_t clear_count;
	/* Super-sampled textures (triangle_setup.comp:677-789): bit 0 set when a texture of the pass
	 * is super-sampled (SUPER_SAMPLED_TEXTURES), bit 1 when the pass is a deduced blur
	 * (gs/gs_renderer.cpp:1848-1885), bits 8..9 the x rate log2 of the sampling grid. */
	uint32_t texture_flags;
} pgs_setup_args;

int pgs_k_triangle_setup(const pgs_setup_args *args, hipStream_t stream);

/* Single-sample heuristic of super-sampled passes: flat primitives in a depth bucket with no
 * 3D geometry sharing their vertices are snapped to the pixel grid; flat Z-tested primitives
 * inside the 3D depth range lose attribute snapping.
 * Replaces single_sample_heuristic.comp dispatched by
 * GSRenderer::dispatch_single_sample_heuristic (gs/gs_renderer.cpp:1979-2040). One 512-thread
 * workgroup per possible bucket; buckets not in `flat_buckets` exit at once. Reads the states
 * from `state_copy` (as setup left them) and writes updated states into `prims`. */
typedef struct pgs_ssaa_heuristic_args
{
	const void *positions;   /* VertexPosition[3 * num_primitives] */
	void *prims;             /* PrimitiveAttribute[num_primitives] */
	const uint32_t *state_copy;
	uint32_t *flat_buckets;
	uint32_t num_primitives;
	uint32_t z_shift_bucket; /* gs/gs_renderer.cpp:2008-2031 */
	uint32_t z_shift;
	uint32_t z_max;
} pgs_ssaa_heuristic_args;

int pgs_k_ssaa_heuristic(const pgs_ssaa_heuristic_args *args, hipStream_t stream);

/* Coarse binning: for each coarse tile, the ascending list of primitives whose conservative
 * edge test overlaps the tile. Replaces binning.comp dispatched by GSRenderer::dispatch_binning
 * (gs/gs_renderer.cpp:2068-2145). One 256-thread workgroup per coarse tile. */
typedef struct pgs_bin_args
{
	const void *setup;
	const void *prims;
	uint16_t *lists;        /* tiles * list_stride entries */
	uint32_t *counts;       /* per tile */
	uint32_t num_primitives;
	uint32_t list_stride;
	int32_t origin_x, origin_y;
	uint32_t tile_log2;
	uint32_t tiles_x, tiles_y;
	uint32_t super_sampled;  /* pass is super-sampled: decide per tile (binning.comp:87-125); bit 1:
	                            the pass samples super-sampled textures (their sprites' tiles
	                            are super-sampled, binning.comp:123-124) */
	uint32_t *tile_modes;    /* per tile: 0 = no work, 1 = single-sampled, 2 = super-sampled */
	/* Optional coarse mask from triangle setup (same grid origin; a coarse tile is
	 * 2^coarse_shift by 2^coarse_shift tiles): only primitives with their bit set are tested.
	 * NULL: every primitive is tested. The lists are the same either way. */
	const uint64_t *coarse_mask;
	uint32_t coarse_shift;
	uint32_t coarse_x, coarse_y;
	uint32_t mask_words;
} pgs_bin_args;

int pgs_k_binning(const pgs_bin_args *args, hipStream_t stream);

/* Shading: primitives of each block's coarse tile are applied to every sample in submission
 * order (coverage, Z, texture, combiner, fog, alpha test, blend, dither, FBMSK) and the results
 * are written back once per tile. Single-sampled tiles run one wave64 per 8x8 pixel block;
 * super-sampled tiles run one wave64 per 64 samples (the samples of 64 / S pixels) and resolve
 * in the wave. In a super-sampled context `vram` holds 2 + S planes of 4 MiB (plane 1 marks
 * pixels with valid samples) and `tile_modes` selects which tiles each launch shades.
 * Replaces ubershader.comp dispatched by GSRenderer::dispatch_shading
 * (gs/gs_renderer.cpp:2489-2682, 2322-2416). */
typedef struct pgs_shade_args
{
	uint8_t *vram;
	const void *setup;
	const void *tattr;
	const void *prims;
	const void *states;      /* StateVector[] */
	const uint16_t *lists;
	const uint32_t *counts;
	const void *textures;    /* TextureDesc[] */
	const uint32_t *tex_images;
	const uint16_t *clut;    /* the CLUT (512 entries), read by on-tile palette feedback */
	uint32_t list_stride;
	int32_t origin_x, origin_y;
	uint32_t tile_log2;
	uint32_t tiles_x, tiles_y;
	uint32_t blocks_x, blocks_y;
	uint32_t fb_psm, fbp, fbw;
	uint32_t z_psm;          /* 0xffffffff: the pass does not read or write depth */
	uint32_t zbp;
	uint32_t rate_log2;      /* 0: single-sampled shading; else the pass's y rate (x: rate_x_log2) */
	uint32_t rate_x_log2;
	const uint32_t *tile_modes; /* NULL: every tile with primitives; else shade tiles whose mode
	                               is 1 (single-sampled launch) or 2 (super-sampled launch) */
	uint32_t has_ref_plane;  /* context is super-sampled: dirty pixels update plane 1 */
	uint32_t waves_per_group; /* waves per workgroup: 1, 2 or 4 (0 means 4) */
	uint32_t single_tiles;   /* super-sampled launch only: 1 also shades the tiles whose mode is 1,
	                            single-sampled, with the first of each block's sample waves (the
	                            pass then needs no single-sampled launch); 0 leaves them */
	/* Passes whose depth buffer aliases the colour buffer (FBP == ZBP) shade in several
	 * launches over primitive ranges (gs/gs_renderer.cpp:2525-2690, the lo/hi primitive index
	 * of ubershader.comp's push constants): */
	const uint8_t *src;      /* NULL: colour and depth are read from vram; else from this copy of
	                            VRAM with the same planes (a single-stepped primitive reads the
	                            VRAM as it was before the primitive) */
	const uint8_t *zsrc;     /* NULL: depth is read where colour is; else from this copy (the
	                            read-only depth cache, gs/gs_renderer.cpp:2184-2232) */
	uint32_t prim_first;     /* shade primitives prim_first .. prim_first + prim_count - 1 of the
	                            tile lists; prim_count 0: every primitive from prim_first */
	uint32_t prim_count;
	uint32_t skip_stores;    /* bit 0: leave colour unstored, bit 1: leave depth unstored */
	uint32_t pad_alias;
	/* Super-sampled textures (ubershader.comp:1155-1212): the phase filter table of the context's
	 * sampling grid (PGS_PHASE_LUT_WORDS words, pgs_phase_lut), read by primitives whose texture
	 * word has TEX_PER_SAMPLE without TEX_SAMPLE_MAPPING / TEX_SAMPLE_RESOLVED; NULL when no
	 * texture of the pass is super-sampled. */
	const uint32_t *phase_lut;
} pgs_shade_args;

int pgs_k_shade(const pgs_shade_args *args, hipStream_t stream);

/* Phase filter table of super-sampled textures for a sampling grid of 2^rate_x x 2^rate_y
 * samples (gs/gs_renderer.cpp:94-313, init_phase_lut): PGS_PHASE_LUT_WORDS words, two per entry,
 * 64 x 64 entries for LINEAR magnification, then 32 x 32 for NEAREST. Entry word 0: sample ids
 * of the four filter taps in bits 0..15, their signed 2-bit texel offsets (x, y) in bits
 * 16..31; word 1: the four weights in unorm8, summing to 255. Host function. */
#define PGS_PHASE_LUT_WORDS (2 * (64 * 64 + 32 * 32))
void pgs_phase_lut(uint32_t rate_x, uint32_t rate_y, uint32_t *out);

/* Work-queue shading of a multi-launch by coarse tile (batched renderer; no counterpart in
 * the reference): a grid of resident workgroups pops tiles (tile numbering: launch l owns
 * tiles [prefix[l], prefix[l + 1])) from `order` with the atomic counter[0] (0 on entry) until
 * counter[1] tiles are taken, and shades every block of each tile. All launches share rate_x /
 * rate_y / waves_per_group. With `costs` set, each tile's shader-clock cycles are stored at
 * costs[tile]. The order only decides which tile runs when: every tile is shaded exactly once
 * by one workgroup, so pixels do not depend on it. */
typedef struct pgs_queue_args
{
	const void *args;          /* device pgs_shade_args[num_launches] */
	const uint32_t *prefix;    /* device tile prefix, num_launches + 1 entries */
	const uint32_t *order;     /* device queue (pgs_k_queue_order): entry = tile << 8 | part << 4 |
	                              (parts - 1); the item covers part `part` of `parts` equal ranges
	                              of the tile's waves */
	uint32_t *counter;         /* device: [0] pop counter (0 on entry), [1] entries in `order` */
	uint32_t *costs;           /* device costs[total] (shader clocks added per item; cleared by the
	                              caller), or NULL */
	uint32_t num_launches;
	uint32_t total;            /* tiles over all launches */
	uint32_t rate_x, rate_y;
	uint32_t waves_per_group;
	uint32_t pad;
} pgs_queue_args;

int pgs_k_shade_queue(const pgs_queue_args *args, hipStream_t stream);

/* Ordered dispatch: the same shading with the hardware dispatcher as the queue. Workgroup b
 * (of `grid`, at least counter[1]) shades chunk order[b] & 0xfff of tile order[b] >> 12, the
 * chunk being waves_per_group consecutive waves of the tile; workgroups from counter[1] on, and
 * those whose entry is 0xffffffff, exit at once. The dispatcher starts workgroups in index order
 * and hands workgroup b to XCD b % 8, so `order` is the start order of every XCD. */
int pgs_k_shade_ordered(const pgs_queue_args *args, uint32_t grid, hipStream_t stream);

/* Features of every coarse tile of a shading multi-launch for the tile-cost model: one wave
 * per tile scans the tile's list. Writes PGS_MLP_INPUTS floats per tile (all zero but the
 * last, constant input, for tiles without work for this launch). */
typedef struct pgs_features_args
{
	const void *args;          /* device pgs_shade_args[num_launches] */
	const uint32_t *prefix;
	float *features;           /* total x PGS_MLP_INPUTS */
	uint32_t *tile_waves;      /* optional: per tile, the shading waves it needs (0 without work) */
	uint32_t num_launches;
	uint32_t total;
} pgs_features_args;

int pgs_k_tile_features(const pgs_features_args *args, hipStream_t stream);

/* Builds the queue of tiles with work (feature 0 nonzero): in tile order (`predicted` NULL)
 * or by predicted cost, largest first (256 buckets of 1/8 in log2 cycles; order within a bucket
 * is arbitrary). A work-queue tile becomes items of at most 16 waves (cost order may split a
 * heavy tile further, up to 16 items); a dispatch-order tile becomes one entry per chunk.
 * Writes the entries to `order`, their number to counter[1] and zeroes counter[0] for the
 * launch that follows. One workgroup places the tiles, a second kernel writes the entries. */
typedef struct pgs_order_args
{
	const float *features;     /* count x PGS_MLP_INPUTS */
	const float *predicted;    /* log2 cycles per tile, or NULL for tile order */
	uint32_t *order;           /* capacity entries */
	uint32_t *counter;         /* [0], [1] as in pgs_queue_args */
	const uint32_t *tile_waves;/* the waves of every tile (pgs_k_tile_features) */
	uint32_t *run_start;       /* scratch, count entries: first entry of each tile */
	uint32_t *run_weight;      /* scratch, count entries: entries of each tile */
	uint32_t count;            /* tiles, below 2^24 (2^20 for dispatch order) */
	uint32_t slots;            /* resident workgroups of the queue launch */
	uint32_t capacity;         /* entries `order` holds: queue, at least count + 2 * slots + 1024
	                              plus the waves of all tiles / 16; dispatch, the waves of all
	                              tiles */
	uint32_t group_waves;      /* dispatch order: waves per workgroup */
	uint32_t dispatch;         /* 1: entries tile << 12 | chunk for pgs_k_shade_ordered, one per
	                              chunk of group_waves waves; 0: work-queue items */
	uint32_t pad;
} pgs_order_args;

int pgs_k_queue_order(const pgs_order_args *args, hipStream_t stream);

/* Writes `value` to `count` words (graph-safe replacement of a memset). */
int pgs_k_fill_u32(uint32_t *dst, uint32_t count, uint32_t value, hipStream_t stream);

/* CLUT load of one TEX0/TEX2 write with CLD that loads.
 * Replaces the clut_write.comp dispatch of GSRenderer::flush_palette_upload
 * (gs/gs_renderer.cpp:3662-3819) for a load queued by GSInterface::handle_clut_upload
 * (gs/gs_interface.cpp:550-797). `clut` is the 512-entry 16-bit CLUT. */
typedef struct pgs_clut_args
{
	const uint8_t *vram;
	uint16_t *clut;
	uint32_t cbp;      /* CLUT base, 256-byte blocks */
	uint32_t cbw;      /* TEXCLUT.CBW (CSM2 only) */
	uint32_t cou, cov; /* TEXCLUT.COU / COV (CSM2 only) */
	uint32_t cpsm;     /* PSMCT32 (PSMCT24 passes as PSMCT32), PSMCT16 or PSMCT16S */
	uint32_t csm;      /* 0 = CSM1, 1 = CSM2 */
	uint32_t csa;      /* bank offset, 0..31 (0..15 for PSMCT32) */
	uint32_t index4;   /* 1 for 4-bit index formats (16 entries) */
} pgs_clut_args;

int pgs_k_clut_write(const pgs_clut_args *args, hipStream_t stream);

/* Decodes the textures of one render pass into RGBA8 images.
 * Replaces the upload.comp dispatches of GSRenderer::upload_texture
 * (gs/gs_renderer.cpp:3458-3598). `decodes` is a device array of TexDecode records
 * (parallel-gs_amd/csrc/texture_layout.hpp), ordered by first workgroup. */
typedef struct pgs_tex_decode_args
{
	const uint8_t *vram;
	const uint16_t *clut;
	uint32_t *images;
	const void *decodes;
	const uint32_t *feedback; /* optional sampler-feedback masks (pgs_k_sampler_feedback): only texels
	                             of marked 8x8 blocks are decoded; NULL decodes every texel */
	uint32_t num_textures;
	uint32_t num_groups;
} pgs_tex_decode_args;

int pgs_k_texture_decode(const pgs_tex_decode_args *args, hipStream_t stream);

/* Sampler feedback: marks, for every texture of a pass, the 8x8 texel blocks of its decoded image
 * that a primitive can sample, so that the decode can skip the rest. The texel box of a primitive
 * comes from its texture coordinates at the corners of its pixel box grown by one pixel (every
 * position the shader evaluates attributes at, snapped ones included), computed with the shader's
 * own edge functions and barycentrics, then taken through the shader's clamp, 1/16 snap, bilinear
 * footprint and wrap. One thread per primitive, atomicOr into masks that must be zero on entry (triangle setup
 * clears them: pgs_setup_args::clear_words). Mask of texture i: PGS_FEEDBACK_WORDS words at
 * i * PGS_FEEDBACK_WORDS, 4 words per row of blocks; images above 1024 texels in either
 * direction are decoded whole. Replaces the sampler_feedback.comp dispatch of
 * GSRenderer::dispatch_texture_analysis (gs/gs_renderer.cpp:3421-3456). */
#define PGS_FEEDBACK_WORDS 512
typedef struct pgs_feedback_args
{
	const void *setup;      /* PrimitiveSetup[num_primitives] */
	const void *prims;      /* PrimitiveAttribute[num_primitives] */
	const void *tattr;      /* TransformedAttributes[num_primitives] (after setup) */
	const void *states;     /* StateVector[] */
	const void *textures;   /* TextureDesc[num_textures] */
	uint32_t *masks;        /* num_textures * PGS_FEEDBACK_WORDS */
	uint32_t num_primitives;
	uint32_t num_textures;
} pgs_feedback_args;

int pgs_k_sampler_feedback(const pgs_feedback_args *args, hipStream_t stream);

/* ---- Tile-cost MLP of the learned tile-queue order (no counterpart in the reference; used
 * by the batched renderer to order its persistent shading queue, never to compute a pixel).
 * 24 features -> 64 -> 64 -> 1, ReLU, BF16 weights and layer inputs, FP32 accumulation.
 * Parameter vector layout: W1[64][24], b1[64], W2[64][64], b2[64], w3[64], b3. */
#define PGS_MLP_INPUTS 24
#define PGS_MLP_HIDDEN 64
#define PGS_MLP_PARAMS 5825
#define PGS_MLP_SAMPLE_FLOATS 281 /* per-sample training scratch: x, h1, h2, dz1, dz2, dy */

typedef struct pgs_mlp_forward_args
{
	const uint16_t *weights; /* BF16 parameter vector */
	const float *features;   /* count x 24 */
	float *out;              /* count predictions */
	uint32_t count;
	uint32_t pad;
} pgs_mlp_forward_args;

int pgs_k_mlp_forward(const pgs_mlp_forward_args *args, hipStream_t stream);

/* Gradient of mean((y - target)^2) over `count` samples into `grads` (FP32, parameter layout). */
typedef struct pgs_mlp_grad_args
{
	const uint16_t *weights;
	const float *features;   /* count x 24 */
	const float *targets;    /* count */
	float *scratch;          /* count x PGS_MLP_SAMPLE_FLOATS */
	float *grads;            /* PGS_MLP_PARAMS */
	float *loss;             /* optional: count squared errors */
	uint32_t count;
	uint32_t pad;
} pgs_mlp_grad_args;

int pgs_k_mlp_gradients(const pgs_mlp_grad_args *args, hipStream_t stream);

/* pgs_k_mlp_gradients without a workspace: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, training loops that forbid
 * allocation). `scratch` is ignored; the per-sample training scratch is a fixed device array
 * of the library (PGS_MLP_NOALLOC_MAX_SAMPLES x PGS_MLP_SAMPLE_FLOATS floats, 4.6 MB per GPU,
 * held until the process ends). count above PGS_MLP_NOALLOC_MAX_SAMPLES returns
 * -hipErrorInvalidValue. The call allocates nothing. Every call on one GPU shares that array:
 * calls that overlap on one GPU (other streams, other threads) race. Results are identical to
 * pgs_k_mlp_gradients. */
#define PGS_MLP_NOALLOC_MAX_SAMPLES 4096u
int pgs_k_mlp_gradients_noalloc(const pgs_mlp_grad_args *args, hipStream_t stream);

typedef struct pgs_mlp_adam_args
{
	float *master;           /* FP32 parameters */
	uint16_t *weights;       /* BF16 copy, refreshed */
	const float *grads;
	float *m, *v;            /* Adam moments */
	float lr, beta1, beta2, eps;
	float beta1_t, beta2_t;  /* beta1^t, beta2^t of this step */
} pgs_mlp_adam_args;

int pgs_k_mlp_adam(const pgs_mlp_adam_args *args, hipStream_t stream);
int pgs_k_mlp_quantize(const float *master, uint16_t *weights, hipStream_t stream);

/* Fused launch (no counterpart in the reference, which records these dispatches one by one):
 * one launch of 256-thread workgroups runs several independent parts, each the work of one
 * launch of the kernels above (VRAM copy or snapshot, CLUT write, texture decode, triangle
 * setup, single-sample heuristic, binning). Part p owns workgroups [begin, begin + groups);
 * parts must not depend on each other (no part reads what another writes), and each part's
 * arguments must satisfy its own launcher's checks. Workgroups per part: copy and snapshot
 * (num_blocks + 3) / 4, CLUT 1, decode num_groups, setup (num_primitives + 255) / 256,
 * heuristic 256, binning tiles_x * tiles_y. Replays use it to run dispatches whose order
 * does not matter as one launch. */
#define PGS_FUSED_VRAM_COPY 0u
#define PGS_FUSED_VRAM_SNAPSHOT 1u
#define PGS_FUSED_CLUT_WRITE 2u
#define PGS_FUSED_TEXTURE_DECODE 3u
#define PGS_FUSED_TRIANGLE_SETUP 4u
#define PGS_FUSED_SSAA_HEURISTIC 5u
#define PGS_FUSED_BINNING 6u
#define PGS_FUSED_MAX_PARTS 8u

typedef struct pgs_fused_part
{
	uint32_t kind;   /* PGS_FUSED_* */
	uint32_t begin;  /* first workgroup of the part */
	uint32_t groups; /* workgroups of the part */
	uint32_t pad;
	union
	{
		pgs_vram_copy_args copy;
		pgs_vram_snapshot_args snapshot;
		pgs_clut_args clut;
		pgs_tex_decode_args decode;
		pgs_setup_args setup;
		pgs_ssaa_heuristic_args heuristic;
		pgs_bin_args bin;
	} args;
} pgs_fused_part;

typedef struct pgs_fused_args
{
	const pgs_fused_part *parts; /* device array, parts in workgroup order */
	uint32_t num_parts;          /* 1 .. PGS_FUSED_MAX_PARTS */
	uint32_t groups;             /* total workgroups (sum over the parts) */
} pgs_fused_args;

int pgs_k_fused(const pgs_fused_args *args, hipStream_t stream);

/* ---- CRTC scanout (vsync) --------------------------------------------------------------
 * Images are RGBA8 (R in the low byte), row pitch = width. Every float step rounds once
 * (division, multiplication, addition in the order written in the kernel source), so the
 * results are defined bit for bit. */

/* Reads one read circuit's DISPFB region into an image.
 * Replaces sample_crtc_circuit + sample_circuit.frag (gs/gs_renderer.cpp:4055-4112,
 * gs/shaders/sample_circuit.frag:40-148). Pixels outside the valid extent are 0. At 4x, 8x
 * and 16x (high-resolution scanout) each 2x2 quad of image pixels reads one single-sampled
 * pixel and averages the quad's share of its super-samples where the reference plane marks
 * them valid. */
typedef struct pgs_scanout_circuit_args
{
	const uint8_t *vram;       /* plane 0 = VRAM, 1 = super-sample reference, 2.. = samples */
	uint32_t *image;
	uint32_t width, height;    /* image extent */
	uint32_t valid_width, valid_height;
	uint32_t psm;              /* DISPFB.PSM: PSMCT32/24/16/16S or PSMZ32/24/16/16S */
	uint32_t base_block;       /* DISPFB.FBP x 32 */
	uint32_t fbw, dbx, dby;
	uint32_t phase, phase_stride;
	uint32_t super_samples;    /* 1, 4, 8 or 16 */
} pgs_scanout_circuit_args;

int pgs_k_scanout_circuit(const pgs_scanout_circuit_args *args, hipStream_t stream);

/* Merges the read circuits over the background colour into one frame.
 * Replaces the merge render pass of GSRenderer::vsync (gs/gs_renderer.cpp:4760-4905) with
 * blit_circuit.frag's linear-clamp sampling: circuit 2 is copied first (SLBG = 0), then
 * circuit 1 is drawn over it replacing (MMOD = 1, ALP = 255), blended with the constant ALP
 * (MMOD = 1) or with its own alpha (MMOD = 0). The filter takes 8-bit sub-texel weights
 * (the weight of texel i0 + 1 is round(256 * frac) / 256). */
typedef struct pgs_scanout_merge_args
{
	uint32_t *merged;
	uint32_t width, height;
	const uint32_t *circuit[2];
	uint32_t circuit_width[2], circuit_height[2];
	int32_t vp_x[2], vp_y[2];
	uint32_t vp_width[2], vp_height[2];
	uint32_t draw;             /* bit 0: circuit 1, bit 1: circuit 2 */
	uint32_t blend;            /* circuit 1: 0 replace, 1 constant ALP, 2 source alpha */
	uint32_t alp;              /* PMODE.ALP */
	uint32_t bgcolor;          /* BGCOLOR: R | G << 8 | B << 16 */
} pgs_scanout_merge_args;

int pgs_k_scanout_merge(const pgs_scanout_merge_args *args, hipStream_t stream);

/* EXTWRITE feedback: writes a scanout image back into VRAM as PSMCT32.
 * Replaces extwrite.comp (gs/gs_renderer.cpp:4921-4999, gs/shaders/extwrite.comp). The source
 * is sampled nearest (texel floor(u * size), clamped); EMODC / EMODA select RGB, luma, YCbCr or
 * alpha conversions. With clear_reference the super-sample reference word of every written
 * pixel is cleared. */
typedef struct pgs_scanout_extwrite_args
{
	uint8_t *vram;
	const uint32_t *view;
	uint32_t view_width, view_height;
	uint32_t width, height;    /* pixels written */
	uint32_t wdx, wdy, exbp, exbw, emoda, emodc;
	uint32_t clear_reference;
	float uv_base[2], uv_scale[2];
} pgs_scanout_extwrite_args;

int pgs_k_scanout_extwrite(const pgs_scanout_extwrite_args *args, hipStream_t stream);

/* FastMAD deinterlace of the last four fields into one frame of twice the field height.
 * Replaces GSRenderer::fastmad_deinterlace + weave.frag (gs/gs_renderer.cpp:5057-5099,
 * gs/shaders/weave.frag). field[0] is the newest field. */
typedef struct pgs_scanout_weave_args
{
	const uint32_t *field[4];
	uint32_t *out;
	uint32_t width, height;    /* output extent: height = 2 x field height */
	uint32_t phase;
} pgs_scanout_weave_args;

int pgs_k_scanout_weave(const pgs_scanout_weave_args *args, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif
