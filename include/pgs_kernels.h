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

/* Host->local and local->local VRAM copies of one transfer batch.
 * Replaces the PREPARE + EXECUTE dispatches of vram_copy.comp issued by
 * GSRenderer::flush_transfer / emit_copy_vram (gs/gs_renderer.cpp:3866-4047, 2721-2827).
 * `blocks` lists every destination 256-byte block of the batch with the ordered slice of
 * `records` (copy index + block origin) that writes it; `descs` holds the copies. */
typedef struct pgs_vram_copy_args
{
	uint8_t *vram;             /* plane 0, 4 MiB */
	uint8_t *ref_plane;        /* super-sample reference plane, or NULL at 1x */
	const uint8_t *shadow;     /* batch-start snapshot for hazardous local copies */
	const uint8_t *payload;    /* host->local payload arena */
	const void *descs;         /* CopyDesc[] */
	const void *records;       /* BlockRecord[] */
	const void *blocks;        /* DestBlock[] */
	uint32_t num_blocks;
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

#ifdef __cplusplus
}
#endif

#endif
