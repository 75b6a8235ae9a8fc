// VRAM transfer kernels for gfx950 (host->local uploads and local->local copies).
//
// Replaces gs/shaders/vram_copy.comp. The reference resolves write-after-write races between
// copies of one batch with an atomic linked list per dword and a second "execute" pass
// (vram_copy.comp:115-197). Here the host groups each dependency level of the batch by
// destination block instead: a wave64 owns one 256-byte VRAM block, lane L owns dword L of it,
// and the wave applies every copy that touches the block in submission order while the dword
// stays in a VGPR. A block is read once and written once, with fully coalesced 256-byte
// accesses, and no atomics are needed. Per-pixel read/merge rules follow vram_copy.comp:199-613.
// A level is latency-bound (a chain of dependent loads per wave): the first record of a block is
// inline in its work item, and a dword's pixels are gathered with all loads in flight at once.
#include "gs_defs.hpp"
#include "pgs_kernels.h"
#include "multi_launch.hpp"
#include "multi_dispatch.hpp"

#include "transfer_bodies.hpp"

namespace pgs
{
namespace
{
using xfer::resolve_blocks_body;
using xfer::snapshot_blocks_body;

// One workgroup per range: 16-byte vector copies. Source (the snapshot) and destination (the
// VRAM planes and CLUT) are separate allocations, so every thread issues four loads before its
// four stores instead of waiting out each load's latency.
__global__ void __launch_bounds__(256) k_restore(pgs_restore_args a)
{
	const pgs_copy_range r = a.ranges[blockIdx.x];
	const uint4 *__restrict__ src = reinterpret_cast<const uint4 *>(r.src);
	uint4 *__restrict__ dst = reinterpret_cast<uint4 *>(r.dst);
	const uint32_t n = r.bytes / 16u;
	uint32_t i = threadIdx.x;
	for (; i + 768u < n; i += 1024u)
	{
		const uint4 v0 = src[i], v1 = src[i + 256u], v2 = src[i + 512u], v3 = src[i + 768u];
		dst[i] = v0;
		dst[i + 256u] = v1;
		dst[i + 512u] = v2;
		dst[i + 768u] = v3;
	}
	for (; i < n; i += 256u)
		dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_resolve_blocks(pgs_vram_copy_args a) { resolve_blocks_body(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) k_snapshot_blocks(pgs_vram_snapshot_args a) { snapshot_blocks_body(a, blockIdx.x, gridDim.x); }
PGS_DEFINE_MULTI(256, k_resolve_blocks_multi, resolve_blocks_body, pgs_vram_copy_args)
PGS_DEFINE_MULTI(256, k_snapshot_blocks_multi, snapshot_blocks_body, pgs_vram_snapshot_args)
}
}

uint32_t pgs_grid_vram_copy(const pgs_vram_copy_args &a) { return (a.num_items + 3u) / 4u; }
uint32_t pgs_grid_vram_snapshot(const pgs_vram_snapshot_args &a) { return (a.num_blocks + 3u) / 4u; }

int pgs_multi_vram_copy(const pgs_vram_copy_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                        hipStream_t stream)
{
	hipLaunchKernelGGL(pgs::k_resolve_blocks_multi, dim3(total), dim3(256), 0, stream, args, prefix, n);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

int pgs_multi_vram_snapshot(const pgs_vram_snapshot_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                            hipStream_t stream)
{
	hipLaunchKernelGGL(pgs::k_snapshot_blocks_multi, dim3(total), dim3(256), 0, stream, args, prefix, n);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

extern "C" int pgs_k_vram_copy(const pgs_vram_copy_args *args, hipStream_t stream)
{
	if (!args || !args->vram || (args->num_items && (!args->items || !args->descs)))
		return -int(hipErrorInvalidValue);
	if (args->num_items == 0)
		return 0;
	const uint32_t groups = pgs_grid_vram_copy(*args); // four waves of one item each
	hipLaunchKernelGGL(pgs::k_resolve_blocks, dim3(groups), dim3(256), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

extern "C" int pgs_k_restore(const pgs_restore_args *args, hipStream_t stream)
{
	if (!args || (args->num_ranges && !args->ranges))
		return -int(hipErrorInvalidValue);
	if (args->num_ranges == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_restore, dim3(args->num_ranges), dim3(256), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}

extern "C" int pgs_k_vram_snapshot(const pgs_vram_snapshot_args *args, hipStream_t stream)
{
	if (!args || !args->vram || !args->shadow || !args->block_list)
		return -int(hipErrorInvalidValue);
	if (args->num_blocks == 0)
		return 0;
	const uint32_t groups = (args->num_blocks + 3u) / 4u;
	hipLaunchKernelGGL(pgs::k_snapshot_blocks, dim3(groups), dim3(256), 0, stream, *args);
	hipError_t err = hipGetLastError();
	return err == hipSuccess ? 0 : -int(err);
}
