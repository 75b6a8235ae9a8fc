// Multi-launch support: one kernel launch runs the workgroups of several launches of the same
// kernel, typically one per GS context of a batch. Launch l owns the launch's workgroups
// [prefix[l], prefix[l + 1]) and its body sees them as 0 .. count - 1, exactly as in a launch
// of its own.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace pgs
{
__device__ __forceinline__ uint32_t find_launch(const uint32_t *prefix, uint32_t n, uint32_t wg)
{
	uint32_t lo = 0, hi = n; // prefix[lo] <= wg < prefix[hi]
	while (hi - lo > 1u)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if (prefix[mid] <= wg)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}
}

// Register budget for 8 waves per SIMD (64 VGPRs). Without it the shading multi-launch
// kernels take 80 VGPRs (6 waves per SIMD); at 64 they do not spill, and the extra waves hide
// more of the memory latency of the per-primitive loads.
#define PGS_OCCUPANCY_8 __attribute__((amdgpu_waves_per_eu(8)))

// Defines `NAME(const ARGS *args, const uint32_t *prefix, uint32_t n)` running BODY for the
// launch that owns each workgroup; the _ATTR form adds a kernel attribute (PGS_OCCUPANCY_8).
#define PGS_DEFINE_MULTI_ATTR(THREADS, ATTR, NAME, BODY, ARGS)                                      \
	__global__ void __launch_bounds__(THREADS) ATTR NAME(const ARGS *args, const uint32_t *prefix, uint32_t n) \
	{                                                                                               \
		const uint32_t l = ::pgs::find_launch(prefix, n, blockIdx.x);                               \
		BODY(args[l], blockIdx.x - prefix[l], prefix[l + 1] - prefix[l]);                            \
	}
#define PGS_DEFINE_MULTI(THREADS, NAME, BODY, ARGS) PGS_DEFINE_MULTI_ATTR(THREADS, , NAME, BODY, ARGS)
