// Scheduling helpers of the batched renderer's work-queue shading (no counterpart in the
// reference): the cost-ordered queue and a graph-safe fill.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "pgs_kernels.h"

namespace pgs
{
namespace
{
constexpr uint32_t kBuckets = 256, kThreads = 1024;

// Counting sort by quantized predicted cost, descending. Positions inside a bucket come from
// LDS atomics, so their order varies between runs; that changes no pixel.
__global__ void __launch_bounds__(kThreads) k_queue_order(pgs_order_args a)
{
	__shared__ uint32_t hist[kBuckets], base[kBuckets];
	for (uint32_t b = threadIdx.x; b < kBuckets; b += kThreads)
		hist[b] = 0;
	__syncthreads();
	auto bucket = [](float p) {
		const float q = fminf(fmaxf(p * 8.0f, 0.0f), float(kBuckets - 1));
		return uint32_t(kBuckets - 1) - uint32_t(q); // largest cost first
	};
	for (uint32_t i = threadIdx.x; i < a.count; i += kThreads)
		atomicAdd(&hist[bucket(a.predicted[i])], 1u);
	__syncthreads();
	if (threadIdx.x == 0)
	{
		uint32_t sum = 0;
		for (uint32_t b = 0; b < kBuckets; b++)
		{
			base[b] = sum;
			sum += hist[b];
		}
		*a.counter = 0;
	}
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < a.count; i += kThreads)
		a.order[atomicAdd(&base[bucket(a.predicted[i])], 1u)] = i;
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *dst, uint32_t count, uint32_t value)
{
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i < count)
		dst[i] = value;
}
}
}

extern "C" int pgs_k_queue_order(const pgs_order_args *a, hipStream_t stream)
{
	if (!a || !a->predicted || !a->order || !a->counter)
		return -int(hipErrorInvalidValue);
	hipLaunchKernelGGL(pgs::k_queue_order, dim3(1), dim3(pgs::kThreads), 0, stream, *a);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

extern "C" int pgs_k_fill_u32(uint32_t *dst, uint32_t count, uint32_t value, hipStream_t stream)
{
	if (!dst)
		return -int(hipErrorInvalidValue);
	if (count == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_fill_u32, dim3((count + 255u) / 256u), dim3(256), 0, stream, dst, count, value);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}
