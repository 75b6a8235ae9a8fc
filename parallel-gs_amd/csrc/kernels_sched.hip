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

// Queue of the tiles with work. Tile order: each thread takes a contiguous chunk, counts its
// tiles with work, and writes them after the exclusive sum of the counts of the chunks before
// it. Cost order: counting sort by quantized predicted cost, descending; positions inside a
// bucket come from LDS atomics, so their order varies between runs, which changes no pixel.
__global__ void __launch_bounds__(kThreads) k_queue_order(pgs_order_args a)
{
	__shared__ uint32_t hist[kThreads > kBuckets ? kThreads : kBuckets];
	__shared__ uint32_t base[kBuckets];
	auto has_work = [&](uint32_t i) { return a.features[size_t(i) * PGS_MLP_INPUTS] != 0.0f; };
	if (!a.predicted)
	{
		const uint32_t chunk = (a.count + kThreads - 1u) / kThreads;
		const uint32_t begin = min(threadIdx.x * chunk, a.count), end = min(begin + chunk, a.count);
		uint32_t n = 0;
		for (uint32_t i = begin; i < end; i++)
			n += has_work(i) ? 1u : 0u;
		hist[threadIdx.x] = n;
		__syncthreads();
		if (threadIdx.x == 0)
		{
			uint32_t sum = 0;
			for (uint32_t t = 0; t < kThreads; t++)
			{
				const uint32_t c = hist[t];
				hist[t] = sum;
				sum += c;
			}
			a.counter[0] = 0;
			a.counter[1] = sum;
		}
		__syncthreads();
		uint32_t pos = hist[threadIdx.x];
		for (uint32_t i = begin; i < end; i++)
			if (has_work(i))
				a.order[pos++] = i;
		return;
	}
	for (uint32_t b = threadIdx.x; b < kBuckets; b += kThreads)
		hist[b] = 0;
	__syncthreads();
	auto bucket = [](float p) {
		const float q = fminf(fmaxf(p * 8.0f, 0.0f), float(kBuckets - 1));
		return uint32_t(kBuckets - 1) - uint32_t(q); // largest cost first
	};
	for (uint32_t i = threadIdx.x; i < a.count; i += kThreads)
		if (has_work(i))
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
		a.counter[0] = 0;
		a.counter[1] = sum;
	}
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < a.count; i += kThreads)
		if (has_work(i))
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
	if (!a || !a->features || !a->order || !a->counter)
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
