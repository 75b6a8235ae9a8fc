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
// it; every tile is one item. Cost order: the predicted costs (2^log2) are summed, a tile
// whose cost is above total / (2 * slots) is split into up to 16 items of equal wave ranges
// (so that no item dominates the makespan of `slots` resident workgroups), and the items are
// counting-sorted by quantized predicted cost per item, largest first. Positions inside a
// bucket come from LDS atomics, so their order varies between runs, which changes no pixel.
// The split needs at most count + 2 * slots entries (ceil(c / t) <= c / t + 1); if rounding
// ever made the items overflow `capacity`, every tile stays whole.
__device__ __forceinline__ uint32_t entry(uint32_t tile, uint32_t part, uint32_t parts)
{
	return (tile << 8) | (part << 4) | (parts - 1u);
}

__global__ void __launch_bounds__(kThreads) k_queue_order(pgs_order_args a)
{
	__shared__ uint32_t hist[kThreads > kBuckets ? kThreads : kBuckets];
	__shared__ uint32_t base[kBuckets];
	__shared__ float partial[kThreads / 64];
	__shared__ float target;
	__shared__ uint32_t split;
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
				a.order[pos++] = entry(i, 0, 1);
		return;
	}

	// Total predicted cost.
	float sum = 0.0f;
	for (uint32_t i = threadIdx.x; i < a.count; i += kThreads)
		if (has_work(i))
			sum += exp2f(fminf(a.predicted[i], 40.0f));
	for (int m = 32; m >= 1; m >>= 1)
		sum += __shfl_xor(sum, m);
	if ((threadIdx.x & 63u) == 0)
		partial[threadIdx.x >> 6] = sum;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		float t = 0.0f;
		for (uint32_t w = 0; w < kThreads / 64; w++)
			t += partial[w];
		target = t / float(2u * max(a.slots, 1u));
		split = 1;
	}
	__syncthreads();
	auto parts_of = [&](float p) -> uint32_t {
		if (!split || !(target > 0.0f))
			return 1u;
		const float c = exp2f(fminf(p, 40.0f));
		return uint32_t(fminf(fmaxf(ceilf(c / target), 1.0f), 16.0f));
	};
	auto bucket = [](float p) {
		const float q = fminf(fmaxf(p * 8.0f, 0.0f), float(kBuckets - 1));
		return uint32_t(kBuckets - 1) - uint32_t(q); // largest cost first
	};
	// log2 of the cost of one item of a tile cut into `parts` items.
	auto item_cost = [](float p, uint32_t parts) { return p - log2f(float(parts)); };
	for (int attempt = 0; attempt < 2; attempt++)
	{
		for (uint32_t b = threadIdx.x; b < kBuckets; b += kThreads)
			hist[b] = 0;
		__syncthreads();
		for (uint32_t i = threadIdx.x; i < a.count; i += kThreads)
			if (has_work(i))
			{
				const float p = a.predicted[i];
				const uint32_t parts = parts_of(p);
				atomicAdd(&hist[bucket(item_cost(p, parts))], parts);
			}
		__syncthreads();
		if (threadIdx.x == 0)
		{
			uint32_t total = 0;
			for (uint32_t b = 0; b < kBuckets; b++)
			{
				base[b] = total;
				total += hist[b];
			}
			if (total > a.capacity)
				split = 0; // rounding overflow: count again with whole tiles
			else
			{
				a.counter[0] = 0;
				a.counter[1] = total;
				hist[0] = 0xffffffffu; // marks success for the other threads
			}
		}
		__syncthreads();
		if (hist[0] == 0xffffffffu)
			break;
		__syncthreads();
	}
	for (uint32_t i = threadIdx.x; i < a.count; i += kThreads)
		if (has_work(i))
		{
			const float p = a.predicted[i];
			const uint32_t parts = parts_of(p);
			const uint32_t pos = atomicAdd(&base[bucket(item_cost(p, parts))], parts);
			for (uint32_t k = 0; k < parts; k++)
				a.order[pos + k] = entry(i, k, parts);
		}
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
	if (!a || !a->features || !a->order || !a->counter || a->count >= (1u << 24) || a->capacity < a->count)
		return -int(hipErrorInvalidValue);
	if (a->predicted && a->capacity - a->count < 2u * a->slots + 1024u)
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
