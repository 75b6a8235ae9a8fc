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

// Queue of the tiles with work. Every tile with work becomes `weight` consecutive entries:
// - work queue: items of at most kItemWaves waves (up to 16 per tile); with predictions a tile
//   whose predicted cost (2^log2) is above t = total / (2 * slots) is split further, up to 16
//   items of equal wave ranges (so that no item dominates the makespan of `slots` resident
//   workgroups);
// - dispatch order: one entry per chunk of group_waves waves.
// Tile order: each thread takes a contiguous range of tiles, sums their weights and starts
// after the exclusive sum of the ranges before it. Predicted order: tiles whose entries are
// heavy (predicted work per entry above t) come first, counting-sorted over 256 buckets of 1/8
// in log2 cycles, largest first, with positions inside a bucket from LDS atomics (their order
// varies between runs, which changes no pixel); the other tiles follow in tile order, which
// keeps neighbouring tiles, which share primitives in L2, together. k_expand_order then writes
// each tile's entries. The split needs at most
// count + 2 * slots entries beyond the base items (ceil(c / t) <= c / t + 1); if rounding ever
// made the items overflow `capacity`, every tile keeps its base items.
constexpr uint32_t kItemWaves = 16;

__global__ void __launch_bounds__(kThreads) k_queue_order(pgs_order_args a)
{
	__shared__ uint32_t hist[kThreads > kBuckets ? kThreads : kBuckets];
	__shared__ uint32_t base[kBuckets];
	__shared__ float partial[kThreads / 64];
	__shared__ float target;
	__shared__ uint32_t split, fits;
	const bool dispatch = a.dispatch != 0;
	if (threadIdx.x == 0)
		fits = 0;
	auto has_work = [&](uint32_t i) { return a.features[size_t(i) * PGS_MLP_INPUTS] != 0.0f; };
	auto base_weight = [&](uint32_t i) {
		const uint32_t w = a.tile_waves[i];
		return dispatch ? (w + a.group_waves - 1u) / a.group_waves
		                : min(16u, max(1u, (w + kItemWaves - 1u) / kItemWaves));
	};
	if (!a.predicted)
	{
		const uint32_t chunk = (a.count + kThreads - 1u) / kThreads;
		const uint32_t begin = min(threadIdx.x * chunk, a.count), end = min(begin + chunk, a.count);
		uint32_t n = 0;
		for (uint32_t i = begin; i < end; i++)
			n += has_work(i) ? base_weight(i) : 0u;
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
			fits = sum <= a.capacity;
			a.counter[0] = 0;
			a.counter[1] = fits ? sum : 0u; // an undersized buffer shades nothing rather than overflow
		}
		__syncthreads();
		uint32_t pos = hist[threadIdx.x];
		for (uint32_t i = begin; i < end; i++)
		{
			const uint32_t w = (fits && has_work(i)) ? base_weight(i) : 0u;
			a.run_start[i] = pos;
			a.run_weight[i] = w;
			pos += w;
		}
		return;
	}

	// Total predicted cost: the split target and the heavy threshold.
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
		split = dispatch ? 0u : 1u;
	}
	__syncthreads();
	auto weight = [&](uint32_t i, float p) -> uint32_t {
		const uint32_t w0 = base_weight(i);
		if (dispatch || !split || !(target > 0.0f))
			return w0;
		const float c = exp2f(fminf(p, 40.0f));
		return max(w0, uint32_t(fminf(fmaxf(ceilf(c / target), 1.0f), 16.0f)));
	};
	// Heavy: one entry's predicted work (a queue item, or a whole tile in dispatch order) is
	// above the target. Heavy tiles go first, largest first; the others follow in tile order,
	// which keeps neighbouring tiles (sharing primitives in L2) together.
	auto entry_cost = [&](float p, uint32_t w) { return dispatch ? p : p - log2f(float(w)); };
	const float heavy_log2 = target > 0.0f ? log2f(target) : 1e30f;
	auto bucket = [](float p) {
		const float q = fminf(fmaxf(p * 8.0f, 0.0f), float(kBuckets - 1));
		return uint32_t(kBuckets - 1) - uint32_t(q); // largest cost first
	};
	const uint32_t chunk = (a.count + kThreads - 1u) / kThreads;
	const uint32_t begin = min(threadIdx.x * chunk, a.count), end = min(begin + chunk, a.count);
	__shared__ uint32_t light_base[kThreads];
	for (int attempt = 0; attempt < 2; attempt++)
	{
		for (uint32_t b = threadIdx.x; b < kBuckets; b += kThreads)
			hist[b] = 0;
		__syncthreads();
		uint32_t light = 0;
		for (uint32_t i = begin; i < end; i++)
			if (has_work(i))
			{
				const float p = a.predicted[i];
				const uint32_t w = weight(i, p);
				const float c = entry_cost(p, w);
				if (c >= heavy_log2)
					atomicAdd(&hist[bucket(c)], w);
				else
					light += w;
			}
		light_base[threadIdx.x] = light;
		__syncthreads();
		if (threadIdx.x == 0)
		{
			uint32_t total = 0;
			for (uint32_t b = 0; b < kBuckets; b++)
			{
				base[b] = total;
				total += hist[b];
			}
			for (uint32_t t = 0; t < kThreads; t++)
			{
				const uint32_t c = light_base[t];
				light_base[t] = total;
				total += c;
			}
			if (total > a.capacity)
				split = 0; // rounding overflow: count again with the base items
			else
			{
				a.counter[0] = 0;
				a.counter[1] = total;
				fits = 1u; // success marker for the other threads
			}
		}
		__syncthreads();
		if (fits == 1u)
			break;
		__syncthreads();
	}
	const bool ok = fits == 1u;
	if (!ok && threadIdx.x == 0)
	{
		// The base items do not fit either (an undersized buffer): shade nothing rather than
		// overflow.
		a.counter[0] = 0;
		a.counter[1] = 0;
	}
	uint32_t light_pos = light_base[threadIdx.x];
	for (uint32_t i = begin; i < end; i++)
	{
		uint32_t w = 0, pos = 0;
		if (ok && has_work(i))
		{
			const float p = a.predicted[i];
			w = weight(i, p);
			const float c = entry_cost(p, w);
			if (c >= heavy_log2)
				pos = atomicAdd(&base[bucket(c)], w);
			else
			{
				pos = light_pos;
				light_pos += w;
			}
		}
		a.run_start[i] = pos;
		a.run_weight[i] = w;
	}
}

// Writes the entries of every tile placed by k_queue_order, one thread per tile.
__global__ void __launch_bounds__(256) k_expand_order(pgs_order_args a)
{
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i >= a.count)
		return;
	const uint32_t w = a.run_weight[i], pos = a.run_start[i];
	for (uint32_t k = 0; k < w; k++)
		a.order[pos + k] = a.dispatch ? (i << 12) | k : (i << 8) | (k << 4) | (w - 1u);
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
	if (!a || !a->features || !a->order || !a->counter || !a->tile_waves || !a->run_start || !a->run_weight ||
	    a->count >= (1u << 24) || a->capacity < a->count)
		return -int(hipErrorInvalidValue);
	if (a->dispatch && (a->count >= (1u << 20) || a->group_waves == 0 || a->group_waves > 16))
		return -int(hipErrorInvalidValue);
	if (!a->dispatch && a->predicted && a->capacity - a->count < 2u * a->slots + 1024u)
		return -int(hipErrorInvalidValue);
	hipLaunchKernelGGL(pgs::k_queue_order, dim3(1), dim3(pgs::kThreads), 0, stream, *a);
	if (a->count)
		hipLaunchKernelGGL(pgs::k_expand_order, dim3((a->count + 255u) / 256u), dim3(256), 0, stream, *a);
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
