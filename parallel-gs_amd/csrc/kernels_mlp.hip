// Tile-cost MLP for the learned tile-queue order (no counterpart in the reference).
//
// 24 features -> 64 -> 64 -> 1 with ReLU. Weights, biases and every layer input are BF16
// (round to nearest even); each output accumulates in FP32 starting from its bias and adding
// w * x in input order. Products of two BF16 values are exact in FP32, so results depend only
// on that fixed summation order and are reproducible bit for bit on the CPU.
//
// Training runs on the GPU with the kernels below: per-sample forward + backward (mean squared
// error, straight-through BF16 rounding of activations), a per-parameter gradient reduction in
// sample order, and Adam on FP32 master weights that refreshes the BF16 copy.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include "pgs_kernels.h"

// Per-sample training scratch of pgs_k_mlp_gradients_noalloc: a fixed device array (each GPU
// the library runs on holds its own copy until the process ends). Host code never passes the
// array's name to a kernel; it passes the device address hipGetSymbolAddress returns for the
// current device.
static __device__ __align__(256) float scratch[size_t(PGS_MLP_NOALLOC_MAX_SAMPLES) * PGS_MLP_SAMPLE_FLOATS];

namespace pgs
{
namespace
{
constexpr uint32_t NI = PGS_MLP_INPUTS, NH = PGS_MLP_HIDDEN;
constexpr uint32_t OFF_W1 = 0, OFF_B1 = NH * NI, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH;
constexpr uint32_t OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH, NP = OFF_B3 + 1;
static_assert(NP == PGS_MLP_PARAMS, "parameter count");

__device__ __forceinline__ float bf16_round(float x)
{
	uint32_t u = __float_as_uint(x);
	u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
	return __uint_as_float(u);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ uint16_t f32_to_bf16(float x) { return uint16_t(__float_as_uint(bf16_round(x)) >> 16); }

// Forward for one sample with weights in LDS (FP32 copies of the BF16 values).
__device__ __forceinline__ float forward_one(const float *w, const float *x_in, float *xb, float *z1, float *h1, float *z2,
                                             float *h2)
{
	for (uint32_t f = 0; f < NI; f++)
		xb[f] = bf16_round(x_in[f]);
	for (uint32_t k = 0; k < NH; k++)
	{
		float acc = w[OFF_B1 + k];
		for (uint32_t f = 0; f < NI; f++)
			acc = __fadd_rn(acc, __fmul_rn(w[OFF_W1 + k * NI + f], xb[f]));
		z1[k] = acc;
		h1[k] = bf16_round(fmaxf(acc, 0.0f));
	}
	for (uint32_t j = 0; j < NH; j++)
	{
		float acc = w[OFF_B2 + j];
		for (uint32_t k = 0; k < NH; k++)
			acc = __fadd_rn(acc, __fmul_rn(w[OFF_W2 + j * NH + k], h1[k]));
		z2[j] = acc;
		h2[j] = bf16_round(fmaxf(acc, 0.0f));
	}
	float y = w[OFF_B3];
	for (uint32_t j = 0; j < NH; j++)
		y = __fadd_rn(y, __fmul_rn(w[OFF_W3 + j], h2[j]));
	return y;
}

__device__ __forceinline__ void stage_weights(const uint16_t *wq, float *w)
{
	for (uint32_t i = threadIdx.x; i < NP; i += blockDim.x)
		w[i] = bf16_to_f32(wq[i]);
	__syncthreads();
}

__global__ void __launch_bounds__(256) k_mlp_forward(pgs_mlp_forward_args a)
{
	__shared__ float w[NP];
	stage_weights(a.weights, w);
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.count)
		return;
	float xb[NI], z1[NH], h1[NH], z2[NH], h2[NH];
	a.out[i] = forward_one(w, a.features + size_t(i) * NI, xb, z1, h1, z2, h2);
}

// Per-sample backward pass. Writes the quantities the gradient reduction needs:
// xb[24], h1[64], h2[64], dz1[64], dz2[64], dy (stride PGS_MLP_SAMPLE_FLOATS).
__global__ void __launch_bounds__(128) k_mlp_backward(pgs_mlp_grad_args a)
{
	__shared__ float w[NP];
	stage_weights(a.weights, w);
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.count)
		return;
	float xb[NI], z1[NH], h1[NH], z2[NH], h2[NH];
	const float y = forward_one(w, a.features + size_t(i) * NI, xb, z1, h1, z2, h2);
	const float dy = __fdiv_rn(__fmul_rn(2.0f, __fsub_rn(y, a.targets[i])), float(a.count));
	float *s = a.scratch + size_t(i) * PGS_MLP_SAMPLE_FLOATS;
	for (uint32_t f = 0; f < NI; f++)
		s[f] = xb[f];
	for (uint32_t k = 0; k < NH; k++)
		s[NI + k] = h1[k];
	for (uint32_t j = 0; j < NH; j++)
		s[NI + NH + j] = h2[j];
	float dz2[NH];
	for (uint32_t j = 0; j < NH; j++)
	{
		dz2[j] = z2[j] > 0.0f ? __fmul_rn(dy, w[OFF_W3 + j]) : 0.0f;
		s[NI + 3 * NH + j] = dz2[j];
	}
	for (uint32_t k = 0; k < NH; k++)
	{
		float acc = 0.0f;
		for (uint32_t j = 0; j < NH; j++)
			acc = __fadd_rn(acc, __fmul_rn(dz2[j], w[OFF_W2 + j * NH + k]));
		s[NI + 2 * NH + k] = z1[k] > 0.0f ? acc : 0.0f;
	}
	s[NI + 4 * NH] = dy;
	if (a.loss)
		a.loss[i] = __fmul_rn(__fsub_rn(y, a.targets[i]), __fsub_rn(y, a.targets[i]));
}

// One thread per parameter: sums its gradient over the samples in order.
__global__ void __launch_bounds__(256) k_mlp_reduce(pgs_mlp_grad_args a)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= NP)
		return;
	const float *s = a.scratch;
	const uint32_t S = PGS_MLP_SAMPLE_FLOATS;
	float g = 0.0f;
	if (p < OFF_B1)
	{
		const uint32_t k = p / NI, f = p % NI;
		for (uint32_t i = 0; i < a.count; i++)
			g = __fadd_rn(g, __fmul_rn(s[i * S + NI + 2 * NH + k], s[i * S + f]));
	}
	else if (p < OFF_W2)
	{
		const uint32_t k = p - OFF_B1;
		for (uint32_t i = 0; i < a.count; i++)
			g = __fadd_rn(g, s[i * S + NI + 2 * NH + k]);
	}
	else if (p < OFF_B2)
	{
		const uint32_t j = (p - OFF_W2) / NH, k = (p - OFF_W2) % NH;
		for (uint32_t i = 0; i < a.count; i++)
			g = __fadd_rn(g, __fmul_rn(s[i * S + NI + 3 * NH + j], s[i * S + NI + k]));
	}
	else if (p < OFF_W3)
	{
		const uint32_t j = p - OFF_B2;
		for (uint32_t i = 0; i < a.count; i++)
			g = __fadd_rn(g, s[i * S + NI + 3 * NH + j]);
	}
	else if (p < OFF_B3)
	{
		const uint32_t j = p - OFF_W3;
		for (uint32_t i = 0; i < a.count; i++)
			g = __fadd_rn(g, __fmul_rn(s[i * S + NI + 4 * NH], s[i * S + NI + NH + j]));
	}
	else
	{
		for (uint32_t i = 0; i < a.count; i++)
			g = __fadd_rn(g, s[i * S + NI + 4 * NH]);
	}
	a.grads[p] = g;
}

// Adam on the FP32 master parameters; refreshes the BF16 copy used by the forward pass.
__global__ void __launch_bounds__(256) k_mlp_adam(pgs_mlp_adam_args a)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= NP)
		return;
	const float g = a.grads[p];
	const float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * g;
	const float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * g * g;
	a.m[p] = m;
	a.v[p] = v;
	const float mh = m / (1.0f - a.beta1_t), vh = v / (1.0f - a.beta2_t);
	const float w = a.master[p] - a.lr * mh / (sqrtf(vh) + a.eps);
	a.master[p] = w;
	a.weights[p] = f32_to_bf16(w);
}

__global__ void __launch_bounds__(256) k_mlp_quantize(const float *master, uint16_t *weights)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p < NP)
		weights[p] = f32_to_bf16(master[p]);
}
}
}

extern "C" int pgs_k_mlp_forward(const pgs_mlp_forward_args *a, hipStream_t stream)
{
	if (!a || !a->weights || !a->features || !a->out)
		return -int(hipErrorInvalidValue);
	if (a->count == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_mlp_forward, dim3((a->count + 255u) / 256u), dim3(256), 0, stream, *a);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

extern "C" int pgs_k_mlp_gradients(const pgs_mlp_grad_args *a, hipStream_t stream)
{
	if (!a || !a->weights || !a->features || !a->targets || !a->scratch || !a->grads)
		return -int(hipErrorInvalidValue);
	if (a->count == 0)
		return 0;
	hipLaunchKernelGGL(pgs::k_mlp_backward, dim3((a->count + 127u) / 128u), dim3(128), 0, stream, *a);
	hipLaunchKernelGGL(pgs::k_mlp_reduce, dim3((PGS_MLP_PARAMS + 255u) / 256u), dim3(256), 0, stream, *a);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

namespace
{
// Device address of `scratch` per device, looked up once per device with that device current.
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
float *scratch_ptr[kMaxDevices] = {};

int scratch_address(float **out)
{
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
		return -int(hipErrorInvalidDevice);
	std::lock_guard<std::mutex> lock(scratch_mutex);
	if (!scratch_ptr[dev])
	{
		void *ptr = nullptr;
		const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (rc != hipSuccess || !ptr)
		{
			fprintf(stderr, "pgs: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev, hipGetErrorString(rc));
			return -int(rc != hipSuccess ? rc : hipErrorInvalidSymbol);
		}
		scratch_ptr[dev] = static_cast<float *>(ptr);
	}
	*out = scratch_ptr[dev];
	return 0;
}
}

extern "C" int pgs_k_mlp_gradients_noalloc(const pgs_mlp_grad_args *a, hipStream_t stream)
{
	if (!a || !a->weights || !a->features || !a->targets || !a->grads)
		return -int(hipErrorInvalidValue);
	if (a->count > PGS_MLP_NOALLOC_MAX_SAMPLES)
	{
		fprintf(stderr, "pgs: pgs_k_mlp_gradients_noalloc: %u samples exceed the fixed scratch of %u samples\n",
		        a->count, PGS_MLP_NOALLOC_MAX_SAMPLES);
		return -int(hipErrorInvalidValue);
	}
	if (a->count == 0)
		return 0;
	float *dev_scratch = nullptr;
	const int rc = scratch_address(&dev_scratch);
	if (rc != 0)
		return rc;
	// The backward kernel writes every scratch float of samples [0, count) before the reduction
	// reads them, so nothing needs clearing between calls.
	pgs_mlp_grad_args b = *a;
	b.scratch = dev_scratch;
	hipLaunchKernelGGL(pgs::k_mlp_backward, dim3((b.count + 127u) / 128u), dim3(128), 0, stream, b);
	hipLaunchKernelGGL(pgs::k_mlp_reduce, dim3((PGS_MLP_PARAMS + 255u) / 256u), dim3(256), 0, stream, b);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

extern "C" int pgs_k_mlp_adam(const pgs_mlp_adam_args *a, hipStream_t stream)
{
	if (!a || !a->master || !a->weights || !a->grads || !a->m || !a->v)
		return -int(hipErrorInvalidValue);
	hipLaunchKernelGGL(pgs::k_mlp_adam, dim3((PGS_MLP_PARAMS + 255u) / 256u), dim3(256), 0, stream, *a);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}

extern "C" int pgs_k_mlp_quantize(const float *master, uint16_t *weights, hipStream_t stream)
{
	if (!master || !weights)
		return -int(hipErrorInvalidValue);
	hipLaunchKernelGGL(pgs::k_mlp_quantize, dim3((PGS_MLP_PARAMS + 255u) / 256u), dim3(256), 0, stream, master, weights);
	return hipGetLastError() == hipSuccess ? 0 : -int(hipErrorLaunchFailure);
}
