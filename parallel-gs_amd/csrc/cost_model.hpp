// Learned tile-cost model (no counterpart in the reference): the MLP of kernels_mlp.hip with
// its parameters in device memory, trained on the GPU from measured per-workgroup shading
// costs and used to order the batched renderer's persistent shading queue.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace pgs
{
class CostModel
{
public:
	CostModel() = default;
	~CostModel();
	CostModel(const CostModel &) = delete;
	CostModel &operator=(const CostModel &) = delete;

	// He-uniform weights from a seeded xoshiro256** stream, zero biases.
	bool init(uint64_t seed);
	void set_params(const float *host);
	void get_params(float *host) const;

	// Predictions for `count` feature rows already in device memory.
	void predict_device(const float *features, uint32_t count, float *out, hipStream_t stream) const;
	void predict(const float *features_host, uint32_t count, float *out_host);
	// Gradient of the mean squared error (host in, host out), for checks.
	void gradients(const float *features_host, const float *targets_host, uint32_t count, float *grads_host);
	// Mini-batch Adam over shuffled samples; returns the mean squared error over all samples
	// after training.
	float train(const float *features_host, const float *targets_host, uint32_t count, uint32_t epochs,
	            uint32_t batch, float lr, uint64_t seed);

	const uint16_t *weights_device() const { return weights_; }

private:
	hipStream_t stream_ = nullptr;
	float *master_ = nullptr;
	uint16_t *weights_ = nullptr;
	float *grads_ = nullptr, *m_ = nullptr, *v_ = nullptr;
	float *x_ = nullptr, *t_ = nullptr, *scratch_ = nullptr, *out_ = nullptr;
	uint32_t cap_ = 0, scratch_cap_ = 0;

	void ensure(uint32_t count, uint32_t batch);
};
}
