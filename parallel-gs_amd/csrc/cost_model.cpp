// Learned tile-cost model; see cost_model.hpp.
#include "cost_model.hpp"
#include "pgs_kernels.h"
#include <cmath>
#include <cstring>
#include <numeric>
#include <algorithm>
#include <stdexcept>

namespace pgs
{
namespace
{
void ok(hipError_t e)
{
	if (e != hipSuccess)
		throw std::runtime_error(hipGetErrorString(e));
}
void ok_rc(int rc)
{
	if (rc != 0)
		throw std::runtime_error("MLP kernel launch failed");
}

struct Xoshiro
{
	uint64_t s[4];
	explicit Xoshiro(uint64_t seed)
	{
		for (auto &v : s)
		{
			seed += 0x9e3779b97f4a7c15ull;
			uint64_t z = seed;
			z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
			z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
			v = z ^ (z >> 31);
		}
	}
	static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
	uint64_t next()
	{
		const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
		s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
		return r;
	}
	double unit() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};
}

CostModel::~CostModel()
{
	for (void *p : { (void *)master_, (void *)weights_, (void *)grads_, (void *)m_, (void *)v_, (void *)x_, (void *)t_,
	                 (void *)scratch_, (void *)out_ })
		if (p)
			(void)hipFree(p);
	if (stream_)
		(void)hipStreamDestroy(stream_);
}

bool CostModel::init(uint64_t seed)
{
	if (!stream_)
		ok(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
	if (!master_)
	{
		ok(hipMalloc(&master_, PGS_MLP_PARAMS * sizeof(float)));
		ok(hipMalloc(&weights_, PGS_MLP_PARAMS * sizeof(uint16_t)));
		ok(hipMalloc(&grads_, PGS_MLP_PARAMS * sizeof(float)));
		ok(hipMalloc(&m_, PGS_MLP_PARAMS * sizeof(float)));
		ok(hipMalloc(&v_, PGS_MLP_PARAMS * sizeof(float)));
	}
	std::vector<float> p(PGS_MLP_PARAMS, 0.0f);
	Xoshiro rng(seed);
	const uint32_t NI = PGS_MLP_INPUTS, NH = PGS_MLP_HIDDEN;
	auto fill = [&](uint32_t off, uint32_t n, uint32_t fan_in) {
		const double lim = std::sqrt(6.0 / fan_in);
		for (uint32_t i = 0; i < n; i++)
			p[off + i] = float((rng.unit() * 2.0 - 1.0) * lim);
	};
	fill(0, NH * NI, NI);                         // W1
	fill(NH * NI + NH, NH * NH, NH);              // W2
	fill(NH * NI + NH + NH * NH + NH, NH, NH);    // w3
	set_params(p.data());
	return true;
}

void CostModel::set_params(const float *host)
{
	ok(hipMemcpy(master_, host, PGS_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
	ok(hipMemsetAsync(m_, 0, PGS_MLP_PARAMS * sizeof(float), stream_));
	ok(hipMemsetAsync(v_, 0, PGS_MLP_PARAMS * sizeof(float), stream_));
	ok_rc(pgs_k_mlp_quantize(master_, weights_, stream_));
	ok(hipStreamSynchronize(stream_));
}

void CostModel::get_params(float *host) const
{
	ok(hipMemcpy(host, master_, PGS_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
}

void CostModel::ensure(uint32_t count, uint32_t batch)
{
	if (count > cap_)
	{
		if (x_) ok(hipFree(x_));
		if (t_) ok(hipFree(t_));
		if (out_) ok(hipFree(out_));
		ok(hipMalloc(&x_, size_t(count) * PGS_MLP_INPUTS * sizeof(float)));
		ok(hipMalloc(&t_, size_t(count) * sizeof(float)));
		ok(hipMalloc(&out_, size_t(count) * sizeof(float)));
		cap_ = count;
	}
	if (batch > scratch_cap_)
	{
		if (scratch_) ok(hipFree(scratch_));
		ok(hipMalloc(&scratch_, size_t(batch) * PGS_MLP_SAMPLE_FLOATS * sizeof(float)));
		scratch_cap_ = batch;
	}
}

void CostModel::predict_device(const float *features, uint32_t count, float *out, hipStream_t stream) const
{
	pgs_mlp_forward_args a{};
	a.weights = weights_;
	a.features = features;
	a.out = out;
	a.count = count;
	ok_rc(pgs_k_mlp_forward(&a, stream));
}

void CostModel::predict(const float *features_host, uint32_t count, float *out_host)
{
	ensure(count, 1);
	ok(hipMemcpy(x_, features_host, size_t(count) * PGS_MLP_INPUTS * sizeof(float), hipMemcpyHostToDevice));
	predict_device(x_, count, out_, stream_);
	ok(hipStreamSynchronize(stream_));
	ok(hipMemcpy(out_host, out_, size_t(count) * sizeof(float), hipMemcpyDeviceToHost));
}

void CostModel::gradients(const float *features_host, const float *targets_host, uint32_t count, float *grads_host)
{
	ensure(count, count);
	ok(hipMemcpy(x_, features_host, size_t(count) * PGS_MLP_INPUTS * sizeof(float), hipMemcpyHostToDevice));
	ok(hipMemcpy(t_, targets_host, size_t(count) * sizeof(float), hipMemcpyHostToDevice));
	pgs_mlp_grad_args g{};
	g.weights = weights_;
	g.features = x_;
	g.targets = t_;
	g.scratch = scratch_;
	g.grads = grads_;
	g.count = count;
	ok_rc(pgs_k_mlp_gradients(&g, stream_));
	ok(hipStreamSynchronize(stream_));
	ok(hipMemcpy(grads_host, grads_, PGS_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
}

float CostModel::train(const float *features_host, const float *targets_host, uint32_t count, uint32_t epochs,
                       uint32_t batch, float lr, uint64_t seed)
{
	if (count == 0 || batch == 0)
		return 0.0f;
	batch = std::min(batch, count);
	ensure(count, batch);
	Xoshiro rng(seed);
	std::vector<uint32_t> order(count);
	std::iota(order.begin(), order.end(), 0u);
	std::vector<float> xs(features_host, features_host + size_t(count) * PGS_MLP_INPUTS);
	std::vector<float> ts(targets_host, targets_host + count);
	ok(hipMemcpyAsync(x_, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
	ok(hipMemcpyAsync(t_, ts.data(), ts.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
	ok(hipStreamSynchronize(stream_)); // the host copies are rewritten by the first shuffle
	const float b1 = 0.9f, b2 = 0.999f;
	float b1t = 1.0f, b2t = 1.0f;
	for (uint32_t e = 0; e < epochs; e++)
	{
		for (uint32_t i = count - 1; i > 0; i--)
			std::swap(order[i], order[uint32_t(rng.next() % (uint64_t(i) + 1u))]);
		for (uint32_t i = 0; i < count; i++)
		{
			memcpy(&xs[size_t(i) * PGS_MLP_INPUTS], features_host + size_t(order[i]) * PGS_MLP_INPUTS,
			       PGS_MLP_INPUTS * sizeof(float));
			ts[i] = targets_host[order[i]];
		}
		ok(hipMemcpyAsync(x_, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
		ok(hipMemcpyAsync(t_, ts.data(), ts.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
		for (uint32_t s = 0; s + batch <= count; s += batch)
		{
			pgs_mlp_grad_args g{};
			g.weights = weights_;
			g.features = x_ + size_t(s) * PGS_MLP_INPUTS;
			g.targets = t_ + s;
			g.scratch = scratch_;
			g.grads = grads_;
			g.count = batch;
			ok_rc(pgs_k_mlp_gradients(&g, stream_));
			b1t *= b1;
			b2t *= b2;
			pgs_mlp_adam_args a{};
			a.master = master_;
			a.weights = weights_;
			a.grads = grads_;
			a.m = m_;
			a.v = v_;
			a.lr = lr;
			a.beta1 = b1;
			a.beta2 = b2;
			a.eps = 1e-8f;
			a.beta1_t = b1t;
			a.beta2_t = b2t;
			ok_rc(pgs_k_mlp_adam(&a, stream_));
		}
		ok(hipStreamSynchronize(stream_));
	}
	// Final mean squared error over all samples (in the last shuffled order).
	predict_device(x_, count, out_, stream_);
	ok(hipStreamSynchronize(stream_));
	std::vector<float> y(count);
	ok(hipMemcpy(y.data(), out_, size_t(count) * sizeof(float), hipMemcpyDeviceToHost));
	double se = 0.0;
	for (uint32_t i = 0; i < count; i++)
		se += double(y[i] - ts[i]) * double(y[i] - ts[i]);
	return float(se / count);
}
}
