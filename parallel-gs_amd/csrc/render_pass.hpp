// Primitive assembly and render-pass batching (host side of drawing).
#pragma once
#include "gs_defs.hpp"
#include "renderer.hpp"
#include <vector>

namespace pgs
{
struct RegisterState;

class RenderPassBuilder
{
public:
	void init(Renderer *renderer, uint32_t sample_log2)
	{
		renderer_ = renderer;
		sample_log2_ = sample_log2;
	}
	bool has_primitives() const { return num_primitives_ != 0; }
	void reset_vertex_queue() { vq_count_ = 0; }
	void flush() {}

private:
	Renderer *renderer_ = nullptr;
	uint32_t sample_log2_ = 0;
	uint32_t num_primitives_ = 0;
	uint32_t vq_count_ = 0;
};
}
