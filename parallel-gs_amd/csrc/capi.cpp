// Context-level C ABI (include/pgs_amd.h): thin wrappers around GSInterface that never
// let a C++ exception cross the boundary.
#include "pgs_amd.h"
#include "gs_interface.hpp"
#include "batch.hpp"
#include "cost_model.hpp"
#include <vector>
#include <exception>
#include <cstdio>
#include <new>
#include <stdexcept>

struct pgs_context
{
	pgs::GSInterface gs;
	uint64_t record_clock = 0; // VRAM write-log time at pgs_begin_recording
};

struct pgs_batch
{
	pgs::Batch batch;
};

struct pgs_cost_model
{
	pgs::CostModel model;
};

namespace
{
template <typename T, typename F>
int ptr_guard(T *p, F &&f)
{
	if (!p)
		return -1;
	try
	{
		return f();
	}
	catch (const std::exception &e)
	{
		fprintf(stderr, "pgs: %s\n", e.what());
		return -1;
	}
}

template <typename F>
int batch_guard(pgs_batch *b, F &&f)
{
	if (!b)
		return -1;
	try
	{
		return f();
	}
	catch (const std::exception &e)
	{
		fprintf(stderr, "pgs: %s\n", e.what());
		return -1;
	}
}

template <typename F>
int guard(pgs_context *ctx, F &&f)
{
	if (!ctx)
		return -1;
	try
	{
		f();
		return 0;
	}
	catch (const std::exception &e)
	{
		fprintf(stderr, "pgs: %s\n", e.what());
		int rc = ctx->gs.renderer().last_error();
		return rc ? rc : -1;
	}
}
}

extern "C"
{
pgs_context *pgs_create(const pgs_options *options)
{
	pgs::GSOptions o;
	if (options)
	{
		switch (options->super_sampling)
		{
		case 2: o.super_sampling = pgs::SuperSampling::X2; break;
		case 4: o.super_sampling = pgs::SuperSampling::X4; break;
		case 8: o.super_sampling = pgs::SuperSampling::X8; break;
		case 16: o.super_sampling = pgs::SuperSampling::X16; break;
		default: o.super_sampling = pgs::SuperSampling::X1; break;
		}
		o.ordered_super_sampling = options->ordered_super_sampling != 0;
		o.super_sampled_textures = options->super_sampled_textures != 0;
		if (options->upload_arena_bytes)
			o.upload_arena_bytes = options->upload_arena_bytes;
		if (options->sampler_feedback > 2u)
			return nullptr;
		o.sampler_feedback = options->sampler_feedback;
		if (options->texture_stream > 1u)
			return nullptr;
		o.texture_stream = options->texture_stream != 0;
		if (hipSetDevice(int(options->device_index)) != hipSuccess)
			return nullptr;
	}
	auto *ctx = new (std::nothrow) pgs_context;
	if (!ctx)
		return nullptr;
	try
	{
		if (!ctx->gs.init(o))
		{
			delete ctx;
			return nullptr;
		}
	}
	catch (const std::exception &)
	{
		delete ctx;
		return nullptr;
	}
	return ctx;
}

void pgs_destroy(pgs_context *ctx) { delete ctx; }

int pgs_reset_context_state(pgs_context *ctx) { return guard(ctx, [&] { ctx->gs.reset_context_state(); }); }

int pgs_gif_transfer(pgs_context *ctx, uint32_t path, const void *data, size_t size)
{
	if (path >= 4 || (size & 15u) != 0 || (!data && size))
		return -1;
	return guard(ctx, [&] { ctx->gs.gif_transfer(path, data, size); });
}

int pgs_write_register(pgs_context *ctx, uint32_t addr, uint64_t payload)
{
	return guard(ctx, [&] { ctx->gs.write_register(addr, payload); });
}

int pgs_flush(pgs_context *ctx) { return guard(ctx, [&] { ctx->gs.flush(); }); }

const void *pgs_map_vram_read(pgs_context *ctx, size_t offset, size_t size)
{
	const void *p = nullptr;
	if (guard(ctx, [&] { p = ctx->gs.map_vram_read(offset, size); }) != 0)
		return nullptr;
	return p;
}

void *pgs_map_vram_write(pgs_context *ctx, size_t offset, size_t size)
{
	void *p = nullptr;
	if (guard(ctx, [&] { p = ctx->gs.map_vram_write(offset, size); }) != 0)
		return nullptr;
	return p;
}

int pgs_end_vram_write(pgs_context *ctx, size_t offset, size_t size)
{
	return guard(ctx, [&] { ctx->gs.end_vram_write(offset, size); });
}

int pgs_read_transfer_fifo(pgs_context *ctx, void *data, uint32_t n)
{
	return guard(ctx, [&] { ctx->gs.read_transfer_fifo(data, n); });
}

int pgs_clobber_register_state(pgs_context *ctx) { return guard(ctx, [&] { ctx->gs.clobber_register_state(); }); }

int pgs_consume_flush_stats(pgs_context *ctx, pgs_flush_stats *out)
{
	return guard(ctx, [&] {
		pgs::FlushStats s = ctx->gs.consume_flush_stats();
		if (out)
		{
			out->transfer_batches = s.transfer_batches;
			out->copies = s.copies;
			out->copy_blocks = s.copy_blocks;
			out->copy_records = s.copy_records;
			out->snapshot_blocks = s.snapshot_blocks;
			out->copy_bytes = s.copy_bytes;
			out->render_passes = s.render_passes;
			out->primitives = s.primitives;
			out->syncs = s.syncs;
			out->textures_decoded = s.textures_decoded;
			out->texture_cache_hits = s.texture_cache_hits;
		}
	});
}

uint32_t pgs_num_planes(pgs_context *ctx) { return ctx ? ctx->gs.renderer().num_planes() : 0; }

int pgs_read_vram_planes(pgs_context *ctx, void *dst)
{
	return guard(ctx, [&] { ctx->gs.read_all_planes(dst); });
}

int pgs_begin_recording(pgs_context *ctx)
{
	return guard(ctx, [&] {
		ctx->gs.flush();
		ctx->gs.render_pass().clear_texture_cache();
		ctx->gs.renderer().begin_recording();
		ctx->record_clock = ctx->gs.render_pass().writes().now();
	});
}

int pgs_end_recording(pgs_context *ctx)
{
	return guard(ctx, [&] {
		ctx->gs.flush();
		ctx->gs.renderer().end_recording();
		// Replays restore only the pages the recording wrote (every VRAM write of the
		// backend, transfers and render passes, is logged with its page span).
		ctx->gs.renderer().set_restore_pages(ctx->gs.render_pass().writes().pages_written_since(ctx->record_clock));
	});
}

int pgs_replay(pgs_context *ctx, int use_graph)
{
	if (!ctx)
		return -1;
	return ctx->gs.renderer().replay(use_graph != 0);
}

int pgs_profile_enable(pgs_context *ctx, int enable)
{
	return guard(ctx, [&] { ctx->gs.renderer().profile_enable(enable != 0); });
}

int pgs_profile_collect(pgs_context *ctx, double *ms, uint64_t *launches, uint32_t kinds)
{
	return guard(ctx, [&] { ctx->gs.renderer().profile_collect(ms, launches, kinds); });
}

int pgs_sync(pgs_context *ctx) { return guard(ctx, [&] { ctx->gs.renderer().sync(); }); }

void *pgs_stream(pgs_context *ctx) { return ctx ? static_cast<void *>(ctx->gs.renderer().stream()) : nullptr; }
}

// Host-side views of the shared swizzle / block-layout code, so that CPU tests can check the
// exact functions the kernels use without a GPU.
extern "C" uint32_t pgs_host_swizzle_addr(uint32_t x, uint32_t y, uint32_t base, uint32_t stride, uint32_t psm)
{
	return pgs::swizzle_addr(x, y, base, stride, psm);
}

extern "C" void pgs_host_block_lane_pixel(uint32_t psm, uint32_t lane, uint32_t k, uint32_t *dx, uint32_t *dy)
{
	pgs::block_lane_pixel(pgs::psm_storage_class(psm), lane, k, *dx, *dy);
}

extern "C"
{
pgs_batch *pgs_batch_create(pgs_context *const *contexts, uint32_t count, uint32_t flags)
{
	if (!contexts || count == 0)
		return nullptr;
	auto *b = new (std::nothrow) pgs_batch;
	if (!b)
		return nullptr;
	try
	{
		std::vector<pgs::Renderer *> r;
		for (uint32_t i = 0; i < count; i++)
		{
			if (!contexts[i])
				throw std::runtime_error("null context");
			r.push_back(&contexts[i]->gs.renderer());
		}
		if (!b->batch.build(r.data(), count, (flags & PGS_BATCH_SEQUENTIAL) == 0))
			throw std::runtime_error("batch build failed");
		return b;
	}
	catch (const std::exception &e)
	{
		fprintf(stderr, "pgs: %s\n", e.what());
		delete b;
		return nullptr;
	}
}

void pgs_batch_destroy(pgs_batch *batch) { delete batch; }

int pgs_batch_apply_tuning(pgs_batch *batch, const pgs_batch_tuning *tuning)
{
	return batch_guard(batch, [&] {
		if (!tuning)
			return -1;
		pgs::BatchTuning t;
		t.tile_log2_1x = tuning->tile_log2_1x;
		t.waves_per_group = tuning->waves_per_group;
		return batch->batch.apply(t) ? 0 : -1;
	});
}

int pgs_batch_autotune(pgs_batch *batch, uint32_t reps, pgs_batch_tuning_result *results, uint32_t max_results)
{
	return batch_guard(batch, [&] {
		std::vector<pgs::TuningResult> table;
		if (!batch->batch.autotune(reps, &table))
			return -1;
		for (uint32_t i = 0; i < table.size() && i < max_results && results; i++)
		{
			results[i].tuning.tile_log2_1x = table[i].tuning.tile_log2_1x;
			results[i].tuning.waves_per_group = table[i].tuning.waves_per_group;
			results[i].ms_per_replay = table[i].ms;
		}
		return int(table.size());
	});
}

int pgs_batch_replay(pgs_batch *batch, int use_graph)
{
	return batch_guard(batch, [&] { return batch->batch.replay(use_graph != 0); });
}

int pgs_batch_sync(pgs_batch *batch)
{
	return batch_guard(batch, [&] {
		batch->batch.sync();
		return 0;
	});
}

void *pgs_batch_stream(pgs_batch *batch) { return batch ? static_cast<void *>(batch->batch.stream()) : nullptr; }

uint32_t pgs_batch_num_launches(pgs_batch *batch) { return batch ? batch->batch.num_launches() : 0; }

int pgs_batch_profile(pgs_batch *batch, double *ms, uint64_t *launches, uint32_t kinds)
{
	return batch_guard(batch, [&] {
		if (!ms || !launches)
			return -1;
		batch->batch.profile(ms, launches, kinds);
		return 0;
	});
}

pgs_cost_model *pgs_cost_model_create(uint64_t seed)
{
	auto *m = new (std::nothrow) pgs_cost_model;
	if (!m)
		return nullptr;
	try
	{
		m->model.init(seed);
		return m;
	}
	catch (const std::exception &e)
	{
		fprintf(stderr, "pgs: %s\n", e.what());
		delete m;
		return nullptr;
	}
}

void pgs_cost_model_destroy(pgs_cost_model *model) { delete model; }

int pgs_cost_model_set_params(pgs_cost_model *model, const float *params)
{
	return ptr_guard(model, [&] {
		if (!params)
			return -1;
		model->model.set_params(params);
		return 0;
	});
}

int pgs_cost_model_get_params(pgs_cost_model *model, float *params)
{
	return ptr_guard(model, [&] {
		if (!params)
			return -1;
		model->model.get_params(params);
		return 0;
	});
}

int pgs_cost_model_predict(pgs_cost_model *model, const float *features, uint32_t count, float *out)
{
	return ptr_guard(model, [&] {
		if (!features || !out)
			return -1;
		model->model.predict(features, count, out);
		return 0;
	});
}

int pgs_cost_model_gradients(pgs_cost_model *model, const float *features, const float *targets, uint32_t count,
                             float *grads)
{
	return ptr_guard(model, [&] {
		if (!features || !targets || !grads || count == 0)
			return -1;
		model->model.gradients(features, targets, count, grads);
		return 0;
	});
}

int pgs_cost_model_train(pgs_cost_model *model, const float *features, const float *targets, uint32_t count,
                         uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *mse)
{
	return ptr_guard(model, [&] {
		if (!features || !targets)
			return -1;
		const float e = model->model.train(features, targets, count, epochs, batch, lr, seed);
		if (mse)
			*mse = e;
		return 0;
	});
}

int pgs_batch_set_schedule(pgs_batch *batch, uint32_t schedule, const pgs_cost_model *model)
{
	return batch_guard(batch, [&] {
		if (schedule > PGS_SCHEDULE_DISPATCH_LEARNED)
			return -1;
		return batch->batch.set_schedule(pgs::Schedule(schedule), model ? &model->model : nullptr) ? 0 : -1;
	});
}

int pgs_batch_collect_costs(pgs_batch *batch, float *features, float *log2_costs, uint32_t max_items)
{
	return batch_guard(batch, [&] {
		if (!features || !log2_costs)
			return -1;
		return int(batch->batch.collect_costs(features, log2_costs, max_items));
	});
}
}
