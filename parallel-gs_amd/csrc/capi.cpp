// Context-level C ABI (include/pgs_amd.h): thin wrappers around GSInterface that never
// let a C++ exception cross the boundary.
#include "pgs_amd.h"
#include "gs_interface.hpp"
#include <exception>
#include <cstdio>
#include <new>

struct pgs_context
{
	pgs::GSInterface gs;
};

namespace
{
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
		ctx->gs.renderer().begin_recording();
	});
}

int pgs_end_recording(pgs_context *ctx)
{
	return guard(ctx, [&] {
		ctx->gs.flush();
		ctx->gs.renderer().end_recording();
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
