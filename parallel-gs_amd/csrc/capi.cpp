// Context-level C ABI (include/pgs_amd.h): thin wrappers around GSInterface that never
// let a C++ exception cross the boundary.
#include "pgs_amd.h"
#include "gs_interface.hpp"
#include "batch.hpp"
#include "cost_model.hpp"
#include "gs_dump.hpp"
#include <vector>
#include <cstring>
#include <exception>
#include <cstdio>
#include <new>
#include <stdexcept>

struct pgs_context
{
	pgs::GSInterface gs;
	uint64_t record_clock = 0; // VRAM write-log time at pgs_begin_recording
	pgs::ScanoutResult last_scanout;
};

struct pgs_dump_reader
{
	pgs_context *ctx;
	pgs::DumpReader reader;
};

struct pgs_dump_writer
{
	pgs_context *ctx;
	pgs::DumpWriter writer;
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

pgs::VSyncInfo to_vsync_info(const pgs_vsync_info &in)
{
	pgs::VSyncInfo v;
	v.phase = in.phase;
	v.force_progressive = in.force_progressive != 0;
	v.overscan = in.overscan != 0;
	v.anti_blur = in.anti_blur != 0;
	v.crtc_offsets = in.crtc_offsets != 0;
	v.adapt_to_internal_horizontal_resolution = in.adapt_to_internal_horizontal_resolution != 0;
	v.raw_circuit_scanout = in.raw_circuit_scanout != 0;
	v.high_resolution_scanout = in.high_resolution_scanout != 0;
	v.skip_deinterlace = in.skip_deinterlace != 0;
	return v;
}

void to_c_scanout(const pgs::ScanoutResult &r, pgs_scanout *out)
{
	memset(out, 0, sizeof(*out));
	out->image = r.image;
	out->width = r.width;
	out->height = r.height;
	out->internal_width = r.internal_width;
	out->internal_height = r.internal_height;
	out->mode_width = r.mode_width;
	out->mode_height = r.mode_height;
	out->high_resolution_scanout = r.high_resolution_scanout;
	out->interlaced = r.interlaced;
	out->double_strike = r.double_strike;
	out->interlace_phase = r.interlace_phase;
	out->sampling_rate_mhz = r.sampling_rate_mhz;
}

template <typename T, typename F>
int handle_guard(T *h, F &&f)
{
	if (!h)
		return PGS_ERROR_INVALID_ARGUMENT;
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
		// Sample counts other than 1/2/4/8/16 make init fail.
		o.super_sampling = pgs::SuperSampling(options->super_sampling ? options->super_sampling : 1u);
		o.ordered_super_sampling = options->ordered_super_sampling != 0;
		o.super_sampled_textures = options->super_sampled_textures != 0;
		o.dynamic_super_sampling = options->dynamic_super_sampling != 0;
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

int pgs_get_register_state(pgs_context *ctx, pgs_register_state *out)
{
	if (!out)
		return PGS_ERROR_INVALID_ARGUMENT;
	return guard(ctx, [&] {
		const pgs::RegisterState &r = ctx->gs.get_register_state();
		pgs_register_state s{};
		for (uint32_t c = 0; c < 2; c++)
		{
			pgs_context_registers &o = s.ctx[c];
			o.tex0 = r.tex0[c]; o.tex1 = r.tex1[c]; o.clamp = r.clamp[c]; o.xyoffset = r.xyoffset[c];
			o.miptbp1 = r.miptbp1[c]; o.miptbp2 = r.miptbp2[c]; o.scissor = r.scissor[c]; o.alpha = r.alpha[c];
			o.test = r.test[c]; o.fba = r.fba[c]; o.frame = r.frame[c]; o.zbuf = r.zbuf[c];
		}
		s.prim = r.prim; s.rgbaq = r.rgbaq; s.st = r.st; s.uv = r.uv; s.fog = r.fog; s.prmodecont = r.prmodecont;
		s.texclut = r.texclut; s.texa = r.texa; s.fogcol = r.fogcol; s.dimx = r.dimx; s.dthe = r.dthe;
		s.colclamp = r.colclamp; s.pabe = r.pabe; s.bitbltbuf = r.bitbltbuf; s.trxpos = r.trxpos;
		s.trxreg = r.trxreg; s.trxdir = r.trxdir; s.scanmsk = r.scanmsk;
		s.internal_q = r.internal_q;
		s.cached_cbp[0] = ctx->gs.cached_cbp()[0];
		s.cached_cbp[1] = ctx->gs.cached_cbp()[1];
		*out = s;
	});
}

int pgs_set_register_state(pgs_context *ctx, const pgs_register_state *in)
{
	if (!in)
		return PGS_ERROR_INVALID_ARGUMENT;
	return guard(ctx, [&] {
		pgs::RegisterState &r = ctx->gs.get_register_state();
		for (uint32_t c = 0; c < 2; c++)
		{
			const pgs_context_registers &o = in->ctx[c];
			r.tex0[c] = o.tex0; r.tex1[c] = o.tex1; r.clamp[c] = o.clamp; r.xyoffset[c] = o.xyoffset;
			r.miptbp1[c] = o.miptbp1; r.miptbp2[c] = o.miptbp2; r.scissor[c] = o.scissor; r.alpha[c] = o.alpha;
			r.test[c] = o.test; r.fba[c] = o.fba; r.frame[c] = o.frame; r.zbuf[c] = o.zbuf;
		}
		r.prim = in->prim; r.rgbaq = in->rgbaq; r.st = in->st; r.uv = in->uv; r.fog = in->fog;
		r.prmodecont = in->prmodecont; r.texclut = in->texclut; r.texa = in->texa; r.fogcol = in->fogcol;
		r.dimx = in->dimx; r.dthe = in->dthe; r.colclamp = in->colclamp; r.pabe = in->pabe;
		r.bitbltbuf = in->bitbltbuf; r.trxpos = in->trxpos; r.trxreg = in->trxreg; r.trxdir = in->trxdir;
		r.scanmsk = in->scanmsk;
		r.internal_q = in->internal_q;
		ctx->gs.cached_cbp()[0] = in->cached_cbp[0];
		ctx->gs.cached_cbp()[1] = in->cached_cbp[1];
	});
}

int pgs_get_gif_path(pgs_context *ctx, uint32_t path, pgs_gif_path *out)
{
	if (!out || path >= 4)
		return PGS_ERROR_INVALID_ARGUMENT;
	return guard(ctx, [&] {
		const pgs::GIFPath &p = ctx->gs.get_gif_path(path);
		out->tag[0] = p.tag_lo;
		out->tag[1] = p.regs;
		out->reg = p.reg;
		out->loop = p.loop;
	});
}

int pgs_set_gif_path(pgs_context *ctx, uint32_t path, const pgs_gif_path *in)
{
	if (!in || path >= 4)
		return PGS_ERROR_INVALID_ARGUMENT;
	return guard(ctx, [&] {
		pgs::GIFPath &p = ctx->gs.get_gif_path(path);
		p.tag_lo = in->tag[0];
		p.nloop = uint32_t(in->tag[0] & 0x7fff);
		p.flg = uint32_t((in->tag[0] >> 58) & 3u);
		p.nreg = uint32_t((in->tag[0] >> 60) & 0xf);
		p.regs = in->tag[1];
		p.reg = in->reg;
		p.loop = in->loop;
	});
}

int pgs_get_priv_register_state(pgs_context *ctx, void *out, size_t size)
{
	if (!out || size > PGS_PRIV_REGISTER_BYTES)
		return PGS_ERROR_INVALID_ARGUMENT;
	return guard(ctx, [&] { memcpy(out, ctx->gs.get_priv_register_state(), size); });
}

int pgs_set_priv_register_state(pgs_context *ctx, const void *in, size_t size)
{
	if (!in || size > PGS_PRIV_REGISTER_BYTES)
		return PGS_ERROR_INVALID_ARGUMENT;
	return guard(ctx, [&] { memcpy(ctx->gs.get_priv_register_state(), in, size); });
}

int pgs_set_super_sampling_rate(pgs_context *ctx, uint32_t super_sampling, uint32_t ordered_grid,
                                uint32_t super_sampled_textures)
{
	if (!ctx)
		return PGS_ERROR_INVALID_ARGUMENT;
	const bool known = super_sampling == 1 || super_sampling == 2 || super_sampling == 4 || super_sampling == 8 ||
	                   super_sampling == 16;
	if (!known)
		return PGS_ERROR_INVALID_ARGUMENT;
	bool ok = false;
	const int rc = guard(ctx, [&] {
		ok = ctx->gs.set_super_sampling_rate(pgs::SuperSampling(super_sampling), ordered_grid != 0,
		                                     super_sampled_textures != 0);
	});
	if (rc != 0)
		return rc;
	return ok ? 0 : PGS_ERROR_INVALID_ARGUMENT;
}

uint32_t pgs_get_super_sampling_rate(pgs_context *ctx) { return ctx ? ctx->gs.super_sampling_rate() : 0; }

int pgs_set_hacks(pgs_context *ctx, const pgs_hacks *hacks)
{
	if (!hacks)
		return PGS_ERROR_INVALID_ARGUMENT;
	const int rc = guard(ctx, [&] {
		pgs::Hacks h;
		h.disable_mipmaps = hacks->disable_mipmaps != 0;
		h.unsynced_readbacks = hacks->unsynced_readbacks != 0;
		ctx->gs.set_hacks(h);
	});
	if (rc != 0)
		return rc;
	return (hacks->backbuffer_promotion || hacks->allow_blend_demote) ? PGS_ERROR_UNSUPPORTED : 0;
}

int pgs_set_debug_mode(pgs_context *ctx, const pgs_debug_mode *mode)
{
	if (!mode)
		return PGS_ERROR_INVALID_ARGUMENT;
	const int rc = guard(ctx, [&] {
		ctx->gs.set_timestamps(mode->timestamps != 0);
		if (mode->disable_sampler_feedback)
			ctx->gs.set_sampler_feedback(2u);
	});
	if (rc != 0)
		return rc;
	return (mode->feedback_render_target || mode->draw_mode) ? PGS_ERROR_UNSUPPORTED : 0;
}

int pgs_set_signal_interface(pgs_context *ctx, const pgs_signal_interface *iface)
{
	return guard(ctx, [&] {
		pgs::SignalCallbacks cb;
		if (iface)
		{
			cb.on_signal = iface->on_signal;
			cb.on_finish = iface->on_finish;
			cb.on_label = iface->on_label;
			cb.user = iface->user;
		}
		ctx->gs.set_signal_interface(cb);
	});
}

double pgs_get_accumulated_timestamps(pgs_context *ctx, uint32_t type)
{
	if (!ctx || type >= PGS_TIMESTAMP_COUNT)
		return -1.0;
	double v = -1.0;
	if (guard(ctx, [&] { v = ctx->gs.get_accumulated_timestamps(pgs::TimestampType(type)); }) != 0)
		return -1.0;
	return v;
}


int pgs_vsync(pgs_context *ctx, const pgs_vsync_info *info, pgs_scanout *result)
{
	if (!ctx || !info || !result)
		return PGS_ERROR_INVALID_ARGUMENT;
	int rc = 0;
	const int g = guard(ctx, [&] {
		pgs::ScanoutResult r;
		rc = ctx->gs.vsync(to_vsync_info(*info), r);
		ctx->last_scanout = r;
		to_c_scanout(r, result);
	});
	if (g != 0)
		return g;
	return rc == pgs::GSInterface::kVsyncUnsupported ? PGS_ERROR_UNSUPPORTED : rc;
}

int pgs_vsync_can_skip(pgs_context *ctx, const pgs_vsync_info *info)
{
	if (!ctx || !info)
		return PGS_ERROR_INVALID_ARGUMENT;
	return ctx->gs.vsync_can_skip(to_vsync_info(*info)) ? 1 : 0;
}

int pgs_read_scanout(pgs_context *ctx, void *dst, size_t size)
{
	if (!ctx || !dst)
		return PGS_ERROR_INVALID_ARGUMENT;
	const pgs::ScanoutResult &r = ctx->last_scanout;
	const size_t bytes = size_t(r.width) * r.height * 4u;
	if (!r.image || size < bytes)
		return PGS_ERROR_INVALID_ARGUMENT;
	hipError_t e = hipStreamSynchronize(ctx->gs.renderer().stream());
	if (e == hipSuccess)
		e = hipMemcpy(dst, r.image, bytes, hipMemcpyDeviceToHost);
	return e == hipSuccess ? 0 : -int(e);
}


pgs_dump_reader *pgs_dump_open(pgs_context *ctx, const char *path, uint32_t raw)
{
	if (!ctx || !path)
		return nullptr;
	try
	{
		auto *r = new pgs_dump_reader{ ctx, {} };
		if (!r->reader.open(path, ctx->gs, raw != 0))
		{
			delete r;
			return nullptr;
		}
		return r;
	}
	catch (const std::exception &e)
	{
		fprintf(stderr, "pgs: %s\n", e.what());
		return nullptr;
	}
}

int pgs_dump_restart(pgs_dump_reader *r)
{
	return handle_guard(r, [&] { return r->reader.restart() ? 0 : -1; });
}

int pgs_dump_iterate_until_vsync(pgs_dump_reader *r, uint32_t high_res_scanout, uint32_t conservative_crtc)
{
	return handle_guard(r, [&] {
		const bool more = r->reader.iterate_until_vsync(high_res_scanout != 0, conservative_crtc != 0);
		const int rc = r->reader.vsync_status();
		if (rc != 0 && rc != pgs::GSInterface::kVsyncUnsupported)
			return rc;
		return more ? 1 : 0;
	});
}

int pgs_dump_consume_vsync_result(pgs_dump_reader *r, pgs_scanout *result)
{
	if (!result)
		return PGS_ERROR_INVALID_ARGUMENT;
	return handle_guard(r, [&] {
		const pgs::ScanoutResult s = r->reader.consume_vsync_result();
		r->ctx->last_scanout = s;
		to_c_scanout(s, result);
		return 0;
	});
}

void pgs_dump_close(pgs_dump_reader *r) { delete r; }

pgs_dump_writer *pgs_dump_writer_create(pgs_context *ctx, const char *path)
{
	if (!ctx || !path)
		return nullptr;
	try
	{
		auto *w = new pgs_dump_writer{ ctx, {} };
		if (!w->writer.init(path, ctx->gs))
		{
			delete w;
			return nullptr;
		}
		return w;
	}
	catch (const std::exception &e)
	{
		fprintf(stderr, "pgs: %s\n", e.what());
		return nullptr;
	}
}

int pgs_dump_write_register(pgs_dump_writer *w, uint32_t addr, uint64_t payload)
{
	return handle_guard(w, [&] { return w->writer.write_register(addr, payload) ? 0 : -1; });
}

int pgs_dump_write_packed(pgs_dump_writer *w, uint64_t prim, const uint32_t *registers, uint32_t num_registers,
                          uint32_t num_loops, const void *data)
{
	if (!registers || (num_loops && !data))
		return PGS_ERROR_INVALID_ARGUMENT;
	return handle_guard(w, [&] {
		return w->writer.write_packed(prim, registers, num_registers, num_loops, data) ? 0 : -1;
	});
}

int pgs_dump_write_vsync(pgs_dump_writer *w, uint32_t field)
{
	return handle_guard(w, [&] { return w->writer.write_vsync(field, w->ctx->gs) ? 0 : -1; });
}

int pgs_dump_write_image_upload(pgs_dump_writer *w, uint32_t addr, uint32_t psm, uint32_t width, uint32_t height,
                                const void *data, size_t size)
{
	if (!data && size)
		return PGS_ERROR_INVALID_ARGUMENT;
	return handle_guard(w, [&] {
		return w->writer.write_image_upload(addr, psm, width, height, data, size) ? 0 : -1;
	});
}

int pgs_dump_writer_close(pgs_dump_writer *w)
{
	if (!w)
		return PGS_ERROR_INVALID_ARGUMENT;
	const int rc = w->writer.close() ? 0 : -1;
	delete w;
	return rc;
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

int pgs_batch_schedule_stats(pgs_batch *batch, pgs_schedule_stats *stats)
{
	return batch_guard(batch, [&] {
		if (!stats)
			return PGS_ERROR_INVALID_ARGUMENT;
		pgs::Batch::ScheduleStats s;
		if (!batch->batch.schedule_stats(s))
			return PGS_ERROR_INVALID_ARGUMENT;
		stats->shading_launches = s.shading_launches;
		stats->tiles = s.tiles;
		stats->entries = s.entries;
		stats->split_tiles = s.split_tiles;
		stats->heavy_first = s.heavy_first;
		return 0;
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
