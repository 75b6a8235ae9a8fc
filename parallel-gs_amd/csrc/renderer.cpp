// Device-side renderer: HBM buffers, upload arena, transfer batching and launch recording.
#include "renderer.hpp"
#include <mutex>
#include "render_pass.hpp"
#include "raster_tables.hpp"
#include <algorithm>
#include <array>
#include <cstdio>
#include <stdexcept>

namespace pgs
{
namespace
{
constexpr size_t kAlign = 256;
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

// Upper bound on the number of 256-byte blocks a rectangle touches.
inline size_t rect_block_bound(uint32_t psm, uint32_t w, uint32_t h)
{
	const Geometry g = psm_geometry(psm);
	size_t bx = (size_t(w) >> g.bw) + 2, by = (size_t(h) >> g.bh) + 2;
	return bx * by;
}

// Arena bytes a copy's work lists can need at batch close: its descriptor plus, per touched
// block, one record and one item, and per source block one snapshot-list entry and one
// snapshot item.
inline size_t copy_meta_bound(const CopyDesc &d)
{
	size_t blocks = rect_block_bound(d.dpsm, d.w, d.h);
	size_t src_blocks = (d.flags & COPY_LOCAL) ? rect_block_bound(d.spsm, d.w, d.h) : 0;
	return sizeof(CopyDesc) + blocks * (sizeof(BlockRecord) + sizeof(CopyItem)) + src_blocks * (4 + sizeof(CopyItem)) +
	       6 * kAlign;
}
}

Renderer::~Renderer()
{
	if (side_)
		(void)hipStreamSynchronize(side_);
	if (stream_)
		(void)hipStreamSynchronize(stream_);
	destroy_graph();
	auto free_arena = [](Arena &a) {
		if (a.host) (void)hipHostFree(a.host);
		if (a.dev) (void)hipFree(a.dev);
		a = {};
	};
	free_arena(arena_);
	for (auto &a : retired_arenas_)
		free_arena(a);
	for (auto &t : timed_)
	{
		event_pool_.push_back(t.start);
		event_pool_.push_back(t.stop);
	}
	for (auto e : event_pool_)
		(void)hipEventDestroy(e);
	for (void *p : { (void *)setup_buf_, (void *)tattr_buf_, (void *)lists_buf_, (void *)counts_buf_, (void *)dummy_buf_,
	                 (void *)clut_buf_, (void *)tex_images_, (void *)state_copy_, (void *)flat_buckets_,
	                 (void *)tile_modes_, (void *)coarse_mask_, (void *)feedback_masks_, (void *)snapshot_,
	                 (void *)restore_dev_, (void *)fused_dev_, (void *)phase_lut_ })
		if (p) (void)hipFree(p);
	if (vram_) (void)hipFree(vram_);
	if (shadow_) (void)hipFree(shadow_);
	if (alias_cache_) (void)hipFree(alias_cache_);
	if (fork_event_) (void)hipEventDestroy(fork_event_);
	if (join_event_) (void)hipEventDestroy(join_event_);
	if (side_) (void)hipStreamDestroy(side_);
	if (stream_) (void)hipStreamDestroy(stream_);
}

void Renderer::check(hipError_t err)
{
	if (err != hipSuccess)
	{
		last_error_ = -int(err);
		fprintf(stderr, "pgs: HIP error %d (%s)\n", int(err), hipGetErrorString(err));
		throw std::runtime_error(hipGetErrorString(err));
	}
}

bool Renderer::init(uint32_t sample_log2, size_t arena_bytes, uint32_t sampler_feedback, bool texture_stream,
                    uint32_t capacity_log2)
{
	try
	{
		sample_log2_ = sample_log2;
		capacity_log2_ = std::max(capacity_log2, sample_log2);
		sampler_feedback_ = sampler_feedback;
		num_planes_ = sample_log2 ? 2u + (1u << sample_log2) : 1u;
		const uint32_t capacity_planes = capacity_log2_ ? 2u + (1u << capacity_log2_) : 1u;
		check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
		if (texture_stream)
		{
			check(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
			check(hipEventCreateWithFlags(&fork_event_, hipEventDisableTiming));
			check(hipEventCreateWithFlags(&join_event_, hipEventDisableTiming));
		}
		check(hipMalloc(&vram_, size_t(capacity_planes) * kVramSize));
		check(hipMemsetAsync(vram_, 0, size_t(capacity_planes) * kVramSize, stream_));
		check(hipMalloc(&shadow_, 2u * size_t(kVramSize))); // two snapshot planes, alternating by level
		arena_.cap = std::max<size_t>(arena_bytes, 16u << 20);
		arena_bytes_ = arena_.cap;
		check(hipHostMalloc(&arena_.host, arena_.cap, hipHostMallocDefault));
		check(hipMalloc(&arena_.dev, arena_.cap));
		level_written_.assign(kNumBlocks, -1);
		level_read_.assign(kNumBlocks, -1);
		check(hipMalloc(&setup_buf_, 65536u * sizeof(PrimitiveSetup)));
		check(hipMalloc(&tattr_buf_, 65536u * sizeof(TransformedAttributes)));
		lists_cap_ = 32u << 20;  // 64 MiB of uint16 entries
		counts_cap_ = 1u << 16;
		check(hipMalloc(&lists_buf_, lists_cap_ * sizeof(uint16_t)));
		check(hipMalloc(&counts_buf_, counts_cap_ * sizeof(uint32_t)));
		check(hipMalloc(&dummy_buf_, 256));
		check(hipMemsetAsync(dummy_buf_, 0, 256, stream_));
		check(hipMalloc(&clut_buf_, 512u * sizeof(uint16_t)));
		check(hipMemsetAsync(clut_buf_, 0, 512u * sizeof(uint16_t), stream_));
		check(hipMalloc(&tex_images_, size_t(kTexImageTexels) * sizeof(uint32_t)));
		check(hipMalloc(&state_copy_, 65536u * sizeof(uint32_t)));
		check(hipMalloc(&flat_buckets_, 8u * sizeof(uint32_t)));
		check(hipMemsetAsync(flat_buckets_, 0, 8u * sizeof(uint32_t), stream_));
		check(hipMalloc(&tile_modes_, counts_cap_ * sizeof(uint32_t)));
		check(hipMalloc(&coarse_mask_, size_t(kMaxCoarseTiles) * kMaxMaskWords * sizeof(uint64_t)));
		check(hipMalloc(&feedback_masks_, size_t(kMaxPassTextures) * PGS_FEEDBACK_WORDS * sizeof(uint32_t)));
		{
			float lut[256];
			make_rcp_float_lut(lut);
			int rc = pgs_k_init_tables(kRcpSeed, lut);
			if (rc != 0)
				check(hipError_t(-rc));
		}
		block_count_.assign(kNumBlocks, 0);
		check(hipStreamSynchronize(stream_));
		return true;
	}
	catch (const std::exception &)
	{
		return false;
	}
}

// ---------------------------------------------------------------------------------------
// Upload arena. Host writes go to pinned memory; arena_upload() copies the not yet uploaded
// range to the device mirror on the stream, ahead of the launches that read it.

uint8_t *Renderer::arena_alloc(size_t bytes, size_t &dev_offset)
{
	bytes = align_up(std::max<size_t>(bytes, 1), kAlign);
	if (!arena_.host)
	{
		// The arena handed to the last recording was detached (end_recording): work after the
		// recording gets a new one of the configured size.
		arena_.cap = std::max(arena_bytes_, align_up(bytes, 1u << 20));
		check(hipHostMalloc(&arena_.host, arena_.cap, hipHostMallocDefault));
		check(hipMalloc(&arena_.dev, arena_.cap));
	}
	if (arena_.used + bytes > arena_.cap)
		arena_make_room(bytes);
	dev_offset = arena_.used;
	uint8_t *p = arena_.host + arena_.used;
	arena_.used += bytes;
	return p;
}

void Renderer::arena_upload()
{
	if (arena_.used > arena_.uploaded)
	{
		check(hipMemcpyAsync(arena_.dev + arena_.uploaded, arena_.host + arena_.uploaded,
		                     arena_.used - arena_.uploaded, hipMemcpyHostToDevice, stream_));
		arena_.uploaded = arena_.used;
		main_pending_ = true;
	}
}

void Renderer::arena_make_room(size_t bytes)
{
	arena_upload();
	if (recording_)
	{
		// Keep the filled arena alive for replay and start a fresh one.
		retired_arenas_.push_back(arena_);
		arena_ = {};
		arena_.cap = std::max<size_t>(retired_arenas_.back().cap, align_up(bytes, 1u << 20));
		check(hipHostMalloc(&arena_.host, arena_.cap, hipHostMallocDefault));
		check(hipMalloc(&arena_.dev, arena_.cap));
		return;
	}

	sync_streams();
	stats_.syncs++;
	if (bytes > arena_.cap)
	{
		(void)hipHostFree(arena_.host);
		(void)hipFree(arena_.dev);
		arena_.cap = align_up(bytes, 1u << 20);
		check(hipHostMalloc(&arena_.host, arena_.cap, hipHostMallocDefault));
		check(hipMalloc(&arena_.dev, arena_.cap));
	}
	arena_.used = 0;
	arena_.uploaded = 0;
}

void Renderer::arena_reset_after_sync()
{
	if (!recording_)
	{
		arena_.used = 0;
		arena_.uploaded = 0;
	}
}

void Renderer::sync()
{
	if (!batch_.empty())
		flush_transfers();
	arena_upload();
	sync_streams();
	stats_.syncs++;
	arena_reset_after_sync();
}

void Renderer::sync_streams()
{
	join_texture_stream();
	check(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------------------------------
// Launches. Every launch goes through the kernel-level C ABI.

hipEvent_t Renderer::get_event()
{
	if (!event_pool_.empty())
	{
		hipEvent_t e = event_pool_.back();
		event_pool_.pop_back();
		return e;
	}
	hipEvent_t e;
	check(hipEventCreate(&e));
	return e;
}

void Renderer::profile_enable(bool enable)
{
	profiling_ = enable;
}

void Renderer::profile_collect(double *ms, uint64_t *launches, uint32_t kinds)
{
	sync_streams();
	for (uint32_t i = 0; i < kinds; i++)
	{
		ms[i] = 0.0;
		launches[i] = 0;
	}
	for (auto &t : timed_)
	{
		float e = 0.0f;
		check(hipEventElapsedTime(&e, t.start, t.stop));
		if (t.kind < kinds)
		{
			ms[t.kind] += e;
			launches[t.kind]++;
		}
		event_pool_.push_back(t.start);
		event_pool_.push_back(t.stop);
	}
	timed_.clear();
}

bool Renderer::on_texture_stream(const Launch &l) const
{
	switch (l.kind)
	{
	case LaunchKind::VramCopy:
	case LaunchKind::VramSnapshot:
	case LaunchKind::ClutWrite: return true;
	// With sampler feedback the decode reads the masks the feedback launch writes after setup.
	case LaunchKind::TexDecode: return l.decode.feedback == nullptr;
	default: return false;
	}
}

void Renderer::launch(const Launch &l)
{
	if (side_)
	{
		if (on_texture_stream(l))
		{
			if (main_pending_)
				issue(Launch{ LaunchKind::Fork, {} });
		}
		else if (side_pending_ && l.kind != LaunchKind::Setup && l.kind != LaunchKind::Binning &&
		         l.kind != LaunchKind::SsaaHeuristic && l.kind != LaunchKind::SamplerFeedback)
			issue(Launch{ LaunchKind::Join, {} });
	}
	issue(l);
}

void Renderer::join_texture_stream()
{
	if (side_ && side_pending_)
		issue(Launch{ LaunchKind::Join, {} });
}

void Renderer::issue(const Launch &l)
{
	if (recording_)
		replay_.push_back(l);
	if (l.kind == LaunchKind::Fork || l.kind == LaunchKind::Join)
	{
		const bool fork = l.kind == LaunchKind::Fork;
		if (side_)
		{
			hipEvent_t e = fork ? fork_event_ : join_event_;
			check(hipEventRecord(e, fork ? stream_ : side_));
			check(hipStreamWaitEvent(fork ? side_ : stream_, e, 0));
		}
		(fork ? main_pending_ : side_pending_) = false;
		return;
	}
	if (recording_)
		recorded_.push_back(l);
	const bool side = side_ && on_texture_stream(l);
	hipStream_t s = side ? side_ : stream_;
	TimedLaunch t{ uint32_t(l.kind), nullptr, nullptr };
	if (profiling_)
	{
		t.start = get_event();
		t.stop = get_event();
		check(hipEventRecord(t.start, s));
	}
	int rc = 0;
	switch (l.kind)
	{
	case LaunchKind::VramCopy: rc = pgs_k_vram_copy(&l.copy, s); break;
	case LaunchKind::VramSnapshot: rc = pgs_k_vram_snapshot(&l.snapshot, s); break;
	case LaunchKind::Setup: rc = pgs_k_triangle_setup(&l.setup, s); break;
	case LaunchKind::Binning: rc = pgs_k_binning(&l.bin, s); break;
	case LaunchKind::Shade: rc = pgs_k_shade(&l.shade, s); break;
	case LaunchKind::ClutWrite: rc = pgs_k_clut_write(&l.clut, s); break;
	case LaunchKind::TexDecode: rc = pgs_k_texture_decode(&l.decode, s); break;
	case LaunchKind::SsaaHeuristic: rc = pgs_k_ssaa_heuristic(&l.heuristic, s); break;
	case LaunchKind::ShadeSuper: rc = pgs_k_shade(&l.shade, s); break;
	case LaunchKind::SamplerFeedback: rc = pgs_k_sampler_feedback(&l.feedback, s); break;
	case LaunchKind::Fused: rc = pgs_k_fused(&l.fused, s); break;
	case LaunchKind::Fork:
	case LaunchKind::Join: break;
	}
	if (rc != 0)
		check(hipError_t(-rc));
	(side ? side_pending_ : main_pending_) = true;
	if (profiling_)
	{
		check(hipEventRecord(t.stop, s));
		timed_.push_back(t);
	}
}

// ---------------------------------------------------------------------------------------
// Transfer batching.
//
// Copies queue until something needs VRAM (a render pass, a CLUT load, a readback, a full
// arena); then each copy gets a dependency level and every level runs as one launch. The
// reference flushes a batch whenever a local copy reads what the batch wrote
// (page_tracker.cpp:221-245, at 8 KiB page granularity); levels at the 256-byte block
// granularity give one launch per step of the read-after-write chain instead, and copies
// that do not depend on that chain join earlier launches. Inside a level, write-after-write
// is resolved by the per-block ordered kernel, and local copies whose source the level
// overwrites read a snapshot taken at level start.

void Renderer::queue_host_copy(const CopyDesc &desc_in, const void *payload, size_t bytes)
{
	CopyDesc desc = desc_in;
	size_t meta = copy_meta_bound(desc);
	size_t need = align_up(bytes, kAlign) + meta;
	if (arena_.used + batch_meta_ + need > arena_.cap)
	{
		if (!batch_.empty())
			flush_transfers();
		arena_make_room(need);
	}

	size_t off;
	uint8_t *dst = arena_alloc(bytes, off);
	memcpy(dst, payload, bytes);
	desc.payload_offset = off;
	desc.payload_bytes = uint32_t(bytes);
	desc.flags &= ~COPY_LOCAL;

	batch_.push_back({ desc });
	batch_meta_ += copy_meta_bound(desc);
}

void Renderer::queue_local_copy(const CopyDesc &desc_in)
{
	CopyDesc desc = desc_in;
	desc.flags |= COPY_LOCAL;
	desc.payload_bytes = 0;
	desc.payload_offset = 0;
	desc.first_pixel = 0;
	desc.end_pixel = desc.w * desc.h;

	size_t meta = copy_meta_bound(desc);
	if (arena_.used + batch_meta_ + meta > arena_.cap)
	{
		if (!batch_.empty())
			flush_transfers();
		arena_make_room(meta);
	}

	batch_.push_back({ desc });
	batch_meta_ += copy_meta_bound(desc);
}

void Renderer::flush_transfers()
{
	if (batch_.empty())
		return;

	// Level of every queued copy, in queue order: after the levels of the copies that last
	// wrote its source blocks (read-after-write), and not before the levels of the copies
	// that last wrote or read its destination blocks (write-after-write, write-after-read:
	// inside one level destination blocks apply copies in queue order and sources that the
	// level overwrites are read from the level-start snapshot).
	std::fill(level_written_.begin(), level_written_.end(), -1);
	std::fill(level_read_.begin(), level_read_.end(), -1);
	std::vector<uint32_t> level(batch_.size());
	uint32_t levels = 0;
	for (size_t j = 0; j < batch_.size(); j++)
	{
		const CopyDesc &d = batch_[j].desc;
		const bool local = (d.flags & COPY_LOCAL) != 0;
		int32_t L = 0;
		if (local)
			for_each_rect_block(d.spsm, d.sbp, d.sbw, d.ssax, d.ssay, d.w, d.h, [&](uint32_t b, uint32_t, uint32_t) {
				L = std::max(L, level_written_[b] + 1);
			});
		for_each_rect_block(d.dpsm, d.dbp, d.dbw, d.dsax, d.dsay, d.w, d.h, [&](uint32_t b, uint32_t, uint32_t) {
			L = std::max(L, std::max(level_written_[b], level_read_[b]));
		});
		for_each_rect_block(d.dpsm, d.dbp, d.dbw, d.dsax, d.dsay, d.w, d.h,
		                    [&](uint32_t b, uint32_t, uint32_t) { level_written_[b] = L; });
		if (local)
			for_each_rect_block(d.spsm, d.sbp, d.sbw, d.ssax, d.ssay, d.w, d.h, [&](uint32_t b, uint32_t, uint32_t) {
				level_read_[b] = std::max(level_read_[b], L);
			});
		level[j] = uint32_t(L);
		levels = std::max(levels, uint32_t(L) + 1u);
	}
	std::vector<std::vector<uint32_t>> members(levels);
	for (uint32_t j = 0; j < uint32_t(batch_.size()); j++)
		members[level[j]].push_back(j);
	std::vector<TransferLevel> plan(levels);
	for (uint32_t i = 0; i < levels; i++)
		plan_transfer_level(members[i], plan[i]);
	for (uint32_t i = 0; i < levels; i++)
		emit_transfer_level(plan[i], i, i + 1u < levels ? &plan[i + 1u] : nullptr);
	batch_.clear();
	batch_meta_ = 0;
}

// The blocks a level writes, and its local copies whose source the level also overwrites:
// those read the level-start snapshot, whose blocks are listed.
void Renderer::plan_transfer_level(const std::vector<uint32_t> &members, TransferLevel &lv)
{
	lv.written.assign(kNumBlocks / 64, 0);
	lv.descs.resize(members.size());
	for (size_t ci = 0; ci < members.size(); ci++)
	{
		lv.descs[ci] = batch_[members[ci]].desc;
		const CopyDesc &d = lv.descs[ci];
		for_each_rect_block(d.dpsm, d.dbp, d.dbw, d.dsax, d.dsay, d.w, d.h,
		                    [&](uint32_t b, uint32_t, uint32_t) { lv.written[b >> 6] |= 1ull << (b & 63); });
	}
	std::vector<uint64_t> snap_mask(kNumBlocks / 64, 0);
	for (CopyDesc &d : lv.descs)
	{
		if (!(d.flags & COPY_LOCAL))
			continue;
		bool hazard = false;
		for_each_rect_block(d.spsm, d.sbp, d.sbw, d.ssax, d.ssay, d.w, d.h, [&](uint32_t b, uint32_t, uint32_t) {
			if (lv.written[b >> 6] & (1ull << (b & 63)))
				hazard = true;
		});
		if (!hazard)
			continue;
		d.flags |= COPY_SOURCE_SHADOW;
		for_each_rect_block(d.spsm, d.sbp, d.sbw, d.ssax, d.ssay, d.w, d.h, [&](uint32_t b, uint32_t, uint32_t) {
			const uint64_t bit = 1ull << (b & 63);
			if (!(snap_mask[b >> 6] & bit))
			{
				snap_mask[b >> 6] |= bit;
				lv.snapshot.push_back(b);
			}
		});
	}
}

// One level of queued copies as one launch. Level i reads its snapshot from shadow plane i & 1;
// the launch of level i - 1 fills it: blocks that level writes store their result there too,
// and the others are copied by extra waves of the same launch (the level does not write them,
// so their value at its end is their value while it runs). Only a first level with a snapshot
// needs a snapshot launch of its own.
void Renderer::emit_transfer_level(const TransferLevel &lv, uint32_t index, const TransferLevel *next)
{
	const uint32_t ncopies = uint32_t(lv.descs.size());
	uint8_t *shadow = shadow_ + size_t(index & 1u) * kVramSize;
	uint8_t *shadow_next = shadow_ + size_t((index + 1u) & 1u) * kVramSize;
	const bool has_next = next && !next->snapshot.empty();

	// Count records per destination block.
	size_t nrec = 0;
	for (const CopyDesc &d : lv.descs)
		for_each_rect_block(d.dpsm, d.dbp, d.dbw, d.dsax, d.dsay, d.w, d.h, [&](uint32_t b, uint32_t, uint32_t) {
			if (block_count_[b]++ == 0)
				touched_blocks_.push_back(b);
			nrec++;
		});
	std::sort(touched_blocks_.begin(), touched_blocks_.end());

	// The next level's snapshot: blocks this level writes keep a flag, the rest are copied.
	std::vector<uint32_t> snap_copy;
	std::vector<uint64_t> next_mask;
	if (has_next)
	{
		next_mask.assign(kNumBlocks / 64, 0);
		for (uint32_t b : next->snapshot)
		{
			next_mask[b >> 6] |= 1ull << (b & 63);
			if (!(lv.written[b >> 6] & (1ull << (b & 63))))
				snap_copy.push_back(b);
		}
	}

	// Items: one per destination block with its first record inline, then the snapshot copies.
	// Further records follow in `recs`, stably ordered (copy order, then raster order).
	const size_t nblocks = touched_blocks_.size(), nitems = nblocks + snap_copy.size();
	const size_t nextra = nrec - nblocks;
	size_t off_desc, off_rec, off_item, off_first = 0;
	auto *descs = reinterpret_cast<CopyDesc *>(arena_alloc(ncopies * sizeof(CopyDesc), off_desc));
	auto *recs = reinterpret_cast<BlockRecord *>(arena_alloc(nextra * sizeof(BlockRecord), off_rec));
	auto *items = reinterpret_cast<CopyItem *>(arena_alloc(nitems * sizeof(CopyItem), off_item));
	uint32_t *first_snap = nullptr;
	if (index == 0 && !lv.snapshot.empty())
		first_snap = reinterpret_cast<uint32_t *>(arena_alloc(lv.snapshot.size() * 4, off_first));

	// block_count_ becomes the index of the block's item (cursor = records written so far in
	// item_fill_, kept in the item's rec_count until the scatter ends).
	uint32_t run = 0;
	for (size_t i = 0; i < nblocks; i++)
	{
		const uint32_t b = touched_blocks_[i];
		const bool to_next = has_next && (next_mask[b >> 6] & (1ull << (b & 63)));
		const uint32_t extra = block_count_[b] - 1u;
		items[i] = { b, to_next ? COPY_ITEM_TO_NEXT_SNAPSHOT : 0u, 0u, 0u, run, 0u, { 0u, 0u } };
		block_count_[b] = uint32_t(i);
		run += extra;
	}
	for (uint32_t ci = 0; ci < ncopies; ci++)
	{
		const CopyDesc &d = lv.descs[ci];
		descs[ci] = d;
		for_each_rect_block(d.dpsm, d.dbp, d.dbw, d.dsax, d.dsay, d.w, d.h, [&](uint32_t b, uint32_t x0, uint32_t y0) {
			CopyItem &it = items[block_count_[b]];
			if (it.pad[0] == 0u)
			{
				it.copy = ci;
				it.xy = x0 | (y0 << 16);
				it.pad[0] = 1u;
			}
			else
				recs[it.rec_begin + it.rec_count++] = { ci, uint16_t(x0), uint16_t(y0) };
		});
	}
	for (size_t i = 0; i < nblocks; i++)
	{
		items[i].pad[0] = 0u;
		block_count_[touched_blocks_[i]] = 0;
	}
	for (size_t i = 0; i < snap_copy.size(); i++)
		items[nblocks + i] = { snap_copy[i], COPY_ITEM_SNAPSHOT, 0u, 0u, 0u, 0u, { 0u, 0u } };
	if (first_snap)
		memcpy(first_snap, lv.snapshot.data(), lv.snapshot.size() * 4);

	arena_upload();

	if (first_snap)
	{
		Launch l{};
		l.kind = LaunchKind::VramSnapshot;
		l.snapshot = { vram_, shadow, reinterpret_cast<const uint32_t *>(arena_.dev + off_first),
		               uint32_t(lv.snapshot.size()), 0 };
		launch(l);
	}

	Launch l{};
	l.kind = LaunchKind::VramCopy;
	l.copy.vram = vram_;
	l.copy.ref_plane = sample_log2_ ? vram_ + kVramSize : nullptr;
	l.copy.shadow = shadow;
	l.copy.shadow_next = has_next ? shadow_next : nullptr;
	l.copy.payload = arena_.dev;
	l.copy.descs = arena_.dev + off_desc;
	l.copy.records = arena_.dev + off_rec;
	l.copy.items = arena_.dev + off_item;
	l.copy.num_items = uint32_t(nitems);
	launch(l);

	stats_.transfer_batches++;
	stats_.copies += ncopies;
	stats_.copy_blocks += touched_blocks_.size();
	stats_.copy_records += nrec;
	stats_.snapshot_blocks += lv.snapshot.size();
	for (const CopyDesc &d : lv.descs)
	{
		uint64_t pixels = (d.flags & COPY_LOCAL) ? uint64_t(d.w) * d.h : uint64_t(d.end_pixel - d.first_pixel);
		stats_.copy_bytes += 2u * pixels * psm_bpp(d.dpsm) / 8u;
	}
	touched_blocks_.clear();
}

// ---------------------------------------------------------------------------------------
// Render passes.

void Renderer::submit_render_pass(const RenderPassDesc &d)
{
	flush_transfers();
	const uint32_t n = d.num_primitives;
	if (n == 0 || n > 65536u)
		return;

	// Sampling rate of the pass (gs/gs_interface.cpp:212-260): channel-sliced 16-bit colour
	// masks, depth read back as a texture on tile and colour/depth aliasing with depth writes
	// render single-sampled.
	const bool ssaa_context = num_planes_ > 1;
	uint32_t rate_x, rate_y;
	rates(rate_x, rate_y);
	if (ssaa_context)
	{
		const bool fb_32_24 = d.fb_psm == PSMCT32 || d.fb_psm == PSMCT24 || d.fb_psm == PSMZ32 || d.fb_psm == PSMZ24;
		const uint32_t m = ~d.opaque_fbmask & 0xf8f8f8u;
		const bool channel_slice = !fb_32_24 && m != 0xf8f8f8u && m != 0;
		const bool aliased = d.z_write && d.zbp == d.fbp;
		// A pass sampling super-sampled textures stays super-sampled.
		if (!d.arrays && (channel_slice || d.z_feedback || aliased))
			rate_x = rate_y = 0;
	}
	const bool super_pass = rate_y != 0;

	// Tile grid. Single-sampled passes may pick any grid (it cannot change a pixel): the pass
	// box rounded down to whole 8x8 VRAM blocks. Super-sampled passes decide super- vs
	// single-sampled shading per coarse tile, so they use the reference's grid: origin at the
	// pass box and the binning-cost tile size (gs/gs_interface.cpp:148-187, 214-215), computed
	// in the same 32-bit arithmetic.
	uint32_t tile_log2 = tile_log2_1x_;
	int32_t ox = d.bb[0] & ~7, oy = d.bb[1] & ~7;
	if (super_pass)
	{
		ox = d.bb[0];
		oy = d.bb[1];
		const uint32_t cost = (uint32_t((d.bb[2] - d.bb[0]) >> 3) + 1u) * (uint32_t((d.bb[3] - d.bb[1]) >> 3) + 1u) * n;
		tile_log2 = cost < 10000u ? 3u : cost < 10000000u ? 4u : cost < 100000000u ? 5u : 6u;
		if (tile_log2 > 3)
			tile_log2--;
		const uint32_t cw = uint32_t((d.bb[2] - d.bb[0]) >> tile_log2) + 1u;
		const uint32_t ch = uint32_t((d.bb[3] - d.bb[1]) >> tile_log2) + 1u;
		uint64_t list_bytes = uint64_t(uint32_t(cw * ch * n)) * 2u;
		while (list_bytes > 48ull * 1024 * 1024)
		{
			tile_log2++;
			list_bytes /= 4;
		}
	}
	auto tiles_for = [&](uint32_t tl, uint32_t &tx, uint32_t &ty) {
		tx = uint32_t((d.bb[2] - ox) >> tl) + 1u;
		ty = uint32_t((d.bb[3] - oy) >> tl) + 1u;
	};
	uint32_t tiles_x, tiles_y;
	tiles_for(tile_log2, tiles_x, tiles_y);
	while ((size_t(tiles_x) * tiles_y * n > lists_cap_ || size_t(tiles_x) * tiles_y > counts_cap_) && tile_log2 < 7)
	{
		if (super_pass)
		{
			// The reference's grid always fits: 48 MiB of lists and at most 256 x 256 tiles.
			last_error_ = -int(hipErrorInvalidValue);
			return;
		}
		tile_log2++;
		tiles_for(tile_log2, tiles_x, tiles_y);
	}

	// First binning level: coarse tiles of 4 x 4 tiles, larger where the grid would need more
	// than kMaxCoarseTiles of them.
	uint32_t coarse_shift = 2;
	while (((((tiles_x - 1u) >> coarse_shift) + 1u) * (((tiles_y - 1u) >> coarse_shift) + 1u)) > kMaxCoarseTiles)
		coarse_shift++;
	const uint32_t coarse_x = ((tiles_x - 1u) >> coarse_shift) + 1u, coarse_y = ((tiles_y - 1u) >> coarse_shift) + 1u;
	const uint32_t mask_words = (n + 63u) / 64u;

	const size_t bytes_pos = size_t(n) * 3u * sizeof(VertexPosition);
	const size_t bytes_attr = size_t(n) * 3u * sizeof(VertexAttribute);
	const size_t bytes_prim = size_t(n) * sizeof(PrimitiveAttribute);
	const size_t bytes_state = size_t(d.num_states) * sizeof(StateVector);
	const size_t bytes_dec = size_t(d.num_textures) * sizeof(TexDecode);
	const size_t bytes_tex = size_t(d.num_textures) * sizeof(TextureDesc);
	// A pass whose depth buffer is its colour buffer (gs/gs_renderer.cpp:2525-2542) shades in
	// several launches; those that must not see colour written earlier in the same launch read
	// depth (and a single-stepped primitive also colour) from a copy of the pass's pages. The
	// copy's block list (every plane the shading kernel reads) goes up with the pass.
	const bool fb_z_alias = d.z_sensitive && d.fbp == d.zbp;
	const uint32_t cache_planes = super_pass ? num_planes_ : 1u;
	std::vector<uint32_t> cache_pages;
	if (fb_z_alias)
	{
		const PageSpan cs = page_span(d.fbp * 32u, d.fbw, d.fb_psm, uint32_t(d.bb[0]), uint32_t(d.bb[1]),
		                              uint32_t(d.bb[2]), uint32_t(d.bb[3]));
		const PageSpan zs = page_span(d.zbp * 32u, d.fbw, d.zbuf_psm, uint32_t(d.bb[0]), uint32_t(d.bb[1]),
		                              uint32_t(d.bb[2]), uint32_t(d.bb[3]));
		const uint32_t first = std::min(cs.first, zs.first), last = std::max(cs.last, zs.last);
		const uint32_t pages = std::min(last - first + 1u, kNumPages);
		for (uint32_t k = 0; k < pages; k++)
			cache_pages.push_back((first + k) % kNumPages);
	}
	const size_t bytes_cache = cache_pages.size() * 32u * cache_planes * sizeof(uint32_t);
	const size_t need = bytes_pos + bytes_attr + bytes_prim + bytes_state + bytes_dec + bytes_tex + bytes_cache +
	                    7 * kAlign;
	if (arena_.used + need > arena_.cap)
		arena_make_room(need);
	size_t off_pos, off_attr, off_prim, off_state;
	memcpy(arena_alloc(bytes_pos, off_pos), d.positions, bytes_pos);
	memcpy(arena_alloc(bytes_attr, off_attr), d.attributes, bytes_attr);
	memcpy(arena_alloc(bytes_prim, off_prim), d.prims, bytes_prim);
	memcpy(arena_alloc(bytes_state, off_state), d.states, bytes_state);
	size_t off_dec = 0, off_tex = 0;
	if (d.num_textures)
	{
		memcpy(arena_alloc(bytes_dec, off_dec), d.decodes, bytes_dec);
		memcpy(arena_alloc(bytes_tex, off_tex), d.tex_descs, bytes_tex);
	}
	size_t off_cache = 0;
	if (bytes_cache)
	{
		uint32_t *blocks = reinterpret_cast<uint32_t *>(arena_alloc(bytes_cache, off_cache));
		for (uint32_t plane = 0; plane < cache_planes; plane++)
			for (uint32_t page : cache_pages)
				for (uint32_t b = 0; b < 32u; b++)
					*blocks++ = plane * kNumBlocks + page * 32u + b;
	}
	arena_upload();

	// Sparse texture decode (sampler feedback, chosen by the pass builder): setup clears the
	// masks, the feedback launch marks the 8x8 blocks the primitives can sample, the decode
	// skips the others.
	const bool feedback = d.num_textures != 0 && d.sparse_decode;
	// A full decode depends only on VRAM and the CLUT: issued first, it runs on the texture
	// stream beside setup, the heuristic and binning. A sparse decode follows the feedback.
	auto decode = [&]() {
		Launch ld{};
		ld.kind = LaunchKind::TexDecode;
		ld.decode.vram = vram_;
		ld.decode.clut = clut_buf_;
		ld.decode.images = tex_images_;
		ld.decode.decodes = arena_.dev + off_dec;
		ld.decode.feedback = feedback ? feedback_masks_ : nullptr;
		ld.decode.num_textures = d.num_textures;
		ld.decode.num_groups = d.tex_groups;
		launch(ld);
	};
	if (d.num_textures && !feedback)
		decode();

	Launch ls{};
	ls.kind = LaunchKind::Setup;
	ls.setup.positions = arena_.dev + off_pos;
	ls.setup.attributes = arena_.dev + off_attr;
	ls.setup.prims = arena_.dev + off_prim;
	ls.setup.setup = setup_buf_;
	ls.setup.tattr = tattr_buf_;
	ls.setup.num_primitives = n;
	ls.setup.rate_log2 = rate_y;
	ls.setup.texture_flags = (d.arrays && super_pass ? 1u : 0u) | rate_x << 8;
	ls.setup.opaque_fbmask = d.opaque_fbmask;
	// Depth format of the single-sample heuristic (gs/gs_renderer.cpp:1918-1940, 2001-2031);
	// 24-bit when the pass does not use depth.
	const uint32_t depth_psm = d.z_sensitive ? d.zbuf_psm : uint32_t(PSMZ24);
	if (super_pass)
	{
		ls.setup.state_copy = state_copy_;
		ls.setup.flat_buckets = flat_buckets_;
		ls.setup.z_bucket_shift = (depth_psm == PSMZ16 || depth_psm == PSMZ16S) ? 8u : depth_psm == PSMZ32 ? 24u : 16u;
	}
	ls.setup.coarse_mask = coarse_mask_;
	ls.setup.coarse_origin_x = ox;
	ls.setup.coarse_origin_y = oy;
	ls.setup.coarse_log2 = tile_log2 + coarse_shift;
	ls.setup.coarse_x = coarse_x;
	ls.setup.coarse_y = coarse_y;
	ls.setup.mask_words = mask_words;
	if (feedback)
	{
		ls.setup.clear_words = feedback_masks_;
		ls.setup.clear_count = d.num_textures * PGS_FEEDBACK_WORDS;
	}
	launch(ls);

	if (feedback)
	{
		Launch lf{};
		lf.kind = LaunchKind::SamplerFeedback;
		lf.feedback.setup = setup_buf_;
		lf.feedback.prims = arena_.dev + off_prim;
		lf.feedback.tattr = tattr_buf_;
		lf.feedback.states = arena_.dev + off_state;
		lf.feedback.textures = arena_.dev + off_tex;
		lf.feedback.masks = feedback_masks_;
		lf.feedback.num_primitives = n;
		lf.feedback.num_textures = d.num_textures;
		launch(lf);
		decode();
	}

	if (super_pass)
	{
		Launch lh{};
		lh.kind = LaunchKind::SsaaHeuristic;
		lh.heuristic.positions = arena_.dev + off_pos;
		lh.heuristic.prims = arena_.dev + off_prim;
		lh.heuristic.state_copy = state_copy_;
		lh.heuristic.flat_buckets = flat_buckets_;
		lh.heuristic.num_primitives = n;
		if (depth_psm == PSMZ16 || depth_psm == PSMZ16S)
		{
			lh.heuristic.z_shift_bucket = 8; lh.heuristic.z_shift = 0; lh.heuristic.z_max = 0xffffu;
		}
		else if (depth_psm == PSMZ32)
		{
			lh.heuristic.z_shift_bucket = 16; lh.heuristic.z_shift = 8; lh.heuristic.z_max = ~0u;
		}
		else
		{
			lh.heuristic.z_shift_bucket = 16; lh.heuristic.z_shift = 0; lh.heuristic.z_max = 0xffffffu;
		}
		launch(lh);
	}

	Launch lb{};
	lb.kind = LaunchKind::Binning;
	lb.bin.setup = setup_buf_;
	lb.bin.prims = arena_.dev + off_prim;
	lb.bin.lists = lists_buf_;
	lb.bin.counts = counts_buf_;
	lb.bin.num_primitives = n;
	lb.bin.list_stride = n;
	lb.bin.origin_x = ox;
	lb.bin.origin_y = oy;
	lb.bin.tile_log2 = tile_log2;
	lb.bin.tiles_x = tiles_x;
	lb.bin.tiles_y = tiles_y;
	lb.bin.super_sampled = super_pass ? (d.arrays ? 3u : 1u) : 0u;
	lb.bin.tile_modes = super_pass ? tile_modes_ : nullptr;
	lb.bin.coarse_mask = coarse_mask_;
	lb.bin.coarse_shift = coarse_shift;
	lb.bin.coarse_x = coarse_x;
	lb.bin.coarse_y = coarse_y;
	lb.bin.mask_words = mask_words;
	launch(lb);

	Launch lsh{};
	lsh.kind = LaunchKind::Shade;
	pgs_shade_args &a = lsh.shade;
	a.vram = vram_;
	a.setup = setup_buf_;
	a.tattr = tattr_buf_;
	a.prims = arena_.dev + off_prim;
	a.states = arena_.dev + off_state;
	a.lists = lists_buf_;
	a.counts = counts_buf_;
	a.textures = d.num_textures ? static_cast<const void *>(arena_.dev + off_tex) : dummy_buf_;
	a.tex_images = d.num_textures ? tex_images_ : dummy_buf_;
	a.clut = reinterpret_cast<const uint16_t *>(clut_buf_);
	a.list_stride = n;
	a.origin_x = ox;
	a.origin_y = oy;
	a.tile_log2 = tile_log2;
	a.tiles_x = tiles_x;
	a.tiles_y = tiles_y;
	a.blocks_x = uint32_t((d.bb[2] - ox) >> 3) + 1u;
	a.blocks_y = uint32_t((d.bb[3] - oy) >> 3) + 1u;
	a.fb_psm = d.fb_psm;
	a.fbp = d.fbp;
	a.fbw = d.fbw;
	a.z_psm = d.z_psm;
	a.zbp = d.zbp;
	a.rate_log2 = 0;
	a.rate_x_log2 = 0;
	a.tile_modes = super_pass ? tile_modes_ : nullptr;
	a.has_ref_plane = ssaa_context ? 1u : 0u;
	if (super_pass)
	{
		// One launch for both kinds of tile: the first sample wave of each block shades the
		// single-sampled tiles.
		lsh.kind = LaunchKind::ShadeSuper;
		a.rate_log2 = rate_y;
		a.rate_x_log2 = rate_x;
		a.single_tiles = 1;
		if (d.arrays)
			a.phase_lut = phase_lut(rate_x, rate_y);
	}
	if (!fb_z_alias)
		launch(lsh);
	else
		shade_aliased_pass(d, lsh, arena_.dev + off_cache,
		                   uint32_t(cache_pages.size()) * 32u * cache_planes);

	stats_.render_passes++;
	stats_.primitives += n;
}

// Phase filter table of a sampling grid for super-sampled textures (gs/gs_renderer.cpp:189-313):
// built once per process and grid on the host, uploaded once per context and grid into its own
// slot, so that recordings keep reading the table of their grid.
const uint32_t *Renderer::phase_lut(uint32_t rate_x, uint32_t rate_y)
{
	static std::mutex mutex;
	static std::vector<uint32_t> tables[kPhaseLutSlots];
	const uint32_t slot = rate_y * 3u + rate_x;
	if (slot >= kPhaseLutSlots)
		throw std::runtime_error("sampling grid without a phase table");
	if (!phase_lut_)
		check(hipMalloc(&phase_lut_, size_t(kPhaseLutSlots) * PGS_PHASE_LUT_WORDS * sizeof(uint32_t)));
	uint32_t *dst = phase_lut_ + size_t(slot) * PGS_PHASE_LUT_WORDS;
	if (!(phase_lut_loaded_ & (1u << slot)))
	{
		std::lock_guard<std::mutex> lock(mutex);
		if (tables[slot].empty())
		{
			tables[slot].resize(PGS_PHASE_LUT_WORDS);
			pgs_phase_lut(rate_x, rate_y, tables[slot].data());
		}
		check(hipMemcpy(dst, tables[slot].data(), PGS_PHASE_LUT_WORDS * sizeof(uint32_t), hipMemcpyHostToDevice));
		phase_lut_loaded_ |= 1u << slot;
	}
	return dst;
}

// Shading of a pass whose depth buffer is its colour buffer (gs/gs_renderer.cpp:2525-2690).
// A primitive's colour may land on another pixel's depth (the two formats swizzle
// differently), so the order of the writes matters:
// * primitives before the first depth-sensitive one shade together (depth unused);
// * if any primitive writes depth, every primitive from the first depth-sensitive one to the
//   last depth writer shades alone, reading the VRAM as it was before it (a copy), storing
//   all colour first and then all depth;
// * later primitives that only test depth shade in runs that read depth from a copy taken
//   when the run starts; a run ends before a depth-sensitive sprite that overlaps the
//   sprites already in the run (gs/gs_renderer.cpp:2418-2475), so the sprite sees their
//   colour as depth;
// * without any, the rest shades without depth.
// Launches that read colour and write it to the same pixel only are race-free; the copies
// make every other read independent of the launch's own writes.
void Renderer::shade_aliased_pass(const RenderPassDesc &d, const Launch &base, const uint8_t *cache_blocks,
                                  uint32_t num_cache_blocks)
{
	if (!alias_cache_)
	{
		const uint32_t capacity_planes = capacity_log2_ ? 2u + (1u << capacity_log2_) : 1u;
		check(hipMalloc(&alias_cache_, size_t(capacity_planes) * kVramSize));
	}
	const uint32_t n = d.num_primitives;
	auto refresh_cache = [&]() {
		Launch lc{};
		lc.kind = LaunchKind::VramSnapshot;
		lc.snapshot = { vram_, alias_cache_, reinterpret_cast<const uint32_t *>(cache_blocks), num_cache_blocks, 0 };
		launch(lc);
	};
	auto shade = [&](uint32_t first, uint32_t count, bool read_copy, bool depth_copy, uint32_t skip, bool no_depth) {
		Launch l = base;
		l.shade.prim_first = first;
		l.shade.prim_count = count;
		l.shade.src = read_copy ? alias_cache_ : nullptr;
		l.shade.zsrc = depth_copy ? alias_cache_ : nullptr;
		l.shade.skip_stores = skip;
		if (no_depth)
			l.shade.z_psm = 0xffffffffu;
		launch(l);
	};

	uint32_t first_z = ~0u, last_zwrite = 0, last_z = 0;
	bool single_step = false, depth_feedback_before = false;
	for (uint32_t i = 0; i < n; i++)
	{
		const uint32_t st = d.prims[i].state;
		if (st & ST_Z_WRITE)
		{
			single_step = true;
			last_zwrite = i;
		}
		if (st & (ST_Z_TEST | ST_Z_WRITE))
		{
			if (first_z == ~0u)
				first_z = i;
			last_z = i;
		}
		else if (first_z == ~0u && (d.prims[i].tex & (TEX_FEEDBACK | TEX_FEEDBACK_DEPTH)) == (TEX_FEEDBACK | TEX_FEEDBACK_DEPTH))
			depth_feedback_before = true;
	}
	if (first_z == ~0u)
	{
		launch(base);
		return;
	}

	// Runs of primitives that read depth from a copy taken as the run starts.
	auto cached_runs = [&](uint32_t lo) {
		int32_t bb[4] = { INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN };
		for (uint32_t i = lo; i < n; i++)
		{
			const PrimitiveAttribute &p = d.prims[i];
			if (!(p.state & ST_SPRITE))
				continue;
			const bool overlap = std::max<int32_t>(bb[0], p.bb[0]) <= std::min<int32_t>(bb[2], p.bb[2]) &&
			                     std::max<int32_t>(bb[1], p.bb[1]) <= std::min<int32_t>(bb[3], p.bb[3]) &&
			                     (p.state & (ST_Z_TEST | ST_Z_WRITE)) != 0;
			if (overlap)
			{
				refresh_cache();
				shade(lo, i - lo, false, true, 0, false);
				lo = i;
				bb[0] = p.bb[0]; bb[1] = p.bb[1]; bb[2] = p.bb[2]; bb[3] = p.bb[3];
			}
			else
			{
				bb[0] = std::min<int32_t>(bb[0], p.bb[0]);
				bb[1] = std::min<int32_t>(bb[1], p.bb[1]);
				bb[2] = std::max<int32_t>(bb[2], p.bb[2]);
				bb[3] = std::max<int32_t>(bb[3], p.bb[3]);
			}
		}
		if (lo < n)
		{
			refresh_cache();
			shade(lo, n - lo, false, true, 0, false);
		}
	};

	if (first_z != 0)
	{
		// Depth is unused here except by on-tile depth feedback, which reads the copy.
		if (depth_feedback_before)
			refresh_cache();
		shade(0, first_z, false, depth_feedback_before, 0, false);
	}
	if (!single_step)
	{
		cached_runs(first_z);
		return;
	}
	for (uint32_t i = first_z; i <= last_zwrite; i++)
	{
		refresh_cache();
		shade(i, 1, true, true, 2u, false); // colour
		shade(i, 1, true, true, 1u, false); // then depth
	}
	if (last_zwrite + 1u < n)
	{
		if (last_z > last_zwrite)
			cached_runs(last_zwrite + 1u);
		else
			shade(last_zwrite + 1u, n - last_zwrite - 1u, false, false, 0, true);
	}
}

// ---------------------------------------------------------------------------------------
// Host VRAM access.

bool Renderer::set_sample_log2(uint32_t sample_log2)
{
	if (sample_log2 > capacity_log2_ || recording_)
		return false;
	sync();
	const uint32_t capacity_planes = capacity_log2_ ? 2u + (1u << capacity_log2_) : 1u;
	if (capacity_planes > 1u)
		check(hipMemset(vram_ + kVramSize, 0, size_t(capacity_planes - 1u) * kVramSize));
	sample_log2_ = sample_log2;
	num_planes_ = sample_log2 ? 2u + (1u << sample_log2) : 1u;
	return true;
}

void Renderer::poll_timings(double *ms, uint32_t kinds, bool wait)
{
	if (wait)
		sync_streams();
	size_t done = 0;
	for (; done < timed_.size(); done++)
	{
		TimedLaunch &t = timed_[done];
		if (!wait && hipEventQuery(t.stop) != hipSuccess)
			break;
		float e = 0.0f;
		check(hipEventElapsedTime(&e, t.start, t.stop));
		if (t.kind < kinds)
			ms[t.kind] += e;
		event_pool_.push_back(t.start);
		event_pool_.push_back(t.stop);
	}
	timed_.erase(timed_.begin(), timed_.begin() + std::ptrdiff_t(done));
}

void Renderer::read_vram(void *dst, uint32_t first_plane, uint32_t planes)
{
	sync();
	check(hipMemcpy(dst, vram_ + size_t(first_plane) * kVramSize, size_t(planes) * kVramSize,
	                hipMemcpyDeviceToHost));
}

void Renderer::write_vram(const void *src, size_t offset, size_t size)
{
	sync();
	check(hipMemcpy(vram_ + offset, src, size, hipMemcpyHostToDevice));
	// A host write invalidates the super-samples of what it overwrites: the reference plane of
	// the range is cleared (gs/gs_renderer.cpp:1478-1492).
	if (sample_log2_)
		check(hipMemset(vram_ + kVramSize + offset, 0, size));
}

// ---------------------------------------------------------------------------------------
// Recording and replay.

// Recording keeps every upload alive and remembers every launch; the VRAM planes and the
// CLUT at the start are snapshotted so that each replay starts from the same state and
// redoes exactly the recorded work.
void Renderer::begin_recording()
{
	sync();
	destroy_graph();
	// The previous recording's arenas hold nothing a later launch reads any more.
	for (auto &a : retired_arenas_)
	{
		if (a.host) (void)hipHostFree(a.host);
		if (a.dev) (void)hipFree(a.dev);
	}
	retired_arenas_.clear();
	recorded_.clear();
	replay_.clear();
	restore_ranges_.clear();
	if (restore_dev_)
		check(hipFree(restore_dev_));
	restore_dev_ = nullptr;
	const size_t vram_bytes = size_t(num_planes_) * kVramSize;
	if (snapshot_ && snapshot_bytes_ < vram_bytes + 1024u)
	{
		check(hipFree(snapshot_));
		snapshot_ = nullptr;
	}
	if (!snapshot_)
	{
		check(hipMalloc(&snapshot_, vram_bytes + 1024u));
		snapshot_bytes_ = vram_bytes + 1024u;
	}
	check(hipMemcpyAsync(snapshot_, vram_, vram_bytes, hipMemcpyDeviceToDevice, stream_));
	check(hipMemcpyAsync(snapshot_ + vram_bytes, clut_buf_, 1024u, hipMemcpyDeviceToDevice, stream_));
	main_pending_ = true;
	recording_ = true;
}

void Renderer::set_restore_pages(const std::vector<uint32_t> &pages)
{
	restore_ranges_.clear();
	const size_t vram_bytes = size_t(num_planes_) * kVramSize;
	constexpr uint32_t kRunPages = 8; // 64 KiB per range
	for (uint32_t plane = 0; plane < num_planes_; plane++)
	{
		const size_t base = size_t(plane) * kVramSize;
		for (size_t i = 0; i < pages.size();)
		{
			size_t j = i + 1;
			while (j < pages.size() && pages[j] == pages[j - 1] + 1 && j - i < kRunPages)
				j++;
			const size_t off = base + size_t(pages[i]) * kPageBytes;
			restore_ranges_.push_back({ snapshot_ + off, vram_ + off, uint32_t((j - i) * kPageBytes), 0 });
			i = j;
		}
	}
	restore_ranges_.push_back({ snapshot_ + vram_bytes, reinterpret_cast<uint8_t *>(clut_buf_), 1024u, 0 });
	if (restore_dev_)
		check(hipFree(restore_dev_));
	restore_dev_ = nullptr;
	check(hipMalloc(&restore_dev_, restore_ranges_.size() * sizeof(pgs_copy_range)));
	check(hipMemcpy(restore_dev_, restore_ranges_.data(), restore_ranges_.size() * sizeof(pgs_copy_range),
	                hipMemcpyHostToDevice));
}

void Renderer::restore_snapshot(hipStream_t stream)
{
	if (restore_dev_)
	{
		pgs_restore_args r{ restore_dev_, uint32_t(restore_ranges_.size()), 0 };
		const int rc = pgs_k_restore(&r, stream);
		if (rc != 0)
			check(hipError_t(-rc));
		return;
	}
	const size_t vram_bytes = size_t(num_planes_) * kVramSize;
	check(hipMemcpyAsync(vram_, snapshot_, vram_bytes, hipMemcpyDeviceToDevice, stream));
	check(hipMemcpyAsync(clut_buf_, snapshot_ + vram_bytes, 1024u, hipMemcpyDeviceToDevice, stream));
}

void Renderer::end_recording()
{
	if (!batch_.empty())
		flush_transfers();
	arena_upload();
	sync_streams();
	recording_ = false;
	// The recorded launches point into the arena: it stays with the recording, and work after
	// the recording allocates a fresh arena on its first upload (arena_make_room).
	retired_arenas_.push_back(arena_);
	arena_ = {};
	if (!side_)
		fuse_replay();
}

// ---------------------------------------------------------------------------------------
// Fused replays. Each launch reads and writes a few resources; a launch's level is one more
// than the highest level of an earlier launch it depends on (reads what that one writes, or
// writes what that one reads or writes). Launches of one level are independent, so the
// fusable ones among them (every kernel but shading and sampler feedback, whose workgroups
// have other shapes) become one launch of 256-thread workgroups. A textured frame's copy,
// CLUT load, decode, setup, heuristic, binning and shading (seven launches) replay as four:
// {copy, setup}, {CLUT, heuristic}, {decode, binning}, shading.

namespace
{
enum : uint32_t
{
	R_VRAM = 1u << 0, R_SHADOW = 1u << 1, R_CLUT = 1u << 2, R_TEXIMG = 1u << 3, R_PRIMS = 1u << 4,
	R_SETUP = 1u << 5, R_FLAT = 1u << 6, R_FEEDBACK = 1u << 7, R_BINS = 1u << 8, R_COUNT = 9
};

// Resources a launch reads and writes. PRIMS: the primitive records (state words); SETUP:
// setup records, attributes, state copies and coarse masks; FLAT: the heuristic's bucket mask;
// BINS: tile lists, counts and modes; TEXIMG: decoded images.
void launch_resources(const Renderer::Launch &l, uint32_t &reads, uint32_t &writes)
{
	using K = Renderer::LaunchKind;
	reads = writes = 0;
	switch (l.kind)
	{
	case K::VramCopy: reads = R_VRAM | R_SHADOW; writes = R_VRAM | (l.copy.shadow_next ? R_SHADOW : 0u); break;
	case K::VramSnapshot: reads = R_VRAM; writes = R_SHADOW; break;
	case K::ClutWrite: reads = R_VRAM; writes = R_CLUT; break;
	case K::TexDecode:
		reads = R_VRAM | R_CLUT | (l.decode.feedback ? R_FEEDBACK : 0u);
		writes = R_TEXIMG;
		break;
	case K::Setup:
		reads = R_PRIMS;
		writes = R_PRIMS | R_SETUP | R_FLAT | (l.setup.clear_words ? R_FEEDBACK : 0u);
		break;
	case K::SamplerFeedback: reads = R_SETUP | R_PRIMS; writes = R_FEEDBACK; break;
	case K::SsaaHeuristic: reads = R_PRIMS | R_SETUP | R_FLAT; writes = R_PRIMS | R_FLAT; break;
	case K::Binning: reads = R_SETUP | R_PRIMS; writes = R_BINS; break;
	case K::Shade:
	case K::ShadeSuper:
		// A shading launch of an aliased pass also reads the VRAM copy (a snapshot buffer).
		reads = R_VRAM | R_CLUT | R_SETUP | R_PRIMS | R_BINS | R_TEXIMG | (l.shade.src || l.shade.zsrc ? R_SHADOW : 0u);
		writes = R_VRAM;
		break;
	default: reads = writes = ~0u; break; // anything else orders against everything
	}
}

// The part of a fusable launch (256-thread workgroups); false for launches that do not fuse.
bool fused_part(const Renderer::Launch &l, pgs_fused_part &p)
{
	using K = Renderer::LaunchKind;
	p = {};
	switch (l.kind)
	{
	case K::VramCopy:
		p.kind = PGS_FUSED_VRAM_COPY; p.args.copy = l.copy; p.groups = (l.copy.num_items + 3u) / 4u;
		break;
	case K::VramSnapshot:
		p.kind = PGS_FUSED_VRAM_SNAPSHOT; p.args.snapshot = l.snapshot; p.groups = (l.snapshot.num_blocks + 3u) / 4u;
		break;
	case K::ClutWrite: p.kind = PGS_FUSED_CLUT_WRITE; p.args.clut = l.clut; p.groups = 1; break;
	case K::TexDecode: p.kind = PGS_FUSED_TEXTURE_DECODE; p.args.decode = l.decode; p.groups = l.decode.num_groups; break;
	case K::Setup:
		p.kind = PGS_FUSED_TRIANGLE_SETUP; p.args.setup = l.setup; p.groups = (l.setup.num_primitives + 255u) / 256u;
		break;
	case K::SsaaHeuristic: p.kind = PGS_FUSED_SSAA_HEURISTIC; p.args.heuristic = l.heuristic; p.groups = 256; break;
	case K::Binning: p.kind = PGS_FUSED_BINNING; p.args.bin = l.bin; p.groups = l.bin.tiles_x * l.bin.tiles_y; break;
	default: return false;
	}
	return true;
}
}

void Renderer::fuse_replay()
{
	if (fused_dev_)
		check(hipFree(fused_dev_));
	fused_dev_ = nullptr;
	std::array<int32_t, R_COUNT> last_read, last_write;
	last_read.fill(-1);
	last_write.fill(-1);
	std::vector<int32_t> level(replay_.size());
	int32_t max_level = -1;
	for (size_t i = 0; i < replay_.size(); i++)
	{
		uint32_t reads, writes;
		launch_resources(replay_[i], reads, writes);
		int32_t lv = 0;
		for (uint32_t r = 0; r < R_COUNT; r++)
		{
			if (reads & (1u << r))
				lv = std::max(lv, last_write[r] + 1);
			if (writes & (1u << r))
				lv = std::max(lv, std::max(last_write[r], last_read[r]) + 1);
		}
		for (uint32_t r = 0; r < R_COUNT; r++)
		{
			if (reads & (1u << r))
				last_read[r] = std::max(last_read[r], lv);
			if (writes & (1u << r))
				last_write[r] = std::max(last_write[r], lv);
		}
		level[i] = lv;
		max_level = std::max(max_level, lv);
	}
	std::vector<std::vector<size_t>> by_level(size_t(max_level + 1));
	for (size_t i = 0; i < replay_.size(); i++)
		by_level[size_t(level[i])].push_back(i);

	std::vector<Launch> out;
	std::vector<pgs_fused_part> parts;
	std::vector<std::pair<size_t, size_t>> fused_spans; // (first part, parts) per fused launch, by out index
	std::vector<size_t> fused_out;
	for (const auto &members : by_level)
	{
		std::vector<pgs_fused_part> group;
		std::vector<size_t> single;
		for (size_t i : members)
		{
			pgs_fused_part p;
			if (fused_part(replay_[i], p))
			{
				if (p.groups)
					group.push_back(p);
			}
			else
				single.push_back(i);
		}
		if (group.size() == 1)
		{
			// One part: its own launch, as recorded.
			for (size_t i : members)
			{
				pgs_fused_part p;
				if (fused_part(replay_[i], p) && p.groups)
					out.push_back(replay_[i]);
			}
		}
		else if (group.size() > 1)
		{
			for (size_t first = 0; first < group.size(); first += PGS_FUSED_MAX_PARTS)
			{
				const size_t n = std::min<size_t>(PGS_FUSED_MAX_PARTS, group.size() - first);
				uint32_t begin = 0;
				for (size_t k = 0; k < n; k++)
				{
					group[first + k].begin = begin;
					begin += group[first + k].groups;
				}
				fused_spans.push_back({ parts.size(), n });
				fused_out.push_back(out.size());
				parts.insert(parts.end(), group.begin() + first, group.begin() + first + n);
				Launch l{};
				l.kind = LaunchKind::Fused;
				l.fused.num_parts = uint32_t(n);
				l.fused.groups = begin;
				out.push_back(l);
			}
		}
		for (size_t i : single)
			out.push_back(replay_[i]);
	}
	if (!parts.empty())
	{
		check(hipMalloc(&fused_dev_, parts.size() * sizeof(pgs_fused_part)));
		check(hipMemcpy(fused_dev_, parts.data(), parts.size() * sizeof(pgs_fused_part), hipMemcpyHostToDevice));
		for (size_t k = 0; k < fused_out.size(); k++)
			out[fused_out[k]].fused.parts = fused_dev_ + fused_spans[k].first;
	}
	replay_ = std::move(out);
}

void Renderer::destroy_graph()
{
	if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
	if (graph_) (void)hipGraphDestroy(graph_);
	graph_exec_ = nullptr;
	graph_ = nullptr;
}

int Renderer::replay(bool use_graph)
{
	try
	{
		if (!snapshot_)
			return -int(hipErrorInvalidValue);
		// The replay list holds the recorded ordering points; the first texture-stream launch
		// is preceded by a Fork (recording starts with a snapshot copy on the main stream), so
		// it waits for the restore below. The closing join gives the main stream the whole
		// replay (and closes the capture's fork).
		join_texture_stream();
		if (!use_graph)
		{
			TimedLaunch t{ kProfileRestore, nullptr, nullptr };
			if (profiling_)
			{
				t.start = get_event();
				t.stop = get_event();
				check(hipEventRecord(t.start, stream_));
			}
			restore_snapshot(stream_);
			if (profiling_)
			{
				check(hipEventRecord(t.stop, stream_));
				timed_.push_back(t);
			}
			for (auto &l : replay_)
				issue(l);
			join_texture_stream();
			return 0;
		}
		if (!graph_exec_)
		{
			check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
			restore_snapshot(stream_);
			for (auto &l : replay_)
				issue(l);
			join_texture_stream();
			check(hipStreamEndCapture(stream_, &graph_));
			check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
		}
		check(hipGraphLaunch(graph_exec_, stream_));
		return 0;
	}
	catch (const std::exception &)
	{
		return last_error_ ? last_error_ : -1;
	}
}
}

namespace pgs
{
void Renderer::clut_load(const pgs_clut_args &args)
{
	Launch l{};
	l.kind = LaunchKind::ClutWrite;
	l.clut = args;
	l.clut.vram = vram_;
	l.clut.clut = clut_buf_;
	launch(l);
}
}
