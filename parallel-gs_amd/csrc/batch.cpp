// Batched replay of many GS contexts; see batch.hpp.
#include "batch.hpp"
#include "multi_dispatch.hpp"
#include <algorithm>
#include <map>
#include <stdexcept>
#include <cstring>
#include <cmath>

namespace pgs
{
namespace
{
void ok(hipError_t e)
{
	if (e != hipSuccess)
		throw std::runtime_error(hipGetErrorString(e));
}

uint32_t grid_of(const Renderer::Launch &l)
{
	using K = Renderer::LaunchKind;
	switch (l.kind)
	{
	case K::VramCopy: return pgs_grid_vram_copy(l.copy);
	case K::VramSnapshot: return pgs_grid_vram_snapshot(l.snapshot);
	case K::Setup: return pgs_grid_triangle_setup(l.setup);
	case K::Binning: return pgs_grid_binning(l.bin);
	case K::Shade: case K::ShadeSuper: return pgs_grid_shade(l.shade);
	case K::ClutWrite: return pgs_grid_clut_write(l.clut);
	case K::TexDecode: return pgs_grid_texture_decode(l.decode);
	case K::SsaaHeuristic: return pgs_grid_ssaa_heuristic(l.heuristic);
	case K::SamplerFeedback: return pgs_grid_sampler_feedback(l.feedback);
	case K::Fork: case K::Join: case K::Fused: break; // not in recorded kernel lists
	}
	return 0;
}

size_t args_size(Renderer::LaunchKind k)
{
	using K = Renderer::LaunchKind;
	switch (k)
	{
	case K::VramCopy: return sizeof(pgs_vram_copy_args);
	case K::VramSnapshot: return sizeof(pgs_vram_snapshot_args);
	case K::Setup: return sizeof(pgs_setup_args);
	case K::Binning: return sizeof(pgs_bin_args);
	case K::Shade: case K::ShadeSuper: return sizeof(pgs_shade_args);
	case K::ClutWrite: return sizeof(pgs_clut_args);
	case K::TexDecode: return sizeof(pgs_tex_decode_args);
	case K::SsaaHeuristic: return sizeof(pgs_ssaa_heuristic_args);
	case K::SamplerFeedback: return sizeof(pgs_feedback_args);
	case K::Fork: case K::Join: case K::Fused: break;
	}
	return 0;
}

// Launches that may share one multi-launch: same kernel, and for shading the same sampling grid.
uint64_t key_of(const Renderer::Launch &l)
{
	uint64_t k = uint64_t(l.kind) << 16;
	if (l.kind == Renderer::LaunchKind::Shade || l.kind == Renderer::LaunchKind::ShadeSuper)
		k |= uint64_t(l.shade.rate_x_log2) << 8 | l.shade.rate_log2 | (l.shade.phase_lut ? 1ull << 12 : 0ull);
	return k;
}

size_t align256(size_t v) { return (v + 255u) & ~size_t(255); }
// Queue entries of a shading launch: its tiles plus what cost-based splitting may add (at most
// 2 * slots, and a launch has far fewer than 65536 resident workgroups).
uint32_t order_capacity(uint32_t tiles) { return tiles + 2u * 65536u + 1024u; }
// Entries of a shading launch's queue or dispatch order: at most one per wave (base items,
// chunks) plus what cost splitting adds.
uint32_t order_entries(uint32_t tiles, uint32_t waves)
{
	const uint64_t n = uint64_t(order_capacity(tiles)) + waves;
	if (n > 0x7fffffffull)
		throw std::runtime_error("batch: too many shading waves for a schedule buffer");
	return uint32_t(n);
}
}

Batch::~Batch()
{
	destroy_graph();
	free_plans();
	if (dev_)
		(void)hipFree(dev_);
	if (restore_dev_)
		(void)hipFree(restore_dev_);
	if (sched_)
		(void)hipFree(sched_);
	if (stream_)
		(void)hipStreamDestroy(stream_);
}

void Batch::destroy_graph()
{
	if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
	if (graph_) (void)hipGraphDestroy(graph_);
	graph_exec_ = nullptr;
	graph_ = nullptr;
}

bool Batch::build(Renderer *const *contexts, uint32_t count, bool merge)
{
	contexts_.assign(contexts, contexts + count);
	for (Renderer *r : contexts_)
		r->sync();
	if (!stream_)
		ok(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));

	// The contexts' restore ranges (pages their recordings write, and the CLUT) in one list,
	// restored by one launch per replay; contexts without such a list restore whole planes.
	if (restore_dev_)
		ok(hipFree(restore_dev_));
	restore_dev_ = nullptr;
	restore_count_ = 0;
	std::vector<pgs_copy_range> ranges;
	bool all_listed = true;
	for (Renderer *r : contexts_)
	{
		if (r->restore_ranges().empty())
			all_listed = false;
		ranges.insert(ranges.end(), r->restore_ranges().begin(), r->restore_ranges().end());
	}
	if (all_listed && !ranges.empty())
	{
		ok(hipMalloc(&restore_dev_, ranges.size() * sizeof(pgs_copy_range)));
		ok(hipMemcpy(restore_dev_, ranges.data(), ranges.size() * sizeof(pgs_copy_range), hipMemcpyHostToDevice));
		restore_count_ = uint32_t(ranges.size());
	}

	merged_.clear();
	if (!merge)
	{
		for (uint32_t c = 0; c < count; c++)
		{
			const auto &rec = contexts_[c]->recorded();
			for (uint32_t i = 0; i < rec.size(); i++)
			{
				Merged m;
				m.kind = rec[i].kind;
				m.refs.push_back({ c, i });
				merged_.push_back(std::move(m));
			}
		}
		return apply(tuning_);
	}

	// Greedy merge: the key most contexts have next goes first; per-context order is kept.
	std::vector<uint32_t> cursor(count, 0);
	for (;;)
	{
		std::map<uint64_t, uint32_t> votes;
		for (uint32_t c = 0; c < count; c++)
		{
			const auto &rec = contexts_[c]->recorded();
			if (cursor[c] < rec.size())
				votes[key_of(rec[cursor[c]])]++;
		}
		if (votes.empty())
			break;
		uint64_t best = votes.begin()->first;
		uint32_t best_n = 0;
		for (auto &v : votes)
			if (v.second > best_n)
			{
				best = v.first;
				best_n = v.second;
			}
		Merged m;
		for (uint32_t c = 0; c < count; c++)
		{
			const auto &rec = contexts_[c]->recorded();
			if (cursor[c] < rec.size() && key_of(rec[cursor[c]]) == best)
			{
				m.kind = rec[cursor[c]].kind;
				m.refs.push_back({ c, cursor[c] });
				cursor[c]++;
			}
		}
		merged_.push_back(std::move(m));
	}
	return apply(tuning_);
}

bool Batch::apply(const BatchTuning &t)
{
	if (t.waves_per_group != 1 && t.waves_per_group != 2 && t.waves_per_group != 4)
		return false;
	if (t.tile_log2_1x != 0 && (t.tile_log2_1x < 3 || t.tile_log2_1x > 6))
		return false;
	using K = Renderer::LaunchKind;
	std::vector<std::vector<Renderer::Launch>> out(contexts_.size());
	for (size_t c = 0; c < contexts_.size(); c++)
	{
		out[c] = contexts_[c]->recorded();
		uint32_t tl = 0, tx = 0, ty = 0, coarse_log2 = 0;
		bool retiled = false;
		for (auto &l : out[c])
		{
			if (l.kind == K::Setup)
				coarse_log2 = l.setup.coarse_log2;
			else if (l.kind == K::Binning)
			{
				retiled = false;
				if (t.tile_log2_1x != 0 && !l.bin.super_sampled)
				{
					const uint32_t old = l.bin.tile_log2;
					const uint32_t ex = l.bin.tiles_x << old, ey = l.bin.tiles_y << old;
					tl = t.tile_log2_1x;
					tx = (ex + (1u << tl) - 1u) >> tl;
					ty = (ey + (1u << tl) - 1u) >> tl;
					if (size_t(tx) * ty * l.bin.list_stride > contexts_[c]->lists_capacity() ||
					    size_t(tx) * ty > contexts_[c]->counts_capacity())
						return false;
					l.bin.tile_log2 = tl;
					l.bin.tiles_x = tx;
					l.bin.tiles_y = ty;
					// The coarse mask keeps the grid of its setup (same origin, coarse tiles of
					// 2^coarse_log2 pixels); the new tiles still nest in it while they are not
					// larger, otherwise every primitive is tested.
					if (l.bin.coarse_mask && coarse_log2 >= tl)
						l.bin.coarse_shift = coarse_log2 - tl;
					else
						l.bin.coarse_mask = nullptr;
					retiled = true;
				}
			}
			else if (l.kind == K::Shade || l.kind == K::ShadeSuper)
			{
				l.shade.waves_per_group = t.waves_per_group;
				if (retiled && !l.shade.tile_modes)
				{
					l.shade.tile_log2 = tl;
					l.shade.tiles_x = tx;
					l.shade.tiles_y = ty;
				}
			}
		}
	}
	launches_ = std::move(out);
	tuning_ = t;
	return upload();
}

bool Batch::upload()
{
	// Layout: per merged launch, its argument blocks, its workgroup prefix and (shading) its
	// coarse-tile prefix.
	auto is_shade = [](const Merged &m) {
		return m.kind == Renderer::LaunchKind::Shade || m.kind == Renderer::LaunchKind::ShadeSuper;
	};
	size_t bytes = 0;
	for (auto &m : merged_)
	{
		m.args_offset = bytes;
		bytes = align256(bytes + args_size(m.kind) * m.refs.size());
		m.prefix_offset = bytes;
		bytes = align256(bytes + sizeof(uint32_t) * (m.refs.size() + 1u));
		m.tile_prefix_offset = bytes;
		if (is_shade(m))
			bytes = align256(bytes + sizeof(uint32_t) * (m.refs.size() + 1u));
		m.run = m.run_n = m.run_total = 0;
		m.in_run = false;
		m.run_all_grids = false;
	}
	// Runs of adjacent shading launches whose contexts are disjoint: passes of different
	// sampling grids that no context orders against each other. Only grids the mixed-grid
	// kernel holds (1x, 2x, 4x, 8x, 16x) join a run.
	auto mixable = [&](const Merged &m) {
		if (!is_shade(m))
			return false;
		const pgs_shade_args &s = launches_[m.refs[0].context][m.refs[0].index].shade;
		const uint32_t rx = s.rate_x_log2, ry = s.rate_log2;
		return !s.phase_lut && (ry == 0 || (ry == 1 && rx <= 1) || (ry == 2 && (rx == 1 || rx == 2)));
	};
	for (size_t i = 0; i < merged_.size();)
	{
		if (!mixable(merged_[i]))
		{
			i++;
			continue;
		}
		std::vector<bool> used(contexts_.size(), false);
		size_t j = i;
		for (; j < merged_.size() && mixable(merged_[j]); j++)
		{
			bool clash = false;
			for (const Ref &r : merged_[j].refs)
				clash = clash || used[r.context];
			if (clash)
				break;
			for (const Ref &r : merged_[j].refs)
				used[r.context] = true;
		}
		if (j - i >= 2)
		{
			Merged &h = merged_[i];
			h.run = uint32_t(j - i);
			for (size_t k = i; k < j; k++)
			{
				h.run_n += uint32_t(merged_[k].refs.size());
				merged_[k].in_run = k > i;
			}
			h.run_args_offset = bytes;
			bytes = align256(bytes + sizeof(pgs_shade_args) * h.run_n);
			h.run_prefix_offset = bytes;
			bytes = align256(bytes + sizeof(uint32_t) * (h.run_n + 1u));
		}
		i = j;
	}
	std::vector<uint8_t> host(bytes);
	for (auto &m : merged_)
	{
		const size_t sz = args_size(m.kind);
		uint32_t *prefix = reinterpret_cast<uint32_t *>(host.data() + m.prefix_offset);
		uint32_t total = 0;
		for (size_t i = 0; i < m.refs.size(); i++)
		{
			const Renderer::Launch &l = launches_[m.refs[i].context][m.refs[i].index];
			memcpy(host.data() + m.args_offset + i * sz, &l.copy, sz);
			prefix[i] = total;
			total += grid_of(l);
		}
		prefix[m.refs.size()] = total;
		m.total = total;
		if (is_shade(m))
		{
			uint32_t *tp = reinterpret_cast<uint32_t *>(host.data() + m.tile_prefix_offset);
			uint32_t tiles = 0;
			uint64_t waves = 0;
			for (size_t i = 0; i < m.refs.size(); i++)
			{
				const Renderer::Launch &l = launches_[m.refs[i].context][m.refs[i].index];
				tp[i] = tiles;
				tiles += l.shade.tiles_x * l.shade.tiles_y;
				const uint64_t B = 1ull << (l.shade.tile_log2 - 3u);
				waves += uint64_t(l.shade.tiles_x) * l.shade.tiles_y * B * B << (l.shade.rate_x_log2 + l.shade.rate_log2);
			}
			tp[m.refs.size()] = tiles;
			m.tiles = tiles;
			m.waves = uint32_t(std::min<uint64_t>(waves, 0xffffffffu));
		}
		if (m.kind == Renderer::LaunchKind::Shade || m.kind == Renderer::LaunchKind::ShadeSuper)
		{
			const Renderer::Launch &l0 = launches_[m.refs[0].context][m.refs[0].index];
			m.rate_x = l0.shade.rate_x_log2 | (l0.shade.phase_lut ? kRateArrays : 0u);
			m.rate_y = l0.shade.rate_log2;
		}
	}
	for (size_t i = 0; i < merged_.size(); i++)
	{
		Merged &h = merged_[i];
		if (h.run < 2)
			continue;
		pgs_shade_args *args = reinterpret_cast<pgs_shade_args *>(host.data() + h.run_args_offset);
		uint32_t *prefix = reinterpret_cast<uint32_t *>(host.data() + h.run_prefix_offset);
		uint32_t at = 0, total = 0;
		for (size_t k = i; k < i + h.run; k++)
			for (const Ref &r : merged_[k].refs)
			{
				const Renderer::Launch &l = launches_[r.context][r.index];
				const uint32_t rx = l.shade.rate_x_log2, ry = l.shade.rate_log2;
				h.run_all_grids = h.run_all_grids || !(ry == 0 || (ry == 1 && rx <= 1));
				args[at] = l.shade;
				prefix[at++] = total;
				total += grid_of(l);
			}
		prefix[at] = total;
		h.run_total = total;
	}
	if (bytes > dev_bytes_)
	{
		if (dev_)
			ok(hipFree(dev_));
		ok(hipMalloc(&dev_, bytes));
		dev_bytes_ = bytes;
	}
	ok(hipMemcpy(dev_, host.data(), bytes, hipMemcpyHostToDevice));

	// Schedule buffers per shading launch: features, predictions, costs, waves per tile, the
	// two per-tile placement words of the order build, the counter, then the entries: at most
	// one per wave (work-queue base items, dispatch chunks) plus what cost splitting adds.
	size_t sb = 0;
	for (auto &m : merged_)
		if (is_shade(m))
		{
			m.sched_offset = sb;
			sb = align256(sb + size_t(m.tiles) * (PGS_MLP_INPUTS + 5u) * 4u + 8u + size_t(order_entries(m.tiles, m.waves)) * 4u);
		}
	if (sb > sched_bytes_)
	{
		if (sched_)
			ok(hipFree(sched_));
		ok(hipMalloc(&sched_, sb));
		sched_bytes_ = sb;
	}
	if (plans_.size() != merged_.size())
	{
		free_plans();
		plans_.resize(merged_.size());
	}
	planned_ = false;
	destroy_graph();
	return true;
}

bool Batch::set_schedule(Schedule s, const CostModel *model)
{
	if ((s == Schedule::QueueLearned || s == Schedule::DispatchLearned) && !model)
		return false;
	schedule_ = s;
	model_ = model;
	planned_ = false;
	destroy_graph();
	return true;
}

void Batch::free_plans()
{
	for (Plan &p : plans_)
		if (p.order)
			(void)hipFree(p.order);
	plans_.clear();
	planned_ = false;
}

// Dispatch order placed per XCD. The dispatcher hands workgroup b to XCD b % 8 in index order,
// so XCD x runs the entries at 8 k + x one after another; unused places hold 0xffffffff (the
// workgroup exits). Each context's tiles with work are cut into eight runs in tile order, run x
// going to XCD x (the screen bands the hardware schedule's remap gives the XCDs, so every XCD
// gets a share of every context). Static order cuts the runs by chunk count and keeps them in
// context order. Learned order cuts them by predicted cost, so that every XCD gets the same
// predicted work, takes the contexts by predicted cost, largest first, and starts, on each XCD,
// the tiles whose predicted cost is above half of what one resident workgroup of the XCD does on
// average, largest first (so that no heavy tile starts last); the other tiles follow in context
// and tile order. A tile's chunks stay together.
void Batch::plan_dispatch(const Merged &m, Plan &plan, const float *features, const float *pred,
                          const uint32_t *tile_waves, uint32_t *counter, uint32_t group_waves)
{
	constexpr uint32_t kXcds = 8;
	const uint32_t n = m.tiles;
	std::vector<float> f(size_t(n) * PGS_MLP_INPUTS), p;
	std::vector<uint32_t> waves(n);
	ok(hipStreamSynchronize(stream_));
	ok(hipMemcpy(f.data(), features, f.size() * sizeof(float), hipMemcpyDeviceToHost));
	ok(hipMemcpy(waves.data(), tile_waves, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
	if (pred)
	{
		p.resize(n);
		ok(hipMemcpy(p.data(), pred, n * sizeof(float), hipMemcpyDeviceToHost));
	}
	auto has_work = [&](uint32_t i) { return f[size_t(i) * PGS_MLP_INPUTS] != 0.0f && waves[i] != 0; };
	auto chunks = [&](uint32_t i) { return (waves[i] + group_waves - 1u) / group_waves; };
	// Static runs are cut by chunks, learned runs by predicted cost.
	auto weight = [&](uint32_t i) {
		return pred ? std::exp2(double(std::min(p[i], 40.0f))) : double(chunks(i));
	};
	std::vector<std::vector<uint32_t>> seq(kXcds); // tiles per XCD, in run order
	double load[kXcds] = {};
	// Contexts (launches of the multi-launch) in launch order, or, learned, by predicted cost,
	// largest first, so that the last contexts to start are the light ones.
	std::vector<uint32_t> firsts(m.refs.size() + 1u, 0u);
	for (size_t k = 0; k < m.refs.size(); k++)
	{
		const pgs_shade_args &sa = launches_[m.refs[k].context][m.refs[k].index].shade;
		firsts[k + 1] = firsts[k] + sa.tiles_x * sa.tiles_y;
	}
	std::vector<uint32_t> launch_order(m.refs.size());
	std::vector<double> launch_cost(m.refs.size(), 0.0);
	for (size_t k = 0; k < m.refs.size(); k++)
	{
		launch_order[k] = uint32_t(k);
		for (uint32_t i = firsts[k]; i < firsts[k + 1]; i++)
			if (has_work(i))
				launch_cost[k] += weight(i);
	}
	if (pred)
		std::stable_sort(launch_order.begin(), launch_order.end(),
		                 [&](uint32_t a, uint32_t b) { return launch_cost[a] > launch_cost[b]; });
	for (uint32_t k : launch_order)
	{
		const uint32_t first = firsts[k], end = firsts[k + 1];
		const double total = launch_cost[k];
		double at = 0.0;
		for (uint32_t i = first; i < end; i++)
			if (has_work(i))
			{
				const double w = weight(i);
				// The run a tile falls in by the midpoint of its weight.
				const uint32_t x = total > 0.0 ? std::min(kXcds - 1u, uint32_t((at + 0.5 * w) * kXcds / total)) : 0u;
				seq[x].push_back(i);
				load[x] += w;
				at += w;
			}
	}
	if (pred)
	{
		const uint32_t slots = std::max(1u, pgs_queue_slots(m.rate_x, m.rate_y, tuning_.waves_per_group, m.tiles) / kXcds);
		for (uint32_t x = 0; x < kXcds; x++)
		{
			const double heavy = 0.5 * load[x] / slots;
			std::stable_partition(seq[x].begin(), seq[x].end(), [&](uint32_t i) { return weight(i) > heavy; });
			const auto mid = std::find_if(seq[x].begin(), seq[x].end(), [&](uint32_t i) { return weight(i) <= heavy; });
			std::stable_sort(seq[x].begin(), mid, [&](uint32_t a, uint32_t b) { return weight(a) > weight(b); });
		}
	}
	size_t longest = 0;
	std::vector<std::vector<uint32_t>> entries(kXcds);
	for (uint32_t x = 0; x < kXcds; x++)
	{
		for (uint32_t i : seq[x])
			for (uint32_t k = 0; k < chunks(i); k++)
				entries[x].push_back((i << 12) | k);
		longest = std::max(longest, entries[x].size());
	}
	plan.tiles = 0;
	plan.heavy = 0;
	for (uint32_t x = 0; x < kXcds; x++)
		plan.tiles += uint32_t(seq[x].size());
	if (pred)
		for (uint32_t x = 0; x < kXcds; x++)
		{
			const uint32_t slots = std::max(1u, pgs_queue_slots(m.rate_x, m.rate_y, tuning_.waves_per_group, m.tiles) / kXcds);
			const double heavy = 0.5 * load[x] / slots;
			for (uint32_t i : seq[x])
				plan.heavy += weight(i) > heavy ? 1u : 0u;
		}
	const size_t grid = longest * kXcds;
	if (grid > 0x7fffffffu)
		throw std::runtime_error("dispatch plan: too many workgroups");
	std::vector<uint32_t> order(grid, 0xffffffffu);
	for (uint32_t x = 0; x < kXcds; x++)
		for (size_t k = 0; k < entries[x].size(); k++)
			order[k * kXcds + x] = entries[x][k];
	if (plan.capacity < grid)
	{
		if (plan.order)
			ok(hipFree(plan.order));
		plan.order = nullptr;
		ok(hipMalloc(&plan.order, std::max<size_t>(grid, 1u) * sizeof(uint32_t)));
		plan.capacity = grid;
	}
	plan.grid = uint32_t(grid);
	if (grid)
		ok(hipMemcpy(plan.order, order.data(), grid * sizeof(uint32_t), hipMemcpyHostToDevice));
	const uint32_t limits[2] = { 0u, uint32_t(grid) };
	ok(hipMemcpy(counter, limits, sizeof(limits), hipMemcpyHostToDevice));
}

// Shading launch under the current schedule.
void Batch::launch_shade(const Merged &m)
{
	const void *args = dev_ + m.args_offset;
	const uint32_t *prefix = reinterpret_cast<const uint32_t *>(dev_ + m.prefix_offset);
	const uint32_t n = uint32_t(m.refs.size());
	if (schedule_ == Schedule::Hardware && !collecting_ && m.run >= 2)
	{
		if (pgs_multi_shade_mixed(reinterpret_cast<const pgs_shade_args *>(dev_ + m.run_args_offset),
		                          reinterpret_cast<const uint32_t *>(dev_ + m.run_prefix_offset), m.run_n,
		                          m.run_total, tuning_.waves_per_group, m.run_all_grids, stream_) != 0)
			throw std::runtime_error("shade multi-launch failed");
		return;
	}
	if (schedule_ == Schedule::Hardware && !collecting_)
	{
		if (pgs_multi_shade(static_cast<const pgs_shade_args *>(args), prefix, n, m.total, m.rate_x, m.rate_y,
		                    tuning_.waves_per_group, stream_) != 0)
			throw std::runtime_error("shade multi-launch failed");
		return;
	}
	uint8_t *base = sched_ + m.sched_offset;
	float *features = reinterpret_cast<float *>(base);
	float *pred = features + size_t(m.tiles) * PGS_MLP_INPUTS;
	uint32_t *costs = reinterpret_cast<uint32_t *>(pred + m.tiles);
	uint32_t *tile_waves = costs + m.tiles;
	uint32_t *run_start = tile_waves + m.tiles;
	uint32_t *run_weight = run_start + m.tiles;
	uint32_t *counter = run_weight + m.tiles;
	uint32_t *order = counter + 2;
	const uint32_t *tile_prefix = reinterpret_cast<const uint32_t *>(dev_ + m.tile_prefix_offset);
	if (m.tiles == 0)
		return;
	const bool dispatch = (schedule_ == Schedule::DispatchStatic || schedule_ == Schedule::DispatchLearned) &&
	                      !collecting_;
	if (m.tiles >= (dispatch ? (1u << 20) : (1u << 24)))
		throw std::runtime_error("scheduled shading: too many tiles");
	const uint32_t group_waves = tuning_.waves_per_group ? tuning_.waves_per_group : 4u;
	Plan &plan = plans_[size_t(&m - merged_.data())];

	pgs_queue_args q{};
	q.args = args;
	q.prefix = tile_prefix;
	q.order = dispatch ? plan.order : order;
	q.counter = counter;
	q.costs = collecting_ ? costs : nullptr;
	q.num_launches = n;
	q.total = m.tiles;
	q.rate_x = m.rate_x;
	q.rate_y = m.rate_y;
	q.waves_per_group = tuning_.waves_per_group;
	if (planned_ && !collecting_)
	{
		// The order of an earlier planning replay: a dispatch plan needs nothing more, a queue
		// only its pop counter reset.
		if (dispatch)
		{
			if (plan.grid && pgs_k_shade_ordered(&q, plan.grid, stream_) != 0)
				throw std::runtime_error("ordered shading launch failed");
			return;
		}
		if (pgs_k_fill_u32(counter, 1, 0u, stream_) != 0 || pgs_k_shade_queue(&q, stream_) != 0)
			throw std::runtime_error("work-queue shading launch failed");
		return;
	}

	// Features of every tile (the queue holds only tiles with work), the model's predictions,
	// the queue, then the resident-workgroup shading.
	pgs_features_args f{};
	f.args = args;
	f.prefix = tile_prefix;
	f.features = features;
	f.tile_waves = tile_waves;
	f.num_launches = n;
	f.total = m.tiles;
	if (pgs_k_tile_features(&f, stream_) != 0)
		throw std::runtime_error("features launch failed");
	const bool learned = (schedule_ == Schedule::QueueLearned || schedule_ == Schedule::DispatchLearned) && !collecting_;
	if (learned)
		model_->predict_device(features, m.tiles, pred, stream_);
	pgs_order_args o{};
	o.features = features;
	o.predicted = learned ? pred : nullptr;
	o.order = order;
	o.counter = counter;
	o.count = m.tiles;
	o.slots = pgs_queue_slots(m.rate_x, m.rate_y, tuning_.waves_per_group, m.tiles);
	o.capacity = order_entries(m.tiles, m.waves);
	o.tile_waves = tile_waves;
	o.run_start = run_start;
	o.run_weight = run_weight;
	o.group_waves = group_waves;
	o.dispatch = dispatch ? 1u : 0u;
	if (dispatch)
	{
		// The host places the chunks per XCD (plan_dispatch) from the features and predictions.
		plan_dispatch(m, plan, features, learned ? pred : nullptr, tile_waves, counter, group_waves);
		q.order = plan.order;
		if (plan.grid && pgs_k_shade_ordered(&q, plan.grid, stream_) != 0)
			throw std::runtime_error("ordered shading launch failed");
		return;
	}
	if (pgs_k_queue_order(&o, stream_) != 0)
		throw std::runtime_error("order launch failed");
	if (collecting_ && pgs_k_fill_u32(costs, m.tiles, 0u, stream_) != 0)
		throw std::runtime_error("cost clear failed");
	if (pgs_k_shade_queue(&q, stream_) != 0)
		throw std::runtime_error("work-queue shading launch failed");
}

uint32_t Batch::collect_costs(float *features_out, float *log2_costs, uint32_t max_items)
{
	collecting_ = true;
	try
	{
		restore(stream_);
		for (auto &m : merged_)
			launch(m);
		ok(hipStreamSynchronize(stream_));
	}
	catch (...)
	{
		collecting_ = false;
		throw;
	}
	collecting_ = false;
	// The collection replay rebuilt the queues in tile order: plan again before the next replay.
	planned_ = false;
	destroy_graph();
	uint32_t written = 0;
	std::vector<float> f;
	std::vector<uint32_t> c;
	for (auto &m : merged_)
	{
		if (!(m.kind == Renderer::LaunchKind::Shade || m.kind == Renderer::LaunchKind::ShadeSuper) || m.tiles == 0)
			continue;
		const uint8_t *base = sched_ + m.sched_offset;
		f.resize(size_t(m.tiles) * PGS_MLP_INPUTS);
		c.resize(m.tiles);
		ok(hipMemcpy(f.data(), base, f.size() * sizeof(float), hipMemcpyDeviceToHost));
		ok(hipMemcpy(c.data(), base + size_t(m.tiles) * (PGS_MLP_INPUTS + 1u) * 4u, c.size() * 4u,
		             hipMemcpyDeviceToHost));
		// Only tiles that were shaded carry a measurement (feature 0 is their primitive count).
		for (uint32_t i = 0; i < m.tiles && written < max_items; i++)
		{
			if (f[size_t(i) * PGS_MLP_INPUTS] == 0.0f)
				continue;
			memcpy(features_out + size_t(written) * PGS_MLP_INPUTS, &f[size_t(i) * PGS_MLP_INPUTS],
			       PGS_MLP_INPUTS * sizeof(float));
			log2_costs[written] = std::log2(float(std::max(c[i], 1u)));
			written++;
		}
	}
	return written;
}

bool Batch::schedule_stats(ScheduleStats &out)
{
	out = {};
	if (schedule_ == Schedule::Hardware || !planned_)
		return false;
	const bool dispatch = schedule_ == Schedule::DispatchStatic || schedule_ == Schedule::DispatchLearned;
	for (size_t k = 0; k < merged_.size(); k++)
	{
		const Merged &m = merged_[k];
		if (!(m.kind == Renderer::LaunchKind::Shade || m.kind == Renderer::LaunchKind::ShadeSuper) || m.tiles == 0)
			continue;
		out.shading_launches++;
		if (dispatch)
		{
			out.tiles += plans_[k].tiles;
			out.heavy_first += plans_[k].heavy;
			out.entries += plans_[k].grid;
			continue;
		}
		// Queue: weights per tile (run_weight) against the base items of its waves.
		const uint8_t *base = sched_ + m.sched_offset;
		const uint32_t *tile_waves = reinterpret_cast<const uint32_t *>(base + size_t(m.tiles) * (PGS_MLP_INPUTS + 2u) * 4u);
		const uint32_t *run_weight = tile_waves + 2u * size_t(m.tiles);
		std::vector<uint32_t> w(m.tiles), rw(m.tiles);
		ok(hipMemcpy(w.data(), tile_waves, m.tiles * 4u, hipMemcpyDeviceToHost));
		ok(hipMemcpy(rw.data(), run_weight, m.tiles * 4u, hipMemcpyDeviceToHost));
		for (uint32_t i = 0; i < m.tiles; i++)
		{
			if (!rw[i])
				continue;
			out.tiles++;
			out.entries += rw[i];
			const uint32_t base_items = std::min(16u, std::max(1u, (w[i] + 15u) / 16u));
			out.split_tiles += rw[i] > base_items ? 1u : 0u;
		}
	}
	return true;
}

uint32_t Batch::num_launches() const
{
	uint32_t n = 0;
	for (const Merged &m : merged_)
		n += (schedule_ == Schedule::Hardware && m.in_run) ? 0u : 1u;
	return n;
}

bool Batch::issues(const Merged &m) const
{
	if (schedule_ == Schedule::Hardware && !collecting_)
	{
		if (m.in_run)
			return false;
		if (m.run >= 2)
			return m.run_total != 0;
	}
	return m.total != 0;
}

void Batch::launch(const Merged &m)
{
	if (!issues(m))
		return;
	using K = Renderer::LaunchKind;
	const void *args = dev_ + m.args_offset;
	const uint32_t *prefix = reinterpret_cast<const uint32_t *>(dev_ + m.prefix_offset);
	const uint32_t n = uint32_t(m.refs.size());
	int rc = 0;
	switch (m.kind)
	{
	case K::VramCopy: rc = pgs_multi_vram_copy(static_cast<const pgs_vram_copy_args *>(args), prefix, n, m.total, stream_); break;
	case K::VramSnapshot:
		rc = pgs_multi_vram_snapshot(static_cast<const pgs_vram_snapshot_args *>(args), prefix, n, m.total, stream_);
		break;
	case K::Setup: rc = pgs_multi_triangle_setup(static_cast<const pgs_setup_args *>(args), prefix, n, m.total, stream_); break;
	case K::Binning: rc = pgs_multi_binning(static_cast<const pgs_bin_args *>(args), prefix, n, m.total, stream_); break;
	case K::Shade: case K::ShadeSuper:
		launch_shade(m);
		break;
	case K::ClutWrite: rc = pgs_multi_clut_write(static_cast<const pgs_clut_args *>(args), prefix, n, m.total, stream_); break;
	case K::TexDecode:
		rc = pgs_multi_texture_decode(static_cast<const pgs_tex_decode_args *>(args), prefix, n, m.total, stream_);
		break;
	case K::SsaaHeuristic:
		rc = pgs_multi_ssaa_heuristic(static_cast<const pgs_ssaa_heuristic_args *>(args), prefix, n, m.total, stream_);
		break;
	case K::SamplerFeedback:
		rc = pgs_multi_sampler_feedback(static_cast<const pgs_feedback_args *>(args), prefix, n, m.total, stream_);
		break;
	case K::Fork: case K::Join: case K::Fused: break;
	}
	if (rc != 0)
		throw std::runtime_error("multi-launch failed");
}

void Batch::restore(hipStream_t stream)
{
	if (restore_dev_)
	{
		pgs_restore_args r{ restore_dev_, restore_count_, 0 };
		if (pgs_k_restore(&r, stream) != 0)
			throw std::runtime_error("restore launch failed");
		return;
	}
	for (Renderer *r : contexts_)
		r->restore_snapshot(stream);
}

int Batch::replay(bool use_graph)
{
	try
	{
		auto issue = [&]() {
			restore(stream_);
			for (auto &m : merged_)
				launch(m);
		};
		if (schedule_ != Schedule::Hardware && !planned_)
		{
			// Planning replay (eager): builds every shading launch's order once.
			if (plans_.size() != merged_.size())
				plans_.resize(merged_.size());
			issue();
			ok(hipStreamSynchronize(stream_));
			planned_ = true;
			destroy_graph();
			return 0;
		}
		if (!use_graph)
		{
			issue();
			return 0;
		}
		if (!graph_exec_)
		{
			ok(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
			issue();
			ok(hipStreamEndCapture(stream_, &graph_));
			ok(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
		}
		ok(hipGraphLaunch(graph_exec_, stream_));
		return 0;
	}
	catch (const std::exception &)
	{
		return -int(hipErrorLaunchFailure);
	}
}

void Batch::sync()
{
	ok(hipStreamSynchronize(stream_));
}

// Autotuner: every candidate (tile size of single-sampled passes x waves per shading
// workgroup) is applied and timed over `reps` graph replays after one warm-up replay; the
// fastest is kept. Candidates whose tile lists do not fit are skipped.
bool Batch::autotune(uint32_t reps, std::vector<TuningResult> *table)
{
	static const uint32_t tiles[] = { 0, 3, 4, 5, 6 };
	static const uint32_t waves[] = { 1, 2, 4 };
	hipEvent_t e0, e1;
	ok(hipEventCreate(&e0));
	ok(hipEventCreate(&e1));
	BatchTuning best = tuning_;
	float best_ms = -1.0f;
	for (uint32_t tl : tiles)
		for (uint32_t w : waves)
		{
			BatchTuning t;
			t.tile_log2_1x = tl;
			t.waves_per_group = w;
			if (!apply(t))
				continue;
			if (replay(true) != 0)
				return false;
			ok(hipEventRecord(e0, stream_));
			for (uint32_t r = 0; r < reps; r++)
				if (replay(true) != 0)
					return false;
			ok(hipEventRecord(e1, stream_));
			ok(hipEventSynchronize(e1));
			float ms = 0.0f;
			ok(hipEventElapsedTime(&ms, e0, e1));
			ms /= float(std::max(reps, 1u));
			if (table)
				table->push_back({ t, ms });
			if (best_ms < 0.0f || ms < best_ms)
			{
				best_ms = ms;
				best = t;
			}
		}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	return apply(best);
}

void Batch::profile(double *ms, uint64_t *launches, uint32_t kinds)
{
	for (uint32_t k = 0; k < kinds; k++)
	{
		ms[k] = 0.0;
		launches[k] = 0;
	}
	std::vector<std::pair<uint32_t, std::pair<hipEvent_t, hipEvent_t>>> ev;
	{
		hipEvent_t a, b;
		ok(hipEventCreate(&a));
		ok(hipEventCreate(&b));
		ok(hipEventRecord(a, stream_));
		restore(stream_);
		ok(hipEventRecord(b, stream_));
		ev.push_back({ Renderer::kProfileRestore, { a, b } });
	}
	for (auto &m : merged_)
	{
		if (!issues(m))
			continue;
		hipEvent_t a, b;
		ok(hipEventCreate(&a));
		ok(hipEventCreate(&b));
		ok(hipEventRecord(a, stream_));
		launch(m);
		ok(hipEventRecord(b, stream_));
		ev.push_back({ uint32_t(m.kind), { a, b } });
	}
	ok(hipStreamSynchronize(stream_));
	for (auto &e : ev)
	{
		float t = 0.0f;
		ok(hipEventElapsedTime(&t, e.second.first, e.second.second));
		if (e.first < kinds)
		{
			ms[e.first] += t;
			launches[e.first]++;
		}
		(void)hipEventDestroy(e.second.first);
		(void)hipEventDestroy(e.second.second);
	}
}
}
