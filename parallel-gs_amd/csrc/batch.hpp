// Batched replay of many GS contexts (no counterpart in the reference, which drives one
// context per Vulkan device queue).
//
// Each context records its work once (Renderer::begin_recording / end_recording). A Batch
// merges the recordings: it repeatedly takes the kernel kind that the most contexts have as
// their next recorded launch and issues one multi-launch for all of them, so every context's
// launches keep their order while 64 contexts cost about as many launches as one. Replay
// restores each context's VRAM snapshot first, so every replay redoes the recorded work, and
// the whole sequence can be captured into one hipGraph.
//
// Tuning knobs applied at build time (an autotuner times candidates): the coarse tile size of
// single-sampled passes (free there; super-sampled passes keep the reference's grid because it
// decides per-tile shading) and the waves per shading workgroup.
#pragma once
#include "renderer.hpp"
#include "cost_model.hpp"
#include <vector>

namespace pgs
{
struct BatchTuning
{
	uint32_t tile_log2_1x = 0;    // 0: keep each pass's recorded tile size; else 3..6
	uint32_t waves_per_group = 4; // 1, 2 or 4
};

// How a batch runs its shading launches: the hardware dispatcher (one multi-launch), or a
// work queue of resident workgroups in static order or in the order of predicted cost.
enum class Schedule : uint32_t { Hardware = 0, QueueStatic = 1, QueueLearned = 2, DispatchStatic = 3, DispatchLearned = 4 };

struct TuningResult
{
	BatchTuning tuning;
	float ms; // per replay
};

class Batch
{
public:
	Batch() = default;
	~Batch();
	Batch(const Batch &) = delete;
	Batch &operator=(const Batch &) = delete;

	// merge = false keeps every launch separate, context after context (the single-stream baseline).
	bool build(Renderer *const *contexts, uint32_t count, bool merge = true);
	// Rewrites the launch arguments for a tuning (no recording needed). Returns false if a
	// context's tile lists would not fit.
	bool apply(const BatchTuning &t);
	int replay(bool use_graph);
	bool autotune(uint32_t reps, std::vector<TuningResult> *table);
	// One non-graph replay with HIP events around every merged launch: summed milliseconds and
	// launch counts per kernel kind (Renderer::LaunchKind numbering).
	void profile(double *ms, uint64_t *launches, uint32_t kinds);
	// QueueLearned needs a model (kept by the caller for the batch's lifetime).
	bool set_schedule(Schedule s, const CostModel *model);
	// One replay with the static work queue that measures every shading workgroup: features
	// (PGS_MLP_INPUTS floats) and log2 shader cycles per item. Returns the item count.
	uint32_t collect_costs(float *features, float *log2_costs, uint32_t max_items);
	// What the current schedule's plan holds (after the planning replay): tiles with work,
	// queue items or dispatch workgroups, queue tiles split beyond their base items, dispatch
	// tiles started first. False before the first replay of a planned schedule.
	struct ScheduleStats
	{
		uint32_t shading_launches = 0, tiles = 0, entries = 0, split_tiles = 0, heavy_first = 0;
	};
	bool schedule_stats(ScheduleStats &out);
	void sync();
	hipStream_t stream() const { return stream_; }
	// Kernel launches per replay (a shading run counts once under the hardware schedule).
	uint32_t num_launches() const;
	uint32_t num_contexts() const { return uint32_t(contexts_.size()); }
	const BatchTuning &tuning() const { return tuning_; }

private:
	struct Ref { uint32_t context, index; };
	struct Merged
	{
		Renderer::LaunchKind kind;
		uint32_t rate_x = 0, rate_y = 0;
		std::vector<Ref> refs;
		size_t args_offset = 0, prefix_offset = 0;
		uint32_t total = 0;
		size_t sched_offset = 0;       // shading launches: features, predictions, order, costs, counter
		size_t tile_prefix_offset = 0; // shading launches: coarse-tile prefix over the launches
		uint32_t tiles = 0;
		uint32_t waves = 0; // shading launches: waves of all tiles (bound on dispatch-order entries)
		// Runs of adjacent shading launches over disjoint contexts: the first launch of a run
		// holds the run's combined argument blocks and prefix (run = launches in the run, >= 2);
		// the others are marked in_run. The hardware schedule issues a run as one launch.
		uint32_t run = 0, run_n = 0, run_total = 0;
		bool in_run = false;
		bool run_all_grids = false; // the run has an 8x or 16x launch
		size_t run_args_offset = 0, run_prefix_offset = 0;
	};

	std::vector<Renderer *> contexts_;
	std::vector<std::vector<Renderer::Launch>> launches_; // per context, after tuning
	std::vector<Merged> merged_;
	BatchTuning tuning_;
	hipStream_t stream_ = nullptr;
	uint8_t *dev_ = nullptr;       // argument blocks + workgroup prefixes of every merged launch
	pgs_copy_range *restore_dev_ = nullptr; // every context's restore ranges, one launch
	void restore(hipStream_t stream);
	uint32_t restore_count_ = 0;
	size_t dev_bytes_ = 0;
	hipGraph_t graph_ = nullptr;
	hipGraphExec_t graph_exec_ = nullptr;
	Schedule schedule_ = Schedule::Hardware;
	const CostModel *model_ = nullptr;
	uint8_t *sched_ = nullptr;    // work-queue buffers of the shading launches
	size_t sched_bytes_ = 0;
	bool collecting_ = false;
	// A schedule's queue or dispatch order depends only on the recording, the tuning, the
	// schedule and the model, so it is built once (a planning replay) and replays then only
	// shade. Dispatch orders are placed per XCD on the host (plans_, one per merged launch).
	bool planned_ = false;
	struct Plan
	{
		uint32_t *order = nullptr; // device: grid entries, entry k of XCD x at 8 k + x
		uint32_t grid = 0;
		size_t capacity = 0;
		uint32_t tiles = 0;        // tiles with work
		uint32_t heavy = 0;        // learned: tiles started first on their XCD
	};
	std::vector<Plan> plans_;
	void free_plans();
	void plan_dispatch(const Merged &m, Plan &plan, const float *features, const float *pred, const uint32_t *tile_waves,
	                   uint32_t *counter, uint32_t group_waves);

	bool upload();
	void destroy_graph();
	// Whether launch(m) issues a kernel (shading runs are issued by their first launch).
	bool issues(const Merged &m) const;
	void launch(const Merged &m);
	void launch_shade(const Merged &m);
};
}
