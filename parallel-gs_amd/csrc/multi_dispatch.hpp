// Host-side grid sizes and multi-launch entry points of every kernel (defined next to the
// kernels). A multi-launch runs several launches of one kernel as one HIP launch.
#pragma once
#include "pgs_kernels.h"
#include <cstdint>

uint32_t pgs_grid_vram_copy(const pgs_vram_copy_args &a);
uint32_t pgs_grid_vram_snapshot(const pgs_vram_snapshot_args &a);
uint32_t pgs_grid_triangle_setup(const pgs_setup_args &a);
uint32_t pgs_grid_ssaa_heuristic(const pgs_ssaa_heuristic_args &a);
uint32_t pgs_grid_binning(const pgs_bin_args &a);
uint32_t pgs_grid_shade(const pgs_shade_args &a);
uint32_t pgs_grid_clut_write(const pgs_clut_args &a);
uint32_t pgs_grid_texture_decode(const pgs_tex_decode_args &a);
uint32_t pgs_grid_sampler_feedback(const pgs_feedback_args &a);

int pgs_multi_vram_copy(const pgs_vram_copy_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                        hipStream_t stream);
int pgs_multi_vram_snapshot(const pgs_vram_snapshot_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                            hipStream_t stream);
int pgs_multi_triangle_setup(const pgs_setup_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                             hipStream_t stream);
int pgs_multi_ssaa_heuristic(const pgs_ssaa_heuristic_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                             hipStream_t stream);
int pgs_multi_binning(const pgs_bin_args *args, const uint32_t *prefix, uint32_t n, uint32_t total, hipStream_t stream);
// Set in the x rate of a shading multi-launch, queue or dispatch order whose launches sample
// super-sampled textures: the kernels instantiated with that sampling path are used.
constexpr uint32_t kRateArrays = 1u << 8;
int pgs_multi_shade(const pgs_shade_args *args, const uint32_t *prefix, uint32_t n, uint32_t total, uint32_t rate_x,
                    uint32_t rate_y, uint32_t waves, hipStream_t stream);
// Shading launches of different sampling grids in one multi-launch: all_grids false allows only
// 1x, 2x and 4x launches, true also 8x and 16x.
int pgs_multi_shade_mixed(const pgs_shade_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                          uint32_t waves, bool all_grids, hipStream_t stream);
// Resident workgroups the work-queue shading launch of these rates uses (0: unsupported).
uint32_t pgs_queue_slots(uint32_t rate_x, uint32_t rate_y, uint32_t waves_per_group, uint32_t total);
int pgs_multi_clut_write(const pgs_clut_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                         hipStream_t stream);
int pgs_multi_texture_decode(const pgs_tex_decode_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                             hipStream_t stream);
int pgs_multi_sampler_feedback(const pgs_feedback_args *args, const uint32_t *prefix, uint32_t n, uint32_t total,
                               hipStream_t stream);
