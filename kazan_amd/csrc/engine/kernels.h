// kernels.h — host launchers of the prebuilt, pipeline-independent kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"

hipError_t kz_launch_fill_u32(hipStream_t st, uint32_t *dst, uint64_t n, uint32_t value);
// Per-tile list state before a batch's binning: counts 0, offsets all ones,
// cut flags 0 (flags may be null).
hipError_t kz_launch_fill_tiles(hipStream_t st, uint32_t *counts, uint32_t *offsets, uint32_t *flags, uint32_t n);
hipError_t kz_launch_fill_rect_u32(hipStream_t st, uint32_t *dst, uint32_t width, int32_t x0, int32_t y0, uint32_t w,
                                   uint32_t h, uint32_t spp, uint32_t value);
// Exclusive scan of counts[0, n) into offsets and cursor (or null); with
// stats, adds the total to the list-entry statistic and flags an overflow
// beyond `capacity` (sort binning; radix binning's last pass writes the
// offsets itself).
hipError_t kz_launch_scan(hipStream_t st, uint32_t *counts, uint32_t *offsets, uint32_t *cursor, uint32_t n,
                          uint32_t capacity, unsigned long long *stats);
// Radix binning, after kz_bin_count wrote f.tri_rect and f.block_pairs.
struct KzRadixScratch {
    uint32_t *chunk_offsets;  // [ceil(tris / KZ_PAIR_CHUNK)]
    uint32_t *ctrl;           // KZ_BIN_WORDS
    uint32_t *tile_flag;      // [tiles], cleared with the counts before a batch
    uint32_t *keys[2], *vals[2];  // [list capacity] each
    uint32_t *table;          // [kz_radix_digits_max() * tstride]
    uint32_t *totals;         // [kz_radix_digits_max()]
    uint32_t tstride;         // kz_radix_blocks(list capacity)
};
uint32_t kz_radix_blocks(uint64_t capacity);
uint32_t kz_radix_digits_max();
// counting-sort passes for `ntiles` keys
uint32_t kz_radix_passes(uint32_t ntiles);
// Pair offsets per chunk and the pairs in primitive order (2 launches).
hipError_t kz_launch_bin_pairs(hipStream_t st, const KzFrame &f, const KzRadixScratch &s);
// The passes, the last into f.tile_lists with f.tile_counts (zeroed by the
// caller); with f.tile_acc (zeroed by the caller), the pair statistics
// (3 launches per pass).
hipError_t kz_launch_bin_radix(hipStream_t st, const KzFrame &f, const KzRadixScratch &s, uint32_t ntiles);
// Binning from the per-triangle pixel boxes: pairs per tile, then the lists.
hipError_t kz_launch_bin_hist(hipStream_t st, const KzFrame &f);
hipError_t kz_launch_bin_scatter(hipStream_t st, const KzFrame &f);
// Cost model: pairs per tile plus their statistics (f.tile_acc, cleared by
// the caller), the per-tile features (f.tile_feat), the MLP (features ->
// predicted fragments per tile).
hipError_t kz_launch_bin_hist_acc(hipStream_t st, const KzFrame &f);
hipError_t kz_launch_tile_features(hipStream_t st, const KzFrame &f, uint32_t ntiles);
hipError_t kz_launch_cost_mlp(hipStream_t st, const KzCostModel *model, const float *feat, uint32_t ntiles,
                              float *pred);
// p.steps Adam steps on state = [W | M | V] (KZ_PARAMS each) against
// targets y[n] from normalised features x[n][16]; loss_out[steps] or null.
hipError_t kz_launch_cost_train(hipStream_t st, float *state, const float *x, const float *y, const KzTrainParams &p,
                                float *loss_out);
// Training data: the Gaussian-cluster scene of `seed` (nc of n triangles in
// the cluster) as 3 * n pos4 + colour4 vertices.
hipError_t kz_launch_gen_cluster(hipStream_t st, float *vertices, uint32_t n, uint32_t nc, uint32_t width,
                                 uint32_t height, double sigma_frac, double radius, uint64_t seed);
hipError_t kz_launch_sort_lists(hipStream_t st, const KzFrame &f, uint32_t ntiles);
// Balanced mode: split levels, split slots and the ordered item queue
// (3 launches over one workgroup per 256 tiles; `table` holds
// kz_plan_table_words(ntiles) words of per-workgroup counts), then the
// sub-tile lists of split tiles on `grid` workgroups.
uint32_t kz_plan_table_words(uint32_t ntiles);
// parts_ready: kz_launch_cost_plan wrote the workgroups' cost parts
hipError_t kz_launch_plan(hipStream_t st, const KzFrame &f, uint32_t ntiles, uint32_t *table, bool parts_ready);
// The balanced form's cost pass with a model: features, predictions and the
// plan's cost parts in one launch (table: the plan's table)
hipError_t kz_launch_cost_plan(hipStream_t st, const KzCostModel *model, const KzFrame &f, uint32_t ntiles,
                               uint32_t *table);
hipError_t kz_launch_split(hipStream_t st, const KzFrame &f, uint32_t grid, uint32_t ntiles);
hipError_t kz_launch_stats_reduce(hipStream_t st, unsigned long long *stats);
hipError_t kz_launch_resolve4(hipStream_t st, const uint32_t *src, uint32_t *dst, uint32_t width, int32_t x0,
                              int32_t y0, uint32_t w, uint32_t h);
