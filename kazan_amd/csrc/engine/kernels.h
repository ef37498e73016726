// kernels.h — host launchers of the prebuilt, pipeline-independent kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"

hipError_t kz_launch_fill_u32(hipStream_t st, uint32_t *dst, uint64_t n, uint32_t value);
hipError_t kz_launch_fill_rect_u32(hipStream_t st, uint32_t *dst, uint32_t width, int32_t x0, int32_t y0, uint32_t w,
                                   uint32_t h, uint32_t spp, uint32_t value);
// Exclusive scan of counts[0, n) (n read from n_dev when it is not null).
hipError_t kz_launch_scan(hipStream_t st, const uint32_t *counts, uint32_t *offsets, uint32_t *cursor, uint32_t n,
                          const uint32_t *n_dev, uint32_t capacity, unsigned long long *stats);
hipError_t kz_launch_bin_scatter(hipStream_t st, const KzFrame &f);
hipError_t kz_launch_sort_lists(hipStream_t st, const uint32_t *counts, const uint32_t *offsets, uint32_t n,
                                const uint32_t *n_dev, uint32_t *lists, uint32_t *tmp, uint32_t capacity);
// Balanced mode: split levels + item bases (item total -> ctrl[1]), item
// geometry, per-item binning, and the most-expensive-first queue.
hipError_t kz_launch_plan(hipStream_t st, const KzFrame &f, uint32_t ntiles, uint32_t *ctrl);
hipError_t kz_launch_plan_emit(hipStream_t st, const KzFrame &f, uint32_t ntiles);
hipError_t kz_launch_item_count(hipStream_t st, const KzFrame &f);
hipError_t kz_launch_item_scatter(hipStream_t st, const KzFrame &f);
hipError_t kz_launch_queue(hipStream_t st, const KzFrame &f, uint32_t *ctrl, uint32_t max_items);
hipError_t kz_launch_stats_reduce(hipStream_t st, unsigned long long *stats);
hipError_t kz_launch_resolve4(hipStream_t st, const uint32_t *src, uint32_t *dst, uint32_t width, int32_t x0,
                              int32_t y0, uint32_t w, uint32_t h);
