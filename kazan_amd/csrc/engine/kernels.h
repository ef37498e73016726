// kernels.h — host launchers of the prebuilt, pipeline-independent kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"

hipError_t kz_launch_fill_u32(hipStream_t st, uint32_t *dst, uint64_t n, uint32_t value);
hipError_t kz_launch_fill_rect_u32(hipStream_t st, uint32_t *dst, uint32_t width, int32_t x0, int32_t y0, uint32_t w,
                                   uint32_t h, uint32_t spp, uint32_t value);
hipError_t kz_launch_bin_scan(hipStream_t st, const KzFrame &f, uint32_t ntiles);
hipError_t kz_launch_bin_scatter(hipStream_t st, const KzFrame &f);
hipError_t kz_launch_bin_sort(hipStream_t st, const KzFrame &f, uint32_t ntiles);
hipError_t kz_launch_stats_reduce(hipStream_t st, unsigned long long *stats);
hipError_t kz_launch_resolve4(hipStream_t st, const uint32_t *src, uint32_t *dst, uint32_t width, int32_t x0,
                              int32_t y0, uint32_t w, uint32_t h);
