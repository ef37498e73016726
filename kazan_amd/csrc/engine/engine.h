// engine.h — the draw engine behind vkQueueSubmit.
//
// The Vulkan layer (csrc/icd) records commands into a list of Ops; a
// submission replays that list on the device's single in-order HIP stream.
// Consecutive draws that share a pipeline, dynamic state and descriptor
// bindings are recorded as one Batch: one binning pass and one raster launch
// for all of them, whatever their number (this is how a command buffer of
// thousands of draws becomes a handful of launches).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "device_common.h"

namespace kz {

// Fixed-function state baked into a pipeline's compiled kernels.
struct FixedState {
    uint32_t tile_shift = 4;
    uint32_t samples = 1;
    uint32_t num_varyings = 0;
    uint32_t cull_mode = 0, front_face = 0;
    uint32_t depth_test = 0, depth_write = 0, depth_op = 7;
    uint32_t blend = 0;
    uint32_t src_color = 1, dst_color = 0, color_op = 0, src_alpha = 1, dst_alpha = 0, alpha_op = 0;
    uint32_t write_mask = 15;
};

// A pipeline's device code: kz_bin_count and kz_raster for one FixedState.
struct PipelineKernels {
    hipModule_t module = nullptr;
    hipFunction_t bin_count = nullptr;
    hipFunction_t raster = nullptr;             // one workgroup per tile
    hipFunction_t raster_persistent = nullptr;  // pops the work-item queue
    uint32_t persistent_blocks = 0;             // resident workgroups that fill the device
};

// Builds the full kernel source (prelude + shared structs + generated shader
// functions + the raster template) and compiles it to a gfx950 code object,
// going through the on-disk code-object cache. Needs no GPU.
bool compile_pipeline_code(const FixedState &fs, const std::string &shader_source, std::vector<char> &code,
                           std::string &log, bool *cache_hit);
std::string pipeline_full_source(const FixedState &fs, const std::string &shader_source);
// Loads the code object on the current device.
bool load_pipeline_kernels(const std::vector<char> &code, PipelineKernels &out, std::string &err);

struct Op {
    enum Kind { Fill, Batch, Copy, Copy2D, Resolve } kind;
    // Fill: ptr = attachment base, rectangle in pixels, spp words per pixel
    uint32_t *fill_ptr = nullptr;
    uint32_t fill_width = 0, fill_spp = 1, fill_value = 0;
    int32_t x0 = 0, y0 = 0;
    uint32_t w = 0, h = 0;
    // Batch
    const PipelineKernels *kernels = nullptr;
    uint32_t tile_shift = 4;
    KzFrame frame{};                // draws/scratch pointers filled at submit
    std::vector<KzDraw> draws;      // host copy, uploaded at end of recording
    KzDraw *d_draws = nullptr;      // device copy owned by the command buffer
    // Copy: linear bytes; Copy2D: pitched rows
    const void *src = nullptr;
    void *dst = nullptr;
    size_t bytes = 0;
    size_t src_pitch = 0, dst_pitch = 0, row_bytes = 0, rows = 0;
    // Resolve: 4x interleaved src -> 1x dst
    const uint32_t *rsrc = nullptr;
    uint32_t *rdst = nullptr;
    uint32_t rwidth = 0;
};

// Kernels of a batch, in launch order, for per-kernel timing.
enum KernelSlot { KS_CLEAR_COUNTS = 0, KS_BIN_COUNT, KS_BIN_SCAN, KS_BIN_SCATTER, KS_BIN_SORT, KS_RASTER, KS_COUNT };

struct SubmitStats {
    unsigned long long counters[KZ_STAT_COUNT] = {};
    float gpu_ms = 0.0f;
    uint32_t batches = 0, launches = 0;
    // with kernel timing on: summed milliseconds and launch count per slot
    float kernel_ms[KS_COUNT] = {};
    uint32_t kernel_launches[KS_COUNT] = {};
};

class Engine {
  public:
    ~Engine();
    bool init(int hip_device, std::string &err);
    hipStream_t stream() const { return stream_; }
    int device() const { return device_; }

    // Replays ops on the stream. Returns hipSuccess or the first error.
    hipError_t begin_submit();
    hipError_t execute(const std::vector<Op> &ops);
    hipError_t end_submit(hipEvent_t fence_event);
    // After the submission's work is complete: read stats, adapt capacity.
    void collect_stats();
    const SubmitStats &last_stats() const { return last_; }

    uint32_t list_capacity() const { return list_capacity_; }

    // Pipeline form, fixed at device creation from KAZAN_PIPELINE:
    //   "plain"    every vkCmdDraw* is its own batch, one raster workgroup
    //              per 16x16 tile in tile order;
    //   "balanced" (default) consecutive compatible draws share one batch,
    //              tiles are split into work items by predicted cost and a
    //              persistent raster kernel takes the most expensive first.
    bool balanced() const { return balanced_; }
    bool merge_draws() const { return balanced_; }

    // Per-kernel timing: an event pair around every batch kernel on the
    // engine's stream. Off by default (it adds a little latency per launch).
    void set_kernel_timing(bool on) { kernel_timing_ = on; }

  private:
    hipError_t ensure_scratch(uint32_t tiles, uint32_t tris);
    hipError_t run_batch(const Op &op);
    hipError_t mark(int slot, bool end);

    bool kernel_timing_ = false;
    bool balanced_ = true;
    uint64_t list_limit_ = 0;
    // balanced-mode scratch, sized for 16 work items per tile
    uint32_t *tile_level_ = nullptr, *tile_item_base_ = nullptr, *items_ = nullptr, *item_counts_ = nullptr,
             *item_offsets_ = nullptr, *item_cursor_ = nullptr, *queue_ = nullptr, *ctrl_ = nullptr;
    uint32_t cap_items_tiles_ = 0;
    hipError_t ensure_item_scratch(uint32_t tiles);
    std::vector<std::pair<int, hipEvent_t>> timing_events_;  // (slot, start) then end, alternating
    std::vector<hipEvent_t> event_pool_;
    size_t events_used_ = 0;

    int device_ = -1;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_start_ = nullptr, ev_end_ = nullptr;
    uint32_t *tile_counts_ = nullptr, *tile_offsets_ = nullptr, *tile_cursor_ = nullptr;
    uint32_t *tile_lists_ = nullptr, *sort_tmp_ = nullptr;
    uint64_t *tri_rect_ = nullptr;
    unsigned long long *stats_d_ = nullptr, *stats_h_ = nullptr;
    uint32_t cap_tiles_ = 0, cap_tris_ = 0, list_capacity_ = 0;
    uint32_t pending_batches_ = 0, pending_launches_ = 0;
    bool submit_open_ = false;
    SubmitStats last_;
};

}  // namespace kz
