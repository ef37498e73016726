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
    bool optimize = true;  // -O3 (false: -O0, the backend API's NoOptimizations)
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
// The code-object cache: directory, and the file name of a pipeline
// variant's code object (the FNV-1a hash of its complete source, the
// compile options and the hiprtc version).
std::string pipeline_code_dir();
std::string pipeline_code_name(const FixedState &fs, const std::string &shader_source);
// With read_cache false the on-disk copy is ignored and replaced (a file
// that failed to load). A cached file that is not a gfx950 code object is
// never returned.
bool compile_pipeline_code(const FixedState &fs, const std::string &shader_source, std::vector<char> &code,
                           std::string &log, bool *cache_hit, bool read_cache = true);
// FNV-1a (64-bit) of a byte range: cache file names and cache-data entry
// checksums.
uint64_t fnv1a64(const void *data, size_t n, uint64_t h = 1469598103934665603ull);
// An ELF64 object for AMDGPU (e_machine 224): the shape of every code object
// hiprtc writes; anything else in a cache is damaged or foreign.
bool is_amdgpu_code_object(const char *data, size_t n);
std::string pipeline_full_source(const FixedState &fs, const std::string &shader_source);
// What one kernel of a gfx950 code object takes, from its kernel descriptor
// (symbol <kernel>.kd): VGPRs including AGPRs (allocation granules of 8),
// scratch bytes per lane, LDS bytes per workgroup, and the resident waves
// per SIMD those allow a one-wave workgroup (512 VGPRs per SIMD lane, 160 KB
// of LDS per CU). False if the code object has no such kernel.
struct KernelResources {
    uint32_t vgprs = 0, scratch = 0, lds = 0, waves = 0;
};
bool kernel_resources(const std::vector<char> &elf, const char *kernel, KernelResources &r);
// A compute pipeline's kernel (generate_compute_source's output with the
// shared device structs), compiled through the same code-object cache.
std::string compute_full_source(const std::string &compute_source);
bool compile_compute_code(const std::string &compute_source, bool optimize, std::vector<char> &code, std::string &log,
                          bool *cache_hit);
// Loads the code object on the current device.
bool load_pipeline_kernels(const std::vector<char> &code, PipelineKernels &out, std::string &err);

// Tile sizes (log2 of the side) a pipeline can run with.
constexpr uint32_t kMinTileShift = 3, kMaxTileShift = 6;

// LDS bytes of one raster workgroup (colour, depth, claims, fragment queue).
inline uint32_t raster_lds_bytes(uint32_t tile_shift, uint32_t samples) {
    const uint32_t tp = 1u << (2 * tile_shift);
    return tp * samples * 8 + tp * 4 + tp * 16;
}

// A pipeline's kernels for every tile size, compiled and loaded on first
// use (the autotuner tries several; code objects come from the cache).
class PipelineVariants {
  public:
    PipelineVariants() = default;
    PipelineVariants(const PipelineVariants &) = delete;
    PipelineVariants &operator=(const PipelineVariants &) = delete;
    ~PipelineVariants();
    FixedState fixed;  // tile_shift: the pipeline's default tile size
    std::string shader_source;
    // null when the tile size does not fit the workgroup's LDS or the
    // compile failed (err says why)
    const PipelineKernels *get(uint32_t tile_shift, std::string *err = nullptr);
    // A code object for one tile size supplied by a VkPipelineCache: get()
    // loads it before asking the on-disk cache or the compiler, and falls
    // back to them when it does not load.
    void preload(uint32_t tile_shift, std::vector<char> code) {
        if (tile_shift <= kMaxTileShift && !tried_[tile_shift]) preloaded_[tile_shift] = std::move(code);
    }
    bool fits(uint32_t tile_shift) const {
        return tile_shift >= kMinTileShift && tile_shift <= kMaxTileShift &&
               raster_lds_bytes(tile_shift, fixed.samples) <= 160u * 1024u;
    }

  private:
    PipelineKernels k_[kMaxTileShift + 1];
    std::vector<char> preloaded_[kMaxTileShift + 1];
    bool tried_[kMaxTileShift + 1] = {};
    std::string err_[kMaxTileShift + 1];
};

// How a batch runs: tile size, the persistent raster's grid divisor, and,
// in the balanced form, whether the plan uses the cost model and whether the
// raster is the persistent one over the planned queue (else one workgroup
// per tile, no plan).
struct BatchConfig {
    uint32_t shift = 4, grid_div = 1;
    bool model = true, persistent = true;
    uint32_t max_level = 2;  // deepest split (2: 16 sub-tiles, 1: 4, 0: none, LPT order only)
    // multiplies the model's fragment weight: the weight is fitted on the
    // training pipelines' fragment cost, and a pipeline whose fragments cost
    // more (a heavier fragment shader, more samples) needs a larger one
    uint32_t frag_scale = 1;
};

// Autotuning of one batch (balanced form): its first submissions time
// candidate configurations with HIP events, one candidate per submission;
// then the fastest is kept. Pixels never depend on the choice.
struct BatchTuning {
    BatchTuning() = default;
    BatchTuning(const BatchTuning &) = delete;
    BatchTuning &operator=(const BatchTuning &) = delete;
    ~BatchTuning();
    bool done = false;
    int phase = 0;  // 0: tile sizes, 1: grid sizes, 2: plan and raster forms, 3: split depth, 4: fragment weight,
                    // 5: every tile size x grid size again, 6: the three fastest twice more each
    std::vector<BatchConfig> cand;
    std::vector<float> ms;           // per candidate, -1 = not timed
    std::vector<std::vector<size_t>> finals;  // the finalists' runs (indices into cand)
    uint32_t next = 0;               // next candidate to try
    int pending = -1;                // candidate whose events are in flight
    hipEvent_t ev[2] = {nullptr, nullptr};
    BatchConfig choice;              // once done
};

struct Op {
    enum Kind { Fill, Batch, Copy, Copy2D, Resolve, Dispatch } kind;
    // Dispatch: a compute pipeline's kernel over groups[3] workgroups of
    // block[3] invocations, with its descriptors and push constants
    hipFunction_t cs_kernel = nullptr;
    uint32_t groups[3] = {0, 0, 0}, block[3] = {1, 1, 1};
    KzComputeFrame cframe{};
    // Fill: ptr = attachment base, rectangle in pixels, spp words per pixel
    uint32_t *fill_ptr = nullptr;
    uint32_t fill_width = 0, fill_spp = 1, fill_value = 0;
    int32_t x0 = 0, y0 = 0;
    uint32_t w = 0, h = 0;
    // Batch
    PipelineVariants *variants = nullptr;
    uint32_t tile_shift = 4;                // the pipeline's default tile size
    std::shared_ptr<BatchTuning> tuning;    // created at the first submission
    KzFrame frame{};                // draws/scratch pointers filled at submit
    std::vector<KzDraw> draws;      // host copy, uploaded at end of recording
    KzDraw *d_draws = nullptr;      // device copy owned by the command buffer
    uint32_t *d_draw_begin = nullptr;  // the draws' first triangles, in the same allocation
    uint32_t *d_block_draw = nullptr;  // then the draw of each 256-triangle block's first triangle
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
enum KernelSlot {
    KS_CLEAR_COUNTS = 0,
    KS_BIN_COUNT,
    KS_BIN_HIST,
    KS_BIN_SCAN,
    KS_BIN_SCATTER,
    KS_BIN_SORT,
    KS_PLAN,
    KS_RASTER,
    KS_COUNT
};

struct SubmitStats {
    unsigned long long counters[KZ_STAT_COUNT] = {};
    float gpu_ms = 0.0f;
    uint32_t batches = 0, launches = 0;
    // with kernel timing on: summed milliseconds and launch count per slot
    float kernel_ms[KS_COUNT] = {};
    uint32_t kernel_launches[KS_COUNT] = {};
};

// A command buffer's op list captured as a hipGraph, valid while the
// engine's scratch buffers and settings are those it was captured with.
struct GraphCache {
    GraphCache() = default;
    GraphCache(const GraphCache &) = delete;
    GraphCache &operator=(const GraphCache &) = delete;
    ~GraphCache() { reset(); }
    void reset();
    hipGraphExec_t exec = nullptr;
    uint64_t generation = 0;
    uint32_t batches = 0, launches = 0;
    uint32_t eager_runs = 0;  // a command buffer is captured from its second submission on
};

class Engine {
  public:
    ~Engine();
    bool init(int hip_device, std::string &err);
    hipStream_t stream() const { return stream_; }
    int device() const { return device_; }

    // Replays ops on the stream. Returns hipSuccess or the first error.
    // With a cache (a command buffer's), once every batch is tuned the ops
    // are captured into a hipGraph on first use and replayed as one graph
    // launch afterwards (KAZAN_GRAPHS=0 turns this off).
    hipError_t begin_submit();
    hipError_t execute(std::vector<Op> &ops, GraphCache *cache = nullptr);
    hipError_t end_submit(hipEvent_t fence_event);
    uint32_t graph_launches() const { return graph_launches_; }
    uint32_t graph_captures() const { return graph_captures_; }
    // the configuration the most recently finished tuning chose (shift 0 =
    // none yet)
    const BatchConfig &last_tuned() const { return last_tuned_; }
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

    // Learned cost model of the balanced form. Loaded at init from
    // $KAZAN_COST_MODEL, else cost_model.bin beside the driver library
    // ("none" disables it; without a model the plan splits by binned
    // triangle count). Returns false and leaves the model unchanged when
    // `blob` is not a valid KzCostModel.
    bool set_cost_model(const void *blob, size_t bytes);
    bool has_cost_model() const { return cost_model_d_ != nullptr; }
    const std::string &cost_model_source() const { return cost_model_src_; }

    // Tile profiling (training data for the cost model): each batch also
    // computes the per-tile features (and, with a model, its predictions),
    // and the static form records every tile's fragments and raster-
    // workgroup cycles. read_tile_profile copies the last batch's features
    // [tiles * KZ_FEATURES], fragments, cycles and predicted fragments
    // [tiles] (zeros where the batch did not produce them); any may be null.
    void set_tile_profile(bool on) {
        if (on != tile_profile_) ++generation_;
        tile_profile_ = on;
    }
    hipError_t read_tile_profile(uint32_t *tiles, float *features, float *frags, float *cycles, float *pred);

    // Autotuning of balanced batches (KAZAN_AUTOTUNE=0 turns it off: every
    // batch then runs at the pipeline's tile size and the full grid).
    bool autotune() const { return autotune_ && balanced_; }

  private:
    hipError_t ensure_scratch(uint32_t tiles, uint32_t tris);
    // the configuration a batch runs with this submission (tuning trial or
    // the tuned choice), and its scratch allocated for it
    hipError_t pick_config(Op &op, BatchConfig &cfg, bool &timed);
    BatchConfig default_config(const Op &op) const;
    hipError_t prepare_batch(const Op &op, uint32_t shift);
    hipError_t run_batch(const Op &op, const BatchConfig &cfg);
    hipError_t execute_eager(std::vector<Op> &ops);
    hipError_t launch_ops(std::vector<Op> &ops);  // tuned configurations, no tuning events
    hipError_t launch_ops_one(const Op &op);      // a non-batch op
    hipError_t mark(int slot, bool end);
    // reallocation of a buffer that kernels (and captured graphs) point at
    template <class T>
    hipError_t regrow(T *&p, uint64_t count);

    bool kernel_timing_ = false;
    bool balanced_ = true;
    bool autotune_ = true, graphs_ = true;
    // the untuned configuration's grid divisor and deepest split
    // (KAZAN_GRID_DIVISOR, KAZAN_SPLIT_LEVEL)
    uint32_t default_grid_div_ = 1, default_max_level_ = 2;
    uint64_t generation_ = 1;  // bumped whenever captured graphs go stale
    uint32_t graph_launches_ = 0, graph_captures_ = 0;
    BatchConfig last_tuned_{0, 0, false, false, 0, 0};
    uint64_t list_limit_ = 0;
    // balanced-mode scratch: up to 16 work items per tile; chunk tables
    // sized from the list capacity; sub-tile lists grown on demand
    uint32_t *tile_level_ = nullptr, *split_slot_ = nullptr, *split_info_ = nullptr, *queue_ = nullptr,
             *ctrl_ = nullptr, *split_counts_ = nullptr, *split_offsets_ = nullptr, *chunk_desc_ = nullptr,
             *chunk_counts_ = nullptr, *split_lists_ = nullptr, *plan_table_ = nullptr;
    uint16_t *split_mask_ = nullptr;  // [max_chunks * KZ_SPLIT_CHUNK]
    uint32_t cap_split_tiles_ = 0, max_chunks_ = 0, split_capacity_ = 0;
    uint32_t split_grid_ = 1024;  // workgroups of the chunk kernels: 4 per CU
    // cost model and tile profile
    KzCostModel *cost_model_d_ = nullptr;
    float frag_weight_ = 0.0f;
    std::string cost_model_src_;
    bool tile_profile_ = false;
    uint32_t *tile_acc_ = nullptr, *tile_cycles_ = nullptr, *tile_frags_ = nullptr, *tri_z_ = nullptr;
    float *tile_feat_ = nullptr, *tile_pred_ = nullptr;
    uint32_t cap_model_tiles_ = 0, cap_model_tris_ = 0, last_profile_tiles_ = 0;
    bool last_profile_pred_ = false, last_profile_labels_ = false;
    hipError_t ensure_model_scratch(uint32_t tiles, uint32_t tris);
    hipError_t ensure_split_scratch(uint32_t tiles);
    std::vector<std::pair<int, hipEvent_t>> timing_events_;  // (slot, start) then end, alternating
    std::vector<hipEvent_t> event_pool_;
    size_t events_used_ = 0;

    int device_ = -1;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_start_ = nullptr, ev_end_ = nullptr;
    uint32_t *tile_counts_ = nullptr, *tile_offsets_ = nullptr, *tile_cursor_ = nullptr;
    uint32_t *tile_lists_ = nullptr, *sort_tmp_ = nullptr;
    uint64_t *tri_rect_ = nullptr;
    uint32_t *tri_draw_ = nullptr;  // draw of each triangle, for batches of several draws
    // Binning: radix (pairs in primitive order, stable radix sort by tile)
    // or, with KAZAN_BINNING=sort, per-tile counts, an unordered scatter and
    // a sort of every list.
    bool radix_ = true;
    uint32_t *block_pairs_ = nullptr, *chunk_offsets_ = nullptr, *bin_ctrl_ = nullptr, *tile_flag_ = nullptr;
    uint32_t *pair_keys_[2] = {nullptr, nullptr}, *pair_vals_[2] = {nullptr, nullptr};
    // (balanced form) the list entries' 4 x 4 sub-tile masks
    uint16_t *list_mask_ = nullptr;
    uint32_t *radix_table_ = nullptr, *radix_totals_ = nullptr;
    // the buffers sized by the list capacity
    hipError_t alloc_lists(uint64_t capacity);
    // setup records written by binning for the raster (KAZAN_SETUP_RECORDS=1);
    // off by default: a record (96-112 B) costs as many HBM bytes as the
    // vertices it replaces, and writing it slows binning more than the
    // skipped setup speeds up the raster
    bool setup_records_ = false;
    uint32_t *tri_rec_ = nullptr;
    uint64_t cap_rec_words_ = 0;
    unsigned long long *stats_d_ = nullptr, *stats_h_ = nullptr;
    uint32_t cap_tiles_ = 0, cap_tris_ = 0, list_capacity_ = 0;
    uint32_t pending_batches_ = 0, pending_launches_ = 0;
    bool submit_open_ = false;
    SubmitStats last_;
};

}  // namespace kz
