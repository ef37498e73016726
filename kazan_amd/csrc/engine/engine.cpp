// engine.cpp — replays recorded command lists on the device's HIP stream.
#include "engine.h"

#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>

#include "kernels.h"

namespace kz {

namespace {

template <class T>
hipError_t grow(T *&p, uint64_t count) {
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        if (e != hipSuccess) return e;
    }
    return hipMalloc((void **)&p, std::max<uint64_t>(count, 1) * sizeof(T));
}

// striped partial counters followed by their totals
constexpr uint64_t kStatWords = (uint64_t)(KZ_STAT_STRIPES + 1) * KZ_STAT_COUNT;

#define KZ_TRY(x)                          \
    do {                                   \
        hipError_t e_ = (x);               \
        if (e_ != hipSuccess) return e_;   \
    } while (0)

}  // namespace

Engine::~Engine() {
    if (device_ < 0) return;
    hipSetDevice(device_);
    if (stream_) hipStreamSynchronize(stream_);
    hipFree(tile_counts_);
    hipFree(tile_offsets_);
    hipFree(tile_cursor_);
    hipFree(tile_lists_);
    hipFree(sort_tmp_);
    hipFree(tri_rect_);
    hipFree(tri_rec_);
    for (uint32_t *p : {tri_draw_, block_pairs_, chunk_offsets_, bin_ctrl_, tile_flag_, pair_keys_[0], pair_keys_[1],
                        pair_vals_[0], pair_vals_[1], radix_table_, radix_totals_})
        hipFree(p);
    for (uint32_t *p : {tile_level_, split_slot_, split_info_, queue_, ctrl_, split_counts_, split_offsets_,
                        chunk_desc_, chunk_counts_, split_lists_, plan_table_, tile_acc_, tile_cycles_, tile_frags_, tri_z_})
        hipFree(p);
    hipFree(split_mask_);
    hipFree(list_mask_);
    hipFree(tile_feat_);
    hipFree(tile_pred_);
    hipFree(cost_model_d_);
    hipFree(stats_d_);
    if (stats_h_) hipHostFree(stats_h_);
    for (hipEvent_t e : event_pool_) hipEventDestroy(e);
    if (ev_start_) hipEventDestroy(ev_start_);
    if (ev_end_) hipEventDestroy(ev_end_);
    if (stream_) hipStreamDestroy(stream_);
}

bool Engine::init(int hip_device, std::string &err) {
    device_ = hip_device;
    hipError_t e;
    if ((e = hipSetDevice(device_)) != hipSuccess || (e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking)) ||
        (e = hipEventCreate(&ev_start_)) || (e = hipEventCreate(&ev_end_)) ||
        (e = hipMalloc((void **)&stats_d_, kStatWords * sizeof(unsigned long long))) ||
        (e = hipHostMalloc((void **)&stats_h_, KZ_STAT_COUNT * sizeof(unsigned long long), hipHostMallocDefault))) {
        err = hipGetErrorString(e);
        return false;
    }
    std::memset(stats_h_, 0, KZ_STAT_COUNT * sizeof(unsigned long long));
    // KAZAN_LIST_CAPACITY_LIMIT caps the per-tile list storage (entries);
    // tiles whose lists do not fit render through the in-order scan path
    if (const char *e = std::getenv("KAZAN_LIST_CAPACITY_LIMIT")) list_limit_ = std::strtoull(e, nullptr, 10);
    if (const char *e = std::getenv("KAZAN_PIPELINE")) balanced_ = std::strcmp(e, "plain") != 0;
    if (const char *e = std::getenv("KAZAN_AUTOTUNE")) autotune_ = std::strcmp(e, "0") != 0;
    if (const char *e = std::getenv("KAZAN_GRAPHS")) graphs_ = std::strcmp(e, "0") != 0;
    if (const char *e = std::getenv("KAZAN_SETUP_RECORDS")) setup_records_ = std::strcmp(e, "0") != 0;
    if (const char *e = std::getenv("KAZAN_BINNING")) radix_ = std::strcmp(e, "sort") != 0;
    if (const char *e = std::getenv("KAZAN_GRID_DIVISOR")) {
        const int d = std::atoi(e);
        default_grid_div_ = d == 2 || d == 4 ? (uint32_t)d : 1u;
    }
    if (const char *e = std::getenv("KAZAN_SPLIT_LEVEL")) default_max_level_ = (uint32_t)std::min(std::max(std::atoi(e), 0), 2);
    if (radix_ &&
        ((e = hipMalloc((void **)&bin_ctrl_, KZ_BIN_WORDS * sizeof(uint32_t))) != hipSuccess ||
         (e = hipMalloc((void **)&radix_totals_, kz_radix_digits_max() * sizeof(uint32_t))) != hipSuccess)) {
        err = hipGetErrorString(e);
        return false;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_) == hipSuccess && cus > 0)
        split_grid_ = 4u * (uint32_t)cus;
    // the cost model: $KAZAN_COST_MODEL, else cost_model.bin beside this library
    std::string path;
    if (const char *e = std::getenv("KAZAN_COST_MODEL")) {
        path = e;
    } else {
        Dl_info info;
        if (dladdr((void *)&kz_launch_fill_u32, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/cost_model.bin";
        }
    }
    if (!path.empty() && path != "none") {
        std::ifstream in(path, std::ios::binary);
        if (in) {
            std::vector<char> blob((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
            if (set_cost_model(blob.data(), blob.size())) cost_model_src_ = path;
        }
    }
    return true;
}

bool Engine::set_cost_model(const void *blob, size_t bytes) {
    KzCostModel m;
    if (!blob || bytes != sizeof m) return false;
    std::memcpy(&m, blob, sizeof m);
    if (m.magic != KZ_COST_MODEL_MAGIC || m.version != 1 || m.bytes != sizeof m) return false;
    if (!cost_model_d_ && hipMalloc((void **)&cost_model_d_, sizeof m) != hipSuccess) {
        cost_model_d_ = nullptr;
        return false;
    }
    if (hipStreamSynchronize(stream_) != hipSuccess ||
        hipMemcpy(cost_model_d_, &m, sizeof m, hipMemcpyHostToDevice) != hipSuccess)
        return false;
    frag_weight_ = m.frag_weight;
    ++generation_;  // captured graphs carry the old model's frag_weight
    cost_model_src_ = "memory";
    return true;
}

hipError_t Engine::ensure_model_scratch(uint32_t tiles, uint32_t tris) {
    if (tiles <= cap_model_tiles_ && tris <= cap_model_tris_) return hipSuccess;
    KZ_TRY(hipStreamSynchronize(stream_));
    if (tiles > cap_model_tiles_) {
        const uint32_t t = std::max(tiles, cap_model_tiles_ * 2);
        KZ_TRY(regrow(tile_acc_, (uint64_t)KZ_ACC * t));
        KZ_TRY(regrow(tile_feat_, (uint64_t)KZ_FEATURES * t));
        KZ_TRY(regrow(tile_pred_, t));
        KZ_TRY(regrow(tile_frags_, t));
        KZ_TRY(regrow(tile_cycles_, t));
        cap_model_tiles_ = t;
    }
    if (tris > cap_model_tris_) {
        const uint32_t t = std::max(tris, cap_model_tris_ + cap_model_tris_ / 2);
        KZ_TRY(regrow(tri_z_, t));
        cap_model_tris_ = t;
    }
    return hipSuccess;
}

namespace {
hipError_t copy_u32_as_float(float *dst, const uint32_t *src, uint32_t n) {
    std::vector<uint32_t> c(n);
    KZ_TRY(hipMemcpy(c.data(), src, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) dst[i] = (float)c[i];
    return hipSuccess;
}
}  // namespace

hipError_t Engine::read_tile_profile(uint32_t *tiles, float *features, float *frags, float *cycles, float *pred) {
    KZ_TRY(hipSetDevice(device_));
    KZ_TRY(hipStreamSynchronize(stream_));
    const uint32_t n = tile_feat_ ? last_profile_tiles_ : 0;
    if (tiles) *tiles = n;
    if (n == 0) return hipSuccess;
    if (features)
        KZ_TRY(hipMemcpy(features, tile_feat_, (size_t)n * KZ_FEATURES * sizeof(float), hipMemcpyDeviceToHost));
    if (frags) {
        if (last_profile_labels_) KZ_TRY(copy_u32_as_float(frags, tile_frags_, n));
        else std::fill(frags, frags + n, 0.0f);
    }
    if (cycles) {
        if (last_profile_labels_) KZ_TRY(copy_u32_as_float(cycles, tile_cycles_, n));
        else std::fill(cycles, cycles + n, 0.0f);
    }
    if (pred) {
        if (last_profile_pred_) KZ_TRY(hipMemcpy(pred, tile_pred_, n * sizeof(float), hipMemcpyDeviceToHost));
        else std::fill(pred, pred + n, 0.0f);
    }
    return hipSuccess;
}

// Called after ensure_scratch (the chunk tables follow the list capacity).
hipError_t Engine::ensure_split_scratch(uint32_t tiles) {
    // a split tile's list fits the list storage, so its chunks number at
    // most list_capacity / KZ_SPLIT_CHUNK plus one partial chunk per tile
    const uint64_t want_chunks = (uint64_t)list_capacity_ / KZ_SPLIT_CHUNK + tiles;
    const uint32_t want_split = std::max(split_capacity_, list_capacity_);
    if (tiles <= cap_split_tiles_ && want_chunks <= max_chunks_ && want_split <= split_capacity_) return hipSuccess;
    KZ_TRY(hipStreamSynchronize(stream_));
    if (tiles > cap_split_tiles_) {
        const uint32_t t = std::max(tiles, cap_split_tiles_ * 2);
        KZ_TRY(regrow(tile_level_, t));
        KZ_TRY(regrow(split_slot_, t));
        KZ_TRY(regrow(split_info_, 2ull * t));
        KZ_TRY(regrow(queue_, 16ull * t));
        KZ_TRY(regrow(split_counts_, 16ull * t));
        KZ_TRY(regrow(split_offsets_, 16ull * t));
        KZ_TRY(regrow(plan_table_, kz_plan_table_words(t)));
        if (!ctrl_) {
            KZ_TRY(hipMalloc((void **)&ctrl_, KZ_CTRL_WORDS * sizeof(uint32_t)));
            KZ_TRY(kz_launch_fill_u32(stream_, ctrl_, KZ_CTRL_WORDS, 0u));
        }
        cap_split_tiles_ = t;
    }
    if (want_chunks > max_chunks_) {
        const uint64_t c = std::min<uint64_t>(std::max<uint64_t>(want_chunks, (uint64_t)max_chunks_ * 3 / 2),
                                              0x0FFFFFFFull);
        KZ_TRY(regrow(chunk_desc_, 4 * c));
        KZ_TRY(regrow(chunk_counts_, 16 * c));
        KZ_TRY(regrow(split_mask_, (uint64_t)KZ_SPLIT_CHUNK * c));
        max_chunks_ = (uint32_t)c;
    }
    if (want_split > split_capacity_) {
        KZ_TRY(regrow(split_lists_, want_split));
        split_capacity_ = want_split;
    }
    return hipSuccess;
}

hipError_t Engine::ensure_scratch(uint32_t tiles, uint32_t tris) {
    uint64_t want_list = std::max<uint64_t>(list_capacity_, 4ull * tris + 65536ull);
    if (list_limit_) want_list = std::min<uint64_t>(want_list, list_limit_);
    if (tiles <= cap_tiles_ && tris <= cap_tris_ && want_list <= list_capacity_) return hipSuccess;
    // buffers may still be in use by queued work
    KZ_TRY(hipStreamSynchronize(stream_));
    if (tiles > cap_tiles_) {
        const uint32_t t = std::max(tiles, cap_tiles_ * 2);
        KZ_TRY(regrow(tile_counts_, t));
        KZ_TRY(regrow(tile_offsets_, t));
        if (radix_) {
            KZ_TRY(regrow(tile_flag_, t));  // cleared with the counts before each batch
        } else {
            KZ_TRY(regrow(tile_cursor_, t));
        }
        cap_tiles_ = t;
    }
    if (tris > cap_tris_) {
        const uint32_t t = std::max(tris, cap_tris_ + cap_tris_ / 2);
        KZ_TRY(regrow(tri_rect_, t));
        KZ_TRY(regrow(tri_draw_, t));
        if (radix_) {
            KZ_TRY(regrow(block_pairs_, (t + KZ_BIN_BLOCK - 1) / KZ_BIN_BLOCK));
            KZ_TRY(regrow(chunk_offsets_, (t + KZ_PAIR_CHUNK - 1) / KZ_PAIR_CHUNK));
        }
        cap_tris_ = t;
    }
    if (want_list > list_capacity_) {
        const uint64_t c = std::min<uint64_t>(std::max<uint64_t>(want_list, (uint64_t)list_capacity_ * 3 / 2),
                                              0xFFFFFFF0ull);
        KZ_TRY(alloc_lists(c));
    }
    return hipSuccess;
}

hipError_t Engine::alloc_lists(uint64_t capacity) {
    list_capacity_ = 0;
    KZ_TRY(regrow(tile_lists_, capacity));
    if (radix_) {
        for (int i = 0; i < 2; ++i) {
            KZ_TRY(regrow(pair_keys_[i], capacity));
            KZ_TRY(regrow(pair_vals_[i], capacity));
        }
        KZ_TRY(regrow(radix_table_, (uint64_t)kz_radix_digits_max() * kz_radix_blocks(capacity)));
        if (balanced_) KZ_TRY(regrow(list_mask_, capacity));
    } else {
        KZ_TRY(regrow(sort_tmp_, capacity));
    }
    list_capacity_ = (uint32_t)capacity;
    return hipSuccess;
}

// Statistics accumulate over all submissions since the last collect_stats():
// a caller that queues K frames and then waits reads the K-frame totals and
// the GPU time from the first submission's start to the last one's end.
hipError_t Engine::begin_submit() {
    KZ_TRY(hipSetDevice(device_));
    if (!submit_open_) {
        KZ_TRY(kz_launch_fill_u32(stream_, (uint32_t *)stats_d_, kStatWords * 2, 0u));
        KZ_TRY(hipEventRecord(ev_start_, stream_));
        pending_batches_ = pending_launches_ = 0;
        submit_open_ = true;
    }
    return hipSuccess;
}

// All allocations a batch needs at tile size 1 << shift (so that a graph
// capture of the batch allocates nothing).
hipError_t Engine::prepare_batch(const Op &op, uint32_t shift) {
    const KzFrame &f = op.frame;
    if (f.total_tris == 0) return hipSuccess;
    const uint32_t ts = 1u << shift;
    const uint32_t ntiles = ((f.width + ts - 1) >> shift) * ((f.height + ts - 1) >> shift);
    KZ_TRY(ensure_scratch(ntiles, f.total_tris));
    if (setup_records_) {
        const uint64_t words = (uint64_t)f.total_tris * KZ_REC_WORDS_FOR(op.variants->fixed.num_varyings);
        if (words > cap_rec_words_) {
            KZ_TRY(hipStreamSynchronize(stream_));
            const uint64_t w = std::max(words, cap_rec_words_ + cap_rec_words_ / 2);
            KZ_TRY(regrow(tri_rec_, w));
            cap_rec_words_ = w;
        }
    }
    if (balanced_) KZ_TRY(ensure_split_scratch(ntiles));
    if ((cost_model_d_ != nullptr && balanced_) || tile_profile_) KZ_TRY(ensure_model_scratch(ntiles, f.total_tris));
    return hipSuccess;
}

hipError_t Engine::run_batch(const Op &op, const BatchConfig &cfg) {
    KzFrame f = op.frame;
    if (f.total_tris == 0) return hipSuccess;
    const uint32_t shift = cfg.shift, grid_div = cfg.grid_div;
    // the balanced form's planned, persistent raster (else one workgroup per tile)
    const bool persistent = balanced_ && cfg.persistent;
    std::string err;
    const PipelineKernels *kern = op.variants->get(shift, &err);
    if (!kern) {
        std::fprintf(stderr, "kazan: tile size %u: %s\n", 1u << shift, err.c_str());
        return hipErrorInvalidValue;
    }
    const uint32_t ts = 1u << shift;
    f.tiles_x = (f.width + ts - 1) >> shift;
    f.tiles_y = (f.height + ts - 1) >> shift;
    const uint32_t ntiles = f.tiles_x * f.tiles_y;
    KZ_TRY(prepare_batch(op, shift));
    f.draws = op.d_draws;
    f.draw_begin = op.d_draw_begin;
    f.block_draw = op.d_block_draw;
    f.tile_counts = tile_counts_;
    f.tile_offsets = tile_offsets_;
    f.tile_cursor = tile_cursor_;
    f.tile_flag = radix_ ? tile_flag_ : nullptr;
    f.bin_ctrl = radix_ ? bin_ctrl_ : nullptr;
    f.tile_lists = tile_lists_;
    f.sort_tmp = sort_tmp_;
    f.list_capacity = list_capacity_;
    f.tri_rect = tri_rect_;
    f.tri_draw = f.ndraws > 1 ? tri_draw_ : nullptr;
    f.tri_rec = setup_records_ ? tri_rec_ : nullptr;
    f.stats = stats_d_;
    f.tile_shift = shift;
    if (persistent) {
        f.balanced = 1;
        f.split_capacity = split_capacity_;
        f.queue = queue_;
        f.ctrl = ctrl_;
        f.tile_level = tile_level_;
        f.split_slot = split_slot_;
        f.split_info = split_info_;
        f.chunk_desc = chunk_desc_;
        f.chunk_counts = chunk_counts_;
        f.split_mask = split_mask_;
        f.split_counts = split_counts_;
        f.split_offsets = split_offsets_;
        f.split_lists = split_lists_;
        f.max_chunks = max_chunks_;
        f.raster_slots = std::max(1u, kern->persistent_blocks / std::max(grid_div, 1u));
        f.split_max_level = std::min(cfg.max_level, 2u);
        // the sub-tile masks travel in the pairs' keys (above 16 tile bits)
        // through the radix sort, so the split reads them beside the lists
        // instead of gathering pixel boxes
        if (radix_ && list_mask_ && f.split_max_level > 0 && shift >= 2 && ntiles < 65536u) f.list_mask = list_mask_;
    }
    // the cost model's inputs (and, in the static form, its labels) when a
    // model will run or tiles are being profiled
    // (predictions also run in a profiled static batch, to measure the
    // model's error against the fragments the rasteriser counts)
    const bool model = cost_model_d_ != nullptr && ((persistent && cfg.model) || tile_profile_);
    const bool acc = model || tile_profile_;
    const bool labels = tile_profile_ && !persistent;
    if (acc) {
        f.tri_z = tri_z_;
        f.tile_acc = tile_acc_;
        f.tile_feat = tile_feat_;
        if (model && persistent) {
            f.tile_pred = tile_pred_;
            f.frag_weight = frag_weight_ * (float)std::max(cfg.frag_scale, 1u);
        }
        if (labels) {
            f.tile_frags = tile_frags_;
            f.tile_cycles = tile_cycles_;
        }
    }
    void *args[] = {&f};
    const uint32_t tri_blocks = (f.total_tris + 255) / 256;
    uint32_t launches = 0;

    // binning: pixel boxes, then the per-tile lists in primitive order
    KZ_TRY(mark(KS_CLEAR_COUNTS, false));
    KZ_TRY(kz_launch_fill_tiles(stream_, tile_counts_, tile_offsets_, radix_ ? tile_flag_ : nullptr, ntiles));
    if (acc) KZ_TRY(kz_launch_fill_u32(stream_, tile_acc_, (uint64_t)KZ_ACC * ntiles, 0u));
    KZ_TRY(mark(KS_CLEAR_COUNTS, true));
    launches += acc ? 2 : 1;
    if (radix_) {
        // pairs in primitive order, a stable radix sort by tile whose last
        // pass writes the lists, counts and offsets
        f.block_pairs = block_pairs_;
        KzRadixScratch rs{chunk_offsets_, bin_ctrl_, tile_flag_, {pair_keys_[0], pair_keys_[1]},
                          {pair_vals_[0], pair_vals_[1]}, radix_table_, radix_totals_,
                          kz_radix_blocks(list_capacity_)};
        KZ_TRY(mark(KS_BIN_COUNT, false));
        KZ_TRY(hipModuleLaunchKernel(kern->bin_count, tri_blocks, 1, 1, KZ_BIN_BLOCK, 1, 1, 0, stream_, args,
                                     nullptr));
        KZ_TRY(mark(KS_BIN_COUNT, true));
        KZ_TRY(mark(KS_BIN_HIST, false));
        KZ_TRY(kz_launch_bin_pairs(stream_, f, rs));
        KZ_TRY(mark(KS_BIN_HIST, true));
        KZ_TRY(mark(KS_BIN_SCATTER, false));
        KZ_TRY(kz_launch_bin_radix(stream_, f, rs, ntiles));
        KZ_TRY(mark(KS_BIN_SCATTER, true));
        launches += 1 + 2 + 3 * kz_radix_passes(ntiles);
    } else {
        // per-tile counts, offsets, an unordered scatter, a sort of each list
        KZ_TRY(mark(KS_BIN_COUNT, false));
        KZ_TRY(hipModuleLaunchKernel(kern->bin_count, tri_blocks, 1, 1, KZ_BIN_BLOCK, 1, 1, 0, stream_, args,
                                     nullptr));
        KZ_TRY(mark(KS_BIN_COUNT, true));
        KZ_TRY(mark(KS_BIN_HIST, false));
        KZ_TRY(acc ? kz_launch_bin_hist_acc(stream_, f) : kz_launch_bin_hist(stream_, f));
        KZ_TRY(mark(KS_BIN_HIST, true));
        KZ_TRY(mark(KS_BIN_SCAN, false));
        KZ_TRY(kz_launch_scan(stream_, tile_counts_, tile_offsets_, tile_cursor_, ntiles, list_capacity_, stats_d_));
        KZ_TRY(mark(KS_BIN_SCAN, true));
        KZ_TRY(mark(KS_BIN_SCATTER, false));
        KZ_TRY(kz_launch_bin_scatter(stream_, f));
        KZ_TRY(mark(KS_BIN_SCATTER, true));
        KZ_TRY(mark(KS_BIN_SORT, false));
        KZ_TRY(kz_launch_sort_lists(stream_, f, ntiles));
        KZ_TRY(mark(KS_BIN_SORT, true));
        launches += 6;
    }

    if (!persistent) {
        if (acc) {
            KZ_TRY(kz_launch_tile_features(stream_, f, ntiles));
            launches += 1;
        }
        if (model) {
            KZ_TRY(kz_launch_cost_mlp(stream_, cost_model_d_, tile_feat_, ntiles, tile_pred_));
            launches += 1;
        }
        KZ_TRY(mark(KS_RASTER, false));
        KZ_TRY(hipModuleLaunchKernel(kern->raster, ntiles, 1, 1, 64, 1, 1, 0, stream_, args, nullptr));
        KZ_TRY(mark(KS_RASTER, true));
        launches += 1;
    } else {
        // predicted costs, split expensive tiles, cut their sub-tile lists,
        // order the items
        KZ_TRY(mark(KS_PLAN, false));
        if (model) {  // features, predictions and the plan's cost parts: one launch
            KZ_TRY(kz_launch_cost_plan(stream_, cost_model_d_, f, ntiles, plan_table_));
            launches += 1;
        } else if (acc) {
            KZ_TRY(kz_launch_tile_features(stream_, f, ntiles));
            launches += 1;
        }
        KZ_TRY(kz_launch_plan(stream_, f, ntiles, plan_table_, model));
        KZ_TRY(kz_launch_split(stream_, f, split_grid_, ntiles));
        KZ_TRY(mark(KS_PLAN, true));
        KZ_TRY(mark(KS_RASTER, false));
        const uint32_t grid = std::min(f.raster_slots, 16 * ntiles);
        KZ_TRY(hipModuleLaunchKernel(kern->raster_persistent, grid, 1, 1, 64, 1, 1, 0, stream_, args, nullptr));
        KZ_TRY(mark(KS_RASTER, true));
        // the plan (2, or 3 with its own cost pass), the split (4), the raster
        launches += (!model && !f.bin_ctrl) ? 8 : 7;
    }
    if (acc) {
        last_profile_tiles_ = ntiles;
        last_profile_pred_ = model;
        last_profile_labels_ = labels;
    }
    pending_batches_ += 1;
    pending_launches_ += launches;
    return hipSuccess;
}

hipError_t Engine::mark(int slot, bool end) {
    if (!kernel_timing_) return hipSuccess;
    if (events_used_ == event_pool_.size()) {
        hipEvent_t e;
        KZ_TRY(hipEventCreate(&e));
        event_pool_.push_back(e);
    }
    hipEvent_t e = event_pool_[events_used_++];
    KZ_TRY(hipEventRecord(e, stream_));
    timing_events_.push_back({end ? -1 - slot : slot, e});
    return hipSuccess;
}

BatchTuning::~BatchTuning() {
    for (hipEvent_t e : ev)
        if (e) hipEventDestroy(e);
}

void GraphCache::reset() {
    if (exec) hipGraphExecDestroy(exec);
    exec = nullptr;
    generation = 0;
}

template <class T>
hipError_t Engine::regrow(T *&p, uint64_t count) {
    ++generation_;
    return grow(p, count);
}

// Tuning candidates of a batch: every tile size whose kernels fit (full
// grid) first; then, at the fastest tile size, a half and a quarter grid.
BatchConfig Engine::default_config(const Op &op) const {
    BatchConfig c;
    c.shift = op.tile_shift;
    c.grid_div = default_grid_div_;
    c.model = true;
    c.persistent = balanced_;
    c.max_level = default_max_level_;
    c.frag_scale = 1;
    return c;
}

namespace {
int fastest(const BatchTuning &t) {
    int best = -1;
    for (size_t i = 0; i < t.cand.size(); ++i)
        if (t.ms[i] >= 0 && (best < 0 || t.ms[i] < t.ms[best])) best = (int)i;
    return best;
}

// The choice: the finalist with the lowest median of its runs (one lucky
// run does not decide), else the fastest run.
int final_choice(const BatchTuning &t) {
    int best = -1;
    float best_ms = 0.0f;
    for (const std::vector<size_t> &runs : t.finals) {
        std::vector<float> v;
        for (size_t i : runs)
            if (t.ms[i] >= 0) v.push_back(t.ms[i]);
        if (v.empty()) continue;
        std::sort(v.begin(), v.end());
        const float med = v[v.size() / 2];
        if (best < 0 || med < best_ms) {
            best = (int)runs[0];
            best_ms = med;
        }
    }
    return best >= 0 ? best : fastest(t);
}
}  // namespace

// Candidates, one HIP-event-timed trial per submission: every tile size
// whose kernels fit (full grid, the plan with the cost model when there is
// one); at the fastest size, a half and a quarter grid; at the fastest of
// those, the plan without the model and the plain raster (one workgroup per
// tile, no plan); at the fastest planned one, splits of at most 4 sub-tiles
// and none (many small triangles are set up once per sub-tile they touch);
// with a cost model, the model's plan with its fragment weight times 4 and
// 16 (a pipeline whose fragments cost more than the training pipelines');
// at the fastest planned one, every other pair of tile size and grid size
// (both were chosen with the deepest splits and the model; with fewer
// splits smaller tiles on a larger grid may pay); then the three fastest
// twice more each. The finalist with the lowest median of its three runs
// is kept.
hipError_t Engine::pick_config(Op &op, BatchConfig &cfg, bool &timed) {
    cfg = default_config(op);
    timed = false;
    if (!autotune()) return hipSuccess;
    const bool model = cost_model_d_ != nullptr;
    if (!op.tuning) {
        op.tuning = std::make_shared<BatchTuning>();
        BatchTuning &t = *op.tuning;
        for (uint32_t s = kMinTileShift; s <= kMaxTileShift; ++s)
            if (op.variants->fits(s)) t.cand.push_back({s, 1u, model, true, 2u});
        t.ms.assign(t.cand.size(), -1.0f);
        KZ_TRY(hipEventCreate(&t.ev[0]));
        KZ_TRY(hipEventCreate(&t.ev[1]));
    }
    BatchTuning &t = *op.tuning;
    if (t.done) {
        cfg = t.choice;
        return hipSuccess;
    }
    if (t.pending >= 0) {  // the previous submission's trial
        KZ_TRY(hipEventSynchronize(t.ev[1]));
        float ms = -1.0f;
        if (hipEventElapsedTime(&ms, t.ev[0], t.ev[1]) == hipSuccess) t.ms[t.pending] = ms;
        t.pending = -1;
    }
    // skip candidates whose kernels do not build
    while (t.next < t.cand.size() && !op.variants->get(t.cand[t.next].shift)) ++t.next;
    while (t.next == t.cand.size() && t.phase < 6) {
        const int best = fastest(t);
        if (best >= 0) {
            const BatchConfig b = t.cand[best];
            if (t.phase == 0) {
                t.cand.push_back({b.shift, 2u, b.model, true, 2u, 1u});
                t.cand.push_back({b.shift, 4u, b.model, true, 2u, 1u});
            } else if (t.phase == 1) {
                if (b.model) t.cand.push_back({b.shift, b.grid_div, false, true, 2u, 1u});
                t.cand.push_back({b.shift, 1u, false, false, 2u, 1u});
            } else if (t.phase == 2) {
                if (b.persistent) {
                    t.cand.push_back({b.shift, b.grid_div, b.model, true, 1u, 1u});
                    t.cand.push_back({b.shift, b.grid_div, b.model, true, 0u, 1u});
                }
            } else if (t.phase == 3) {
                if (b.persistent && model) {
                    // (splitting by triangle counts may not have paid; by
                    // weighted fragments it may)
                    t.cand.push_back({b.shift, b.grid_div, true, true, 2u, 4u});
                    t.cand.push_back({b.shift, b.grid_div, true, true, 2u, 16u});
                }
            } else if (t.phase == 4) {
                if (b.persistent)
                    for (uint32_t s = kMinTileShift; s <= kMaxTileShift; ++s)
                        for (uint32_t d : {1u, 2u, 4u})
                            if ((s != b.shift || d != b.grid_div) && op.variants->fits(s))
                                t.cand.push_back({s, d, b.model, true, b.max_level, b.frag_scale});
            } else {
                // the three fastest twice more each (a near tie is decided
                // by the median of three runs, not by one run's noise)
                std::vector<size_t> order;
                for (size_t i = 0; i < t.cand.size(); ++i)
                    if (t.ms[i] >= 0) order.push_back(i);
                std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return t.ms[x] < t.ms[y]; });
                for (size_t k = 0; k < order.size() && k < 3; ++k) {
                    t.finals.push_back({order[k], t.cand.size(), t.cand.size() + 1});
                    t.cand.push_back(t.cand[order[k]]);
                    t.cand.push_back(t.cand[order[k]]);
                }
            }
            t.ms.resize(t.cand.size(), -1.0f);
        }
        ++t.phase;
    }
    if (t.next < t.cand.size()) {
        cfg = t.cand[t.next];
        t.pending = (int)t.next++;
        timed = true;
        return hipSuccess;
    }
    const int best = final_choice(t);
    t.done = true;
    if (best >= 0) t.choice = t.cand[best];
    else t.choice = default_config(op);
    cfg = t.choice;
    last_tuned_ = cfg;
    return hipSuccess;
}

hipError_t Engine::execute(std::vector<Op> &ops, GraphCache *cache) {
    // Graphs pay off for command buffers that draw and are submitted again:
    // the first submission (and every tuning trial) runs eagerly.
    bool graphable = cache != nullptr && graphs_ && !kernel_timing_;
    bool draws = false;
    for (Op &op : ops)
        if (op.kind == Op::Batch) {
            draws = true;
            if (autotune() && !(op.tuning && op.tuning->done)) graphable = false;
        }
    if (!graphable || !draws || (cache->eager_runs == 0 && !cache->exec)) {
        if (cache) ++cache->eager_runs;
        return execute_eager(ops);
    }
    // every allocation before the capture
    for (Op &op : ops)
        if (op.kind == Op::Batch) {
            BatchConfig cfg;
            bool timed;
            KZ_TRY(pick_config(op, cfg, timed));
            if (!op.variants->get(cfg.shift)) return execute_eager(ops);
            KZ_TRY(prepare_batch(op, cfg.shift));
        }
    if (!cache->exec || cache->generation != generation_) {
        cache->reset();
        const uint32_t b0 = pending_batches_, l0 = pending_launches_;
        KZ_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
        const uint64_t gen = generation_;
        const hipError_t e = launch_ops(ops);
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(stream_, &graph);
        cache->batches = pending_batches_ - b0;
        cache->launches = pending_launches_ - l0;
        pending_batches_ = b0;
        pending_launches_ = l0;
        if (e != hipSuccess || e2 != hipSuccess || gen != generation_) {
            if (graph) hipGraphDestroy(graph);
            if (e != hipSuccess) return e;
            if (e2 != hipSuccess) return e2;
            return execute_eager(ops);  // something was reallocated mid-capture
        }
        const hipError_t e3 = hipGraphInstantiate(&cache->exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e3 != hipSuccess) {
            cache->exec = nullptr;
            return e3;
        }
        cache->generation = generation_;
        ++graph_captures_;
    }
    KZ_TRY(hipGraphLaunch(cache->exec, stream_));
    pending_batches_ += cache->batches;
    pending_launches_ += cache->launches;
    ++graph_launches_;
    return hipSuccess;
}

hipError_t Engine::execute_eager(std::vector<Op> &ops) {
    for (Op &op : ops)
        if (op.kind == Op::Batch) {
            BatchConfig cfg;
            bool timed;
            KZ_TRY(pick_config(op, cfg, timed));
            if (timed) {
                // a trial times the batch, not the loading of a tile size's
                // kernels or the growth of its scratch (host work and syncs)
                if (op.variants->get(cfg.shift)) KZ_TRY(prepare_batch(op, cfg.shift));
                KZ_TRY(hipEventRecord(op.tuning->ev[0], stream_));
            }
            KZ_TRY(run_batch(op, cfg));
            if (timed) KZ_TRY(hipEventRecord(op.tuning->ev[1], stream_));
        } else {
            KZ_TRY(launch_ops_one(op));
        }
    return hipSuccess;
}

hipError_t Engine::launch_ops(std::vector<Op> &ops) {
    for (Op &op : ops) {
        if (op.kind == Op::Batch) {
            const BatchTuning *t = op.tuning.get();
            KZ_TRY(run_batch(op, t && t->done ? t->choice : default_config(op)));
        } else {
            KZ_TRY(launch_ops_one(op));
        }
    }
    return hipSuccess;
}

hipError_t Engine::launch_ops_one(const Op &op) {
    {
        switch (op.kind) {
        case Op::Fill:
            KZ_TRY(kz_launch_fill_rect_u32(stream_, op.fill_ptr, op.fill_width, op.x0, op.y0, op.w, op.h, op.fill_spp,
                                           op.fill_value));
            ++pending_launches_;
            break;
        case Op::Batch: break;
        case Op::Copy:
            if (op.bytes) KZ_TRY(hipMemcpyAsync(op.dst, op.src, op.bytes, hipMemcpyDefault, stream_));
            break;
        case Op::Copy2D:
            if (op.rows && op.row_bytes)
                KZ_TRY(hipMemcpy2DAsync(op.dst, op.dst_pitch, op.src, op.src_pitch, op.row_bytes, op.rows,
                                        hipMemcpyDefault, stream_));
            break;
        case Op::Resolve:
            KZ_TRY(kz_launch_resolve4(stream_, op.rsrc, op.rdst, op.rwidth, op.x0, op.y0, op.w, op.h));
            ++pending_launches_;
            break;
        case Op::Dispatch: {
            size_t size = sizeof op.cframe;
            void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<KzComputeFrame *>(&op.cframe),
                              HIP_LAUNCH_PARAM_BUFFER_SIZE, &size, HIP_LAUNCH_PARAM_END};
            KZ_TRY(hipModuleLaunchKernel(op.cs_kernel, op.groups[0], op.groups[1], op.groups[2], op.block[0],
                                         op.block[1], op.block[2], 0, stream_, nullptr, config));
            ++pending_launches_;
            break;
        }
        }
    }
    return hipSuccess;
}

hipError_t Engine::end_submit(hipEvent_t fence_event) {
    KZ_TRY(hipEventRecord(ev_end_, stream_));
    if (fence_event) KZ_TRY(hipEventRecord(fence_event, stream_));
    return hipSuccess;
}

void Engine::collect_stats() {
    if (!submit_open_) return;
    // the striped counters are folded and read once per wait, not once per
    // submission (they accumulate over the submissions in between)
    if (kz_launch_stats_reduce(stream_, stats_d_) != hipSuccess ||
        hipMemcpyAsync(stats_h_, stats_d_ + (size_t)KZ_STAT_STRIPES * KZ_STAT_COUNT,
                       KZ_STAT_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_) != hipSuccess)
        std::memset(stats_h_, 0, KZ_STAT_COUNT * sizeof(unsigned long long));
    hipEventSynchronize(ev_end_);
    hipStreamSynchronize(stream_);
    std::memcpy(last_.counters, stats_h_, sizeof(last_.counters));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev_start_, ev_end_) == hipSuccess) last_.gpu_ms = ms;
    last_.batches = pending_batches_;
    last_.launches = pending_launches_;
    for (int k = 0; k < KS_COUNT; ++k) {
        last_.kernel_ms[k] = 0.0f;
        last_.kernel_launches[k] = 0;
    }
    for (size_t i = 0; i + 1 < timing_events_.size(); i += 2) {
        const int slot = timing_events_[i].first;
        float kms = 0.0f;
        if (hipEventElapsedTime(&kms, timing_events_[i].second, timing_events_[i + 1].second) == hipSuccess) {
            last_.kernel_ms[slot] += kms;
            last_.kernel_launches[slot] += 1;
        }
    }
    timing_events_.clear();
    events_used_ = 0;
    submit_open_ = false;
    // A submission whose lists overflowed rendered correctly through the
    // in-order scan fallback; size the lists for the next one.
    if (last_.counters[KZ_STAT_LIST_OVERFLOW]) {
        uint64_t need = last_.counters[KZ_STAT_LIST_ENTRIES] + last_.counters[KZ_STAT_LIST_ENTRIES] / 4;
        if (list_limit_) need = std::min<uint64_t>(need, list_limit_);
        if (need > list_capacity_ && need < 0xFFFFFFF0ull) {
            hipStreamSynchronize(stream_);
            if (alloc_lists(need) != hipSuccess) list_capacity_ = 0, cap_tris_ = 0;  // next ensure_scratch reallocates
        }
    }
    // Likewise for sub-tile lists (their overflow fallback filters the
    // parent tile's list): the device keeps the largest batch total.
    if (ctrl_ && last_.counters[KZ_STAT_SPLIT_ENTRIES] > split_capacity_) {
        uint32_t mx = 0;
        if (hipMemcpy(&mx, ctrl_ + KZ_CTRL_SPLIT_MAX, sizeof mx, hipMemcpyDeviceToHost) == hipSuccess &&
            mx > split_capacity_) {
            const uint64_t need = std::min<uint64_t>((uint64_t)mx + mx / 4, 0xFFFFFFF0ull);
            if (regrow(split_lists_, need) == hipSuccess)
                split_capacity_ = (uint32_t)need;
            else
                split_capacity_ = 0;  // next ensure_split_scratch reallocates
        }
    }
}

}  // namespace kz
