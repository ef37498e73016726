// engine.cpp — replays recorded command lists on the device's HIP stream.
#include "engine.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

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
    hipFree(tile_level_);
    hipFree(tile_item_base_);
    hipFree(items_);
    hipFree(item_counts_);
    hipFree(item_offsets_);
    hipFree(item_cursor_);
    hipFree(queue_);
    hipFree(ctrl_);
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
    return true;
}

hipError_t Engine::ensure_item_scratch(uint32_t tiles) {
    if (tiles <= cap_items_tiles_) return hipSuccess;
    KZ_TRY(hipStreamSynchronize(stream_));
    const uint32_t t = std::max(tiles, cap_items_tiles_ * 2);
    const uint64_t items = 16ull * t;
    KZ_TRY(grow(tile_level_, t));
    KZ_TRY(grow(tile_item_base_, t));
    KZ_TRY(grow(items_, items));
    KZ_TRY(grow(item_counts_, items));
    KZ_TRY(grow(item_offsets_, items));
    KZ_TRY(grow(item_cursor_, items));
    KZ_TRY(grow(queue_, items));
    if (!ctrl_) KZ_TRY(hipMalloc((void **)&ctrl_, 128 * sizeof(uint32_t)));
    cap_items_tiles_ = t;
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
        KZ_TRY(grow(tile_counts_, t));
        KZ_TRY(grow(tile_offsets_, t));
        KZ_TRY(grow(tile_cursor_, t));
        cap_tiles_ = t;
    }
    if (tris > cap_tris_) {
        const uint32_t t = std::max(tris, cap_tris_ + cap_tris_ / 2);
        KZ_TRY(grow(tri_rect_, t));
        cap_tris_ = t;
    }
    if (want_list > list_capacity_) {
        const uint64_t c = std::min<uint64_t>(std::max<uint64_t>(want_list, (uint64_t)list_capacity_ * 3 / 2),
                                              0xFFFFFFF0ull);
        KZ_TRY(grow(tile_lists_, c));
        KZ_TRY(grow(sort_tmp_, c));
        list_capacity_ = (uint32_t)c;
    }
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

hipError_t Engine::run_batch(const Op &op) {
    KzFrame f = op.frame;
    if (f.total_tris == 0) return hipSuccess;
    const uint32_t ntiles = f.tiles_x * f.tiles_y;
    KZ_TRY(ensure_scratch(ntiles, f.total_tris));
    f.draws = op.d_draws;
    f.tile_counts = tile_counts_;
    f.tile_offsets = tile_offsets_;
    f.tile_cursor = tile_cursor_;
    f.tile_lists = tile_lists_;
    f.sort_tmp = sort_tmp_;
    f.list_capacity = list_capacity_;
    f.tri_rect = tri_rect_;
    f.stats = stats_d_;
    f.tile_shift = op.tile_shift;
    void *args[] = {&f};
    const uint32_t tri_blocks = (f.total_tris + 255) / 256;

    if (!balanced_) {
        KZ_TRY(mark(KS_CLEAR_COUNTS, false));
        KZ_TRY(kz_launch_fill_u32(stream_, tile_counts_, ntiles, 0u));
        KZ_TRY(mark(KS_CLEAR_COUNTS, true));
        KZ_TRY(mark(KS_BIN_COUNT, false));
        KZ_TRY(hipModuleLaunchKernel(op.kernels->bin_count, tri_blocks, 1, 1, 256, 1, 1, 0, stream_, args, nullptr));
        KZ_TRY(mark(KS_BIN_COUNT, true));
        KZ_TRY(mark(KS_BIN_SCAN, false));
        KZ_TRY(kz_launch_scan(stream_, tile_counts_, tile_offsets_, tile_cursor_, ntiles, nullptr, list_capacity_,
                              stats_d_));
        KZ_TRY(mark(KS_BIN_SCAN, true));
        KZ_TRY(mark(KS_BIN_SCATTER, false));
        KZ_TRY(kz_launch_bin_scatter(stream_, f));
        KZ_TRY(mark(KS_BIN_SCATTER, true));
        KZ_TRY(mark(KS_BIN_SORT, false));
        KZ_TRY(kz_launch_sort_lists(stream_, tile_counts_, tile_offsets_, ntiles, nullptr, tile_lists_, sort_tmp_,
                                    list_capacity_));
        KZ_TRY(mark(KS_BIN_SORT, true));
        KZ_TRY(mark(KS_RASTER, false));
        KZ_TRY(hipModuleLaunchKernel(op.kernels->raster, ntiles, 1, 1, 64, 1, 1, 0, stream_, args, nullptr));
        KZ_TRY(mark(KS_RASTER, true));
        pending_batches_ += 1;
        pending_launches_ += 6;
        return hipSuccess;
    }

    // balanced: tiles -> work items by predicted cost, per-item lists, queue
    KZ_TRY(ensure_item_scratch(ntiles));
    const uint32_t max_items = 16 * ntiles;
    f.balanced = 1;
    f.tile_level = tile_level_;
    f.tile_item_base = tile_item_base_;
    f.items = items_;
    f.item_counts = item_counts_;
    f.item_offsets = item_offsets_;
    f.item_cursor = item_cursor_;
    f.queue = queue_;
    f.queue_head = ctrl_;
    KZ_TRY(mark(KS_CLEAR_COUNTS, false));
    KZ_TRY(kz_launch_fill_u32(stream_, tile_counts_, ntiles, 0u));
    KZ_TRY(kz_launch_fill_u32(stream_, ctrl_, 128, 0u));
    KZ_TRY(mark(KS_CLEAR_COUNTS, true));
    KZ_TRY(mark(KS_BIN_COUNT, false));
    KZ_TRY(hipModuleLaunchKernel(op.kernels->bin_count, tri_blocks, 1, 1, 256, 1, 1, 0, stream_, args, nullptr));
    KZ_TRY(mark(KS_BIN_COUNT, true));
    KZ_TRY(mark(KS_BIN_SCAN, false));
    KZ_TRY(kz_launch_plan(stream_, f, ntiles, ctrl_));
    KZ_TRY(kz_launch_plan_emit(stream_, f, ntiles));
    KZ_TRY(kz_launch_item_count(stream_, f));
    KZ_TRY(kz_launch_scan(stream_, item_counts_, item_offsets_, item_cursor_, max_items, ctrl_ + 1, list_capacity_,
                          stats_d_));
    KZ_TRY(mark(KS_BIN_SCAN, true));
    KZ_TRY(mark(KS_BIN_SCATTER, false));
    KZ_TRY(kz_launch_item_scatter(stream_, f));
    KZ_TRY(mark(KS_BIN_SCATTER, true));
    KZ_TRY(mark(KS_BIN_SORT, false));
    KZ_TRY(kz_launch_sort_lists(stream_, item_counts_, item_offsets_, max_items, ctrl_ + 1, tile_lists_, sort_tmp_,
                                list_capacity_));
    KZ_TRY(kz_launch_queue(stream_, f, ctrl_, max_items));
    KZ_TRY(mark(KS_BIN_SORT, true));
    KZ_TRY(mark(KS_RASTER, false));
    KZ_TRY(hipModuleLaunchKernel(op.kernels->raster_persistent, std::min(op.kernels->persistent_blocks, max_items), 1,
                                 1, 64, 1, 1, 0, stream_, args, nullptr));
    KZ_TRY(mark(KS_RASTER, true));
    pending_batches_ += 1;
    pending_launches_ += 14;
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

hipError_t Engine::execute(const std::vector<Op> &ops) {
    for (const Op &op : ops) {
        switch (op.kind) {
        case Op::Fill:
            KZ_TRY(kz_launch_fill_rect_u32(stream_, op.fill_ptr, op.fill_width, op.x0, op.y0, op.w, op.h, op.fill_spp,
                                           op.fill_value));
            ++pending_launches_;
            break;
        case Op::Batch: KZ_TRY(run_batch(op)); break;
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
        }
    }
    return hipSuccess;
}

hipError_t Engine::end_submit(hipEvent_t fence_event) {
    KZ_TRY(hipEventRecord(ev_end_, stream_));
    KZ_TRY(kz_launch_stats_reduce(stream_, stats_d_));
    KZ_TRY(hipMemcpyAsync(stats_h_, stats_d_ + (size_t)KZ_STAT_STRIPES * KZ_STAT_COUNT,
                          KZ_STAT_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          stream_));
    if (fence_event) KZ_TRY(hipEventRecord(fence_event, stream_));
    return hipSuccess;
}

void Engine::collect_stats() {
    if (!submit_open_) return;
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
            if (grow(tile_lists_, need) == hipSuccess && grow(sort_tmp_, need) == hipSuccess)
                list_capacity_ = (uint32_t)need;
            else
                list_capacity_ = 0, cap_tris_ = 0;  // next ensure_scratch reallocates
        }
    }
}

}  // namespace kz
