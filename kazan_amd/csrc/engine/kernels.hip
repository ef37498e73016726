// kernels.hip — the pipeline-independent kernels of the draw path (gfx950).
//
//   kz_fill_u32 / kz_fill_rect_u32  attachment clears and counter resets (a
//                    fill kernel rather than a memset, so command buffers can
//                    be captured into hipGraphs)
//   kz_scan          exclusive prefix sum of per-list counts
//   kz_bin_scatter   writes triangle ids into the per-tile lists
//   kz_sort_lists    sorts each list into primitive order (bitonic in LDS;
//                    lists longer than one LDS load are sorted in LDS-sized
//                    runs and merged with merge-path in HBM)
//   kz_plan_*        load-balanced mode: split each tile into 1, 4 or 16
//                    work items by predicted cost, bin triangles per item,
//                    order the items most expensive first
//   kz_resolve4      4x -> 1x box resolve, (s0+s1+s2+s3+2) >> 2 per channel
//   kz_stats_reduce  folds the striped statistics counters
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace {

constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortCap = 4096;
constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (uint32_t d = 32; d > 0; d >>= 1) {
        const uint32_t o = __shfl_xor(v, d, 64);
        v = v > o ? v : o;
    }
    return v;
}

// One atomic per distinct counter per wave.
__device__ __forceinline__ void wave_count(uint32_t *counts, uint32_t idx, uint32_t lane) {
    uint64_t pending = __ballot(idx != kNone);
    while (pending) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
        const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)idx, leader);
        const uint64_t peers = __ballot(idx == t) & pending;
        if (lane == leader) atomicAdd(&counts[t], (uint32_t)__popcll(peers));
        pending &= ~peers;
    }
}

// Reserves one slot per lane with an index, one atomic per distinct index
// per wave; returns this lane's slot (or kNone).
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *cursor, uint32_t idx, uint32_t lane) {
    uint64_t pending = __ballot(idx != kNone);
    uint32_t my_slot = kNone;
    while (pending) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
        const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)idx, leader);
        const uint64_t peers = __ballot(idx == t) & pending;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&cursor[t], (uint32_t)__popcll(peers));
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
        if ((peers >> lane) & 1ull) my_slot = base + rank_in(peers);
        pending &= ~peers;
    }
    return my_slot;
}

__device__ __forceinline__ void unpack_box(uint64_t r, uint32_t &x0, uint32_t &y0, uint32_t &x1, uint32_t &y1) {
    x0 = (uint32_t)(r & 0xFFFF);
    y0 = (uint32_t)((r >> 16) & 0xFFFF);
    x1 = (uint32_t)((r >> 32) & 0xFFFF);
    y1 = (uint32_t)(r >> 48);
}

__global__ void kz_fill_u32(uint32_t *dst, uint64_t n, uint32_t value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = value;
}

// Fills a rectangle of a W-wide attachment with `spp` 32-bit values per pixel.
__global__ void kz_fill_rect_u32(uint32_t *dst, uint32_t width, int32_t x0, int32_t y0, uint32_t w, uint32_t h,
                                 uint32_t spp, uint32_t value) {
    const uint64_t n = (uint64_t)w * h * spp;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t p = i / spp, s = i % spp;
        const uint64_t y = y0 + p / w, x = x0 + p % w;
        dst[(y * width + x) * spp + s] = value;
    }
}

// Exclusive prefix sum of counts[0, n) into offsets and cursor (n from
// n_dev when given). Adds the total to *total_out and, when `capacity` is
// exceeded, counts an overflow.
__global__ __launch_bounds__(kScanThreads) void kz_scan(const uint32_t *counts, uint32_t *offsets, uint32_t *cursor,
                                                        uint32_t n_host, const uint32_t *n_dev, uint32_t capacity,
                                                        unsigned long long *stats, uint32_t *total_out) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    __shared__ uint32_t s_carry;
    const uint32_t n = n_dev ? *n_dev : n_host;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < n; b += kScanThreads) {
        const uint32_t i = b + tid;
        const uint32_t v = i < n ? counts[i] : 0u;
        uint32_t x = v;
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(x, d, 64);
            if (lane >= d) x += o;
        }
        if (lane == 63) s_wave[wave] = x;
        __syncthreads();
        uint32_t wbase = 0, total = 0;
        for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
            const uint32_t t = s_wave[w];
            if (w < wave) wbase += t;
            total += t;
        }
        const uint32_t carry = s_carry;
        if (i < n) {
            const uint32_t excl = carry + wbase + x - v;
            offsets[i] = excl;
            if (cursor) cursor[i] = excl;
        }
        __syncthreads();
        if (tid == 0) s_carry = carry + total;
        __syncthreads();
    }
    if (tid == 0 && total_out) *total_out = s_carry;
    if (tid == 0 && stats) {
        atomicAdd(&stats[KZ_STAT_LIST_ENTRIES], (unsigned long long)s_carry);
        if (s_carry > capacity) atomicAdd(&stats[KZ_STAT_LIST_OVERFLOW], 1ull);
    }
}

// Static mode: one list per screen tile.
__global__ __launch_bounds__(256) void kz_bin_scatter(KzFrame f) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = lane_id();
    const uint64_t r = g < f.total_tris ? f.tri_rect[g] : KZ_RECT_EMPTY;
    uint32_t tx0 = 1, tx1 = 0, ty0 = 1, ty1 = 0;
    if (r != KZ_RECT_EMPTY) {
        uint32_t x0, y0, x1, y1;
        unpack_box(r, x0, y0, x1, y1);
        const uint32_t sh = f.tile_shift;
        tx0 = x0 >> sh;
        tx1 = x1 >> sh;
        ty0 = y0 >> sh;
        ty1 = y1 >> sh;
    }
    const uint32_t ntiles = tx0 <= tx1 ? (tx1 - tx0 + 1) * (ty1 - ty0 + 1) : 0u;
    const uint32_t steps = wave_max_u32(ntiles);
    uint32_t tx = tx0, ty = ty0;
    for (uint32_t k = 0; k < steps; ++k) {
        const bool act = k < ntiles;
        const uint32_t slot = wave_reserve(f.tile_cursor, act ? ty * f.tiles_x + tx : kNone, lane);
        if (act) {
            if (slot < f.list_capacity) f.tile_lists[slot] = g;
            if (++tx > tx1) {
                tx = tx0;
                ++ty;
            }
        }
    }
}

__device__ void bitonic_lds(uint32_t *s, uint32_t P) {
    for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += kSortThreads) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t a = s[i], b = s[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
}

__device__ uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

__device__ void sort_run_lds(uint32_t *g, uint32_t n, uint32_t *s) {
    const uint32_t P = next_pow2(n);
    for (uint32_t i = threadIdx.x; i < P; i += kSortThreads) s[i] = i < n ? g[i] : 0xFFFFFFFFu;
    __syncthreads();
    bitonic_lds(s, P);
    for (uint32_t i = threadIdx.x; i < n; i += kSortThreads) g[i] = s[i];
    __syncthreads();
}

// Merges adjacent sorted runs of width w from src into dst (merge path:
// each output element finds its split point by binary search; keys are
// unique triangle ids).
__device__ void merge_pass(const uint32_t *src, uint32_t *dst, uint32_t n, uint32_t w) {
    for (uint32_t o = threadIdx.x; o < n; o += kSortThreads) {
        const uint32_t pair = o / (2 * w);
        const uint32_t start = pair * 2 * w;
        const uint32_t a = min(w, n - start);
        const uint32_t b = min(w, n - start - a);
        const uint32_t *A = src + start, *B = A + a;
        const uint32_t k = o - start;
        uint32_t lo = k > b ? k - b : 0, hi = min(k, a);
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (A[mid] < B[k - 1 - mid]) lo = mid + 1;
            else hi = mid;
        }
        const uint32_t i = lo, j = k - lo;
        dst[o] = (i < a && (j >= b || A[i] < B[j])) ? A[i] : B[j];
    }
    __syncthreads();
}

// Sorts lists [0, n) (n from n_dev when given), grid-stride over lists.
__global__ __launch_bounds__(kSortThreads) void kz_sort_lists(const uint32_t *counts, const uint32_t *offsets,
                                                               uint32_t n_host, const uint32_t *n_dev,
                                                               uint32_t *lists, uint32_t *tmp, uint32_t capacity) {
    __shared__ uint32_t s[kSortCap];
    const uint32_t nl = n_dev ? *n_dev : n_host;
    for (uint32_t l = blockIdx.x; l < nl; l += gridDim.x) {
        const uint32_t n = counts[l];
        const uint32_t off = offsets[l];
        if (n <= 1 || (uint64_t)off + n > (uint64_t)capacity) continue;
        uint32_t *list = lists + off;
        if (n <= kSortCap) {
            sort_run_lds(list, n, s);
            continue;
        }
        for (uint32_t r = 0; r < n; r += kSortCap) sort_run_lds(list + r, min(kSortCap, n - r), s);
        uint32_t *src = list, *dst = tmp + off;
        for (uint32_t w = kSortCap; w < n; w <<= 1) {
            merge_pass(src, dst, n, w);
            uint32_t *t = src;
            src = dst;
            dst = t;
        }
        if (src != list) {
            for (uint32_t i = threadIdx.x; i < n; i += kSortThreads) list[i] = src[i];
            __syncthreads();
        }
    }
}

// --- load-balanced mode -------------------------------------------------

// Split level and item count per tile, from the predicted cost (or, with no
// cost model, the binned triangle count).
__global__ void kz_plan_levels(KzFrame f, uint32_t ntiles) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const float cost = f.tile_cost ? f.tile_cost[t] : (float)f.tile_counts[t];
    uint32_t level = cost > KZ_SPLIT_16 ? 2u : cost > KZ_SPLIT_4 ? 1u : 0u;
    if (f.tile_counts[t] == 0) level = 0;
    while (level > 0 && f.tile_shift < level + 1) --level;  // sub-tiles at least 2 px
    f.tile_level[t] = level;
    f.tile_item_base[t] = f.tile_counts[t] ? (1u << (2 * level)) : 0u;  // items per tile, scanned next
}

// Writes the item geometry table; items of empty tiles are not created.
__global__ void kz_plan_emit(KzFrame f, uint32_t ntiles) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles || f.tile_counts[t] == 0) return;
    const uint32_t level = f.tile_level[t], base = f.tile_item_base[t];
    for (uint32_t s = 0; s < (1u << (2 * level)); ++s) {
        f.items[base + s] = KZ_ITEM_PACK(t, level, s);
        f.item_counts[base + s] = 0;
    }
}

// Walks the work items a triangle's pixel box overlaps, one per call.
struct ItemWalk {
    uint32_t px0, py0, px1, py1;
    uint32_t tx0, tx1, ty1, tx, ty;
    uint32_t sx, sy, sx0, sx1, sy1, base, level;
    bool done;

    __device__ void enter(const KzFrame &f) {
        const uint32_t t = ty * f.tiles_x + tx;
        level = f.tile_level[t];
        base = f.tile_item_base[t];
        const uint32_t sh = f.tile_shift - level;  // sub-tile side = 1 << sh
        const uint32_t X0 = tx << f.tile_shift, Y0 = ty << f.tile_shift;
        const uint32_t last = (1u << level) - 1;
        sx0 = px0 > X0 ? (px0 - X0) >> sh : 0;
        sx1 = min((px1 - X0) >> sh, last);
        sy = py0 > Y0 ? (py0 - Y0) >> sh : 0;
        sy1 = min((py1 - Y0) >> sh, last);
        sx = sx0;
    }
    __device__ void init(const KzFrame &f, uint64_t r) {
        done = r == KZ_RECT_EMPTY;
        if (done) return;
        unpack_box(r, px0, py0, px1, py1);
        tx0 = px0 >> f.tile_shift;
        tx1 = px1 >> f.tile_shift;
        ty1 = py1 >> f.tile_shift;
        tx = tx0;
        ty = py0 >> f.tile_shift;
        enter(f);
    }
    __device__ uint32_t next(const KzFrame &f) {
        if (done) return kNone;
        const uint32_t item = base + (sy << level) + sx;
        if (++sx > sx1) {
            sx = sx0;
            if (++sy > sy1) {
                if (++tx > tx1) {
                    tx = tx0;
                    if (++ty > ty1) done = true;
                }
                if (!done) enter(f);
            }
        }
        return item;
    }
};

__global__ __launch_bounds__(256) void kz_item_count(KzFrame f) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = lane_id();
    ItemWalk w;
    w.init(f, g < f.total_tris ? f.tri_rect[g] : KZ_RECT_EMPTY);
    while (__ballot(!w.done)) wave_count(f.item_counts, w.next(f), lane);
}

__global__ __launch_bounds__(256) void kz_item_scatter(KzFrame f) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = lane_id();
    ItemWalk w;
    w.init(f, g < f.total_tris ? f.tri_rect[g] : KZ_RECT_EMPTY);
    while (__ballot(!w.done)) {
        const uint32_t slot = wave_reserve(f.item_cursor, w.next(f), lane);
        if (slot != kNone && slot < f.list_capacity) f.tile_lists[slot] = g;
    }
}

// LPT queue: items bucketed by floor(log2(count + 1)), largest bucket first.
// ctrl layout: [0] queue head, [1] item count, [2, 34) bucket counts,
// [34, 66) bucket cursors.
__device__ __forceinline__ uint32_t cost_bucket(uint32_t count) { return 31u - (31u - __clz(count + 1)); }

__global__ __launch_bounds__(256) void kz_queue_count(KzFrame f, uint32_t *ctrl) {
    const uint32_t n = ctrl[1], lane = lane_id();
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        wave_count(ctrl + 2, i < n ? cost_bucket(f.item_counts[i]) : kNone, lane);
    }
}

__global__ void kz_queue_scan(uint32_t *ctrl) {
    if (threadIdx.x != 0) return;
    uint32_t s = 0;
    for (uint32_t b = 0; b < 32; ++b) {
        ctrl[34 + b] = s;
        s += ctrl[2 + b];
    }
    ctrl[0] = 0;
}

__global__ __launch_bounds__(256) void kz_queue_emit(KzFrame f, uint32_t *ctrl) {
    const uint32_t n = ctrl[1], lane = lane_id();
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        const uint32_t slot = wave_reserve(ctrl + 34, i < n ? cost_bucket(f.item_counts[i]) : kNone, lane);
        if (i < n) f.queue[slot] = i;
    }
}

__global__ void kz_resolve4(const uint32_t *src, uint32_t *dst, uint32_t width, int32_t x0, int32_t y0, uint32_t w,
                            uint32_t h) {
    const uint64_t n = (uint64_t)w * h;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t y = y0 + i / w, x = x0 + i % w;
        const uint64_t p = y * width + x;
        const uint4 v = *reinterpret_cast<const uint4 *>(src + p * 4);
        uint32_t out = 0;
        for (int c = 0; c < 4; ++c) {
            const uint32_t sh = 8 * c;
            const uint32_t s = ((v.x >> sh) & 255u) + ((v.y >> sh) & 255u) + ((v.z >> sh) & 255u) +
                               ((v.w >> sh) & 255u) + 2u;
            out |= (s >> 2) << sh;
        }
        dst[p] = out;
    }
}

// Sums the KZ_STAT_STRIPES partial counters into the totals that follow them.
__global__ void kz_stats_reduce(unsigned long long *stats) {
    const uint32_t c = threadIdx.x;
    if (c >= KZ_STAT_COUNT) return;
    unsigned long long s = 0;
    for (uint32_t k = 0; k < KZ_STAT_STRIPES; ++k) s += stats[(size_t)k * KZ_STAT_COUNT + c];
    stats[(size_t)KZ_STAT_STRIPES * KZ_STAT_COUNT + c] = s;
}

uint32_t grid_for(uint64_t n, uint32_t block) {
    uint64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    return (uint32_t)(g == 0 ? 1 : g);
}

}  // namespace

hipError_t kz_launch_fill_u32(hipStream_t st, uint32_t *dst, uint64_t n, uint32_t value) {
    if (n == 0) return hipSuccess;
    kz_fill_u32<<<grid_for(n, 256), 256, 0, st>>>(dst, n, value);
    return hipGetLastError();
}

hipError_t kz_launch_fill_rect_u32(hipStream_t st, uint32_t *dst, uint32_t width, int32_t x0, int32_t y0, uint32_t w,
                                   uint32_t h, uint32_t spp, uint32_t value) {
    const uint64_t n = (uint64_t)w * h * spp;
    if (n == 0) return hipSuccess;
    kz_fill_rect_u32<<<grid_for(n, 256), 256, 0, st>>>(dst, width, x0, y0, w, h, spp, value);
    return hipGetLastError();
}

hipError_t kz_launch_scan(hipStream_t st, const uint32_t *counts, uint32_t *offsets, uint32_t *cursor, uint32_t n,
                          const uint32_t *n_dev, uint32_t capacity, unsigned long long *stats) {
    kz_scan<<<1, kScanThreads, 0, st>>>(counts, offsets, cursor, n, n_dev, capacity, stats, nullptr);
    return hipGetLastError();
}

hipError_t kz_launch_bin_scatter(hipStream_t st, const KzFrame &f) {
    if (f.total_tris == 0) return hipSuccess;
    kz_bin_scatter<<<(f.total_tris + 255) / 256, 256, 0, st>>>(f);
    return hipGetLastError();
}

hipError_t kz_launch_sort_lists(hipStream_t st, const uint32_t *counts, const uint32_t *offsets, uint32_t n,
                                const uint32_t *n_dev, uint32_t *lists, uint32_t *tmp, uint32_t capacity) {
    const uint32_t grid = n < 8192u ? (n ? n : 1u) : 8192u;
    kz_sort_lists<<<grid, kSortThreads, 0, st>>>(counts, offsets, n, n_dev, lists, tmp, capacity);
    return hipGetLastError();
}

hipError_t kz_launch_plan(hipStream_t st, const KzFrame &f, uint32_t ntiles, uint32_t *ctrl) {
    const uint32_t g = (ntiles + 255) / 256;
    kz_plan_levels<<<g, 256, 0, st>>>(f, ntiles);
    // items per tile -> first item of each tile; the total lands in ctrl[1]
    kz_scan<<<1, kScanThreads, 0, st>>>(f.tile_item_base, f.tile_item_base, nullptr, ntiles, nullptr, 0, nullptr,
                                        ctrl + 1);
    return hipGetLastError();
}

hipError_t kz_launch_plan_emit(hipStream_t st, const KzFrame &f, uint32_t ntiles) {
    kz_plan_emit<<<(ntiles + 255) / 256, 256, 0, st>>>(f, ntiles);
    return hipGetLastError();
}

hipError_t kz_launch_item_count(hipStream_t st, const KzFrame &f) {
    if (f.total_tris == 0) return hipSuccess;
    kz_item_count<<<(f.total_tris + 255) / 256, 256, 0, st>>>(f);
    return hipGetLastError();
}

hipError_t kz_launch_item_scatter(hipStream_t st, const KzFrame &f) {
    if (f.total_tris == 0) return hipSuccess;
    kz_item_scatter<<<(f.total_tris + 255) / 256, 256, 0, st>>>(f);
    return hipGetLastError();
}

hipError_t kz_launch_queue(hipStream_t st, const KzFrame &f, uint32_t *ctrl, uint32_t max_items) {
    const uint32_t g = max_items / 256 + 1 < 1024 ? max_items / 256 + 1 : 1024;
    kz_queue_count<<<g, 256, 0, st>>>(f, ctrl);
    kz_queue_scan<<<1, 64, 0, st>>>(ctrl);
    kz_queue_emit<<<g, 256, 0, st>>>(f, ctrl);
    return hipGetLastError();
}

hipError_t kz_launch_stats_reduce(hipStream_t st, unsigned long long *stats) {
    kz_stats_reduce<<<1, 64, 0, st>>>(stats);
    return hipGetLastError();
}

hipError_t kz_launch_resolve4(hipStream_t st, const uint32_t *src, uint32_t *dst, uint32_t width, int32_t x0,
                              int32_t y0, uint32_t w, uint32_t h) {
    if ((uint64_t)w * h == 0) return hipSuccess;
    kz_resolve4<<<grid_for((uint64_t)w * h, 256), 256, 0, st>>>(src, dst, width, x0, y0, w, h);
    return hipGetLastError();
}
