// kernels.hip — the pipeline-independent kernels of the draw path (gfx950).
//
//   kz_fill_u32 / kz_fill_rect_u32  attachment clears and counter resets (a
//                    fill kernel rather than a memset, so command buffers can
//                    be captured into hipGraphs)
//   kz_pair_scan, kz_pair_write, kz_radix_hist / scan / scatter  radix
//                    binning (the default): (triangle, tile) pairs in
//                    primitive order, sorted by tile with a stable LSD radix
//                    sort whose last pass writes the lists, the tile counts
//                    and the cost model's pair statistics
//   kz_scan          exclusive prefix sum of per-list counts
//   kz_bin_hist, kz_bin_scatter, kz_sort_short / kz_sort_lists  the other
//                    binning: pairs per tile, ids scattered into the lists
//                    in no order, each list sorted into primitive order (up
//                    to 256 ids in one wave's registers; longer ones bitonic
//                    in LDS, beyond 16K ids in runs merged with merge path)
//   kz_bin_hist_acc, kz_tile_features, kz_cost_mlp, kz_cost_train  the
//                    learned cost model: statistics, features, bf16 MFMA
//                    inference and training
//   kz_plan, kz_split_*  load-balanced mode: split expensive tiles into 4 or
//                    16 sub-tiles, cut their lists into ordered sub-tile
//                    lists, order all work items most expensive first
//   kz_resolve4      4x -> 1x box resolve, (s0+s1+s2+s3+2) >> 2 per channel
//   kz_stats_reduce  folds the striped statistics counters
//
// Binning aggregates in LDS before it touches HBM counters: a workgroup
// takes a chunk of consecutive triangles, counts its (triangle, tile) pairs
// per tile in an LDS hash table, and then adds one total per distinct tile.
// A skewed scene (thousands of triangles in a few tiles) thus costs one HBM
// atomic per (workgroup, tile), not one per wave and tile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "device_common.h"
#include "kazan/kazan_tools.h"
#include "kernels.h"

namespace {

constexpr uint32_t kScanThreads = 256;  // single-workgroup scans: 256 threads x 8 per step launch and sync faster than 1024
constexpr uint32_t kSortThreads = 1024;
constexpr uint32_t kSortCap = 16384;  // ids one workgroup sorts in LDS (64 KB)
constexpr uint32_t kNone = 0xFFFFFFFFu;

constexpr uint32_t kBinThreads = 512;
constexpr uint32_t kHashLog = 12;
constexpr uint32_t kHashSize = 1u << kHashLog;  // 32 KB of LDS: keys + values
constexpr uint32_t kHashProbes = 32;
constexpr uint32_t kAccThreads = 1024;
constexpr uint32_t kHashLogAcc = 11;
constexpr uint32_t kHashSizeAcc = 1u << kHashLogAcc;  // 64 KB: keys, values, 6 packed statistics words

constexpr uint32_t kSplitThreads = 256;

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// the statistics stripe of this workgroup
__device__ __forceinline__ unsigned long long *kz_stats_stripe_of(const KzFrame &f) {
    return f.stats + (size_t)(blockIdx.x & (KZ_STAT_STRIPES - 1)) * KZ_STAT_COUNT;
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

__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

__device__ __forceinline__ void unpack_box(uint64_t r, uint32_t &x0, uint32_t &y0, uint32_t &x1, uint32_t &y1) {
    x0 = (uint32_t)(r & 0xFFFF);
    y0 = (uint32_t)((r >> 16) & 0xFFFF);
    x1 = (uint32_t)((r >> 32) & 0xFFFF);
    y1 = (uint32_t)(r >> 48);
}

// Exclusive scan of in[0, n) into out (in == out allowed) by one workgroup of
// kScanThreads; returns the total to every thread. A step takes kScanPer
// consecutive values per thread, loaded together (the scan is bound by the
// latency of its loads, not by arithmetic).
constexpr uint32_t kScanPer = 8;

__device__ uint32_t block_exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t *cursor, uint32_t n,
                                         uint32_t *s_wave, uint32_t *s_carry) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) *s_carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < n; b += kScanThreads * kScanPer) {
        const uint32_t i0 = b + tid * kScanPer;
        uint32_t v[kScanPer], sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < kScanPer; ++j) v[j] = i0 + j < n ? in[i0 + j] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < kScanPer; ++j) sum += v[j];
        uint32_t x = sum;
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
        uint32_t run = *s_carry + wbase + x - sum;
#pragma unroll
        for (uint32_t j = 0; j < kScanPer; ++j) {
            if (i0 + j < n) {
                out[i0 + j] = run;
                if (cursor) cursor[i0 + j] = run;
            }
            run += v[j];
        }
        __syncthreads();
        if (tid == 0) *s_carry += total;
        __syncthreads();
    }
    return *s_carry;
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

// Exclusive prefix sum of counts[0, n) into offsets and cursor. With stats,
// adds the total to the list-entry statistic and, when `capacity` is
// exceeded, counts an overflow.
__global__ __launch_bounds__(kScanThreads) void kz_scan(uint32_t *counts, uint32_t *offsets, uint32_t *cursor,
                                                        uint32_t n, uint32_t capacity, unsigned long long *stats) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    __shared__ uint32_t s_carry;
    const uint32_t total = block_exclusive_scan(counts, offsets, cursor, n, s_wave, &s_carry);
    if (threadIdx.x == 0 && stats) {
        atomicAdd(&stats[KZ_STAT_LIST_ENTRIES], (unsigned long long)total);
        if (total > capacity) atomicAdd(&stats[KZ_STAT_LIST_OVERFLOW], 1ull);
    }
}

// A batch's per-tile list state before binning: counts 0, offsets all ones
// (lowered by radix binning's last pass), cut flags 0 (flags may be null).
__global__ void kz_fill_tiles(uint32_t *counts, uint32_t *offsets, uint32_t *flags, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        counts[i] = 0u;
        offsets[i] = 0xFFFFFFFFu;
        if (flags) flags[i] = 0u;
    }
}

// --- binning -----------------------------------------------------------

// LDS open-addressing tables of 1 << LOG slots: keys[] (kNone = free) and
// per-slot counters, linear probing over at most kHashProbes slots.
template <uint32_t LOG>
__device__ __forceinline__ uint32_t hash_home(uint32_t key) {
    return (key * 0x9E3779B1u) >> (32 - LOG);
}

// Adds `add` to the counter of `key`; returns its slot, or kNone when the
// key's probe window is full of other keys (the table is never emptied
// while in use, so a key that did not fit never will).
template <uint32_t LOG>
__device__ __forceinline__ uint32_t hash_add(uint32_t *keys, uint32_t *vals, uint32_t key, uint32_t add) {
    uint32_t h = hash_home<LOG>(key);
    for (uint32_t p = 0; p < kHashProbes; ++p) {
        uint32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (k == kNone) {
            const uint32_t old = atomicCAS(&keys[h], kNone, key);
            k = old == kNone ? key : old;
        }
        if (k == key) {
            atomicAdd(&vals[h], add);
            return h;
        }
        h = (h + 1) & ((1u << LOG) - 1);
    }
    return kNone;
}

// Slot of `key` in a table no longer being written, or kNone.
template <uint32_t LOG>
__device__ __forceinline__ uint32_t hash_find(const uint32_t *keys, uint32_t key) {
    uint32_t h = hash_home<LOG>(key);
    for (uint32_t p = 0; p < kHashProbes; ++p) {
        const uint32_t k = keys[h];
        if (k == key) return h;
        if (k == kNone) return kNone;
        h = (h + 1) & ((1u << LOG) - 1);
    }
    return kNone;
}

template <uint32_t LOG>
__device__ __forceinline__ void hash_clear(uint32_t *keys, uint32_t *vals) {
    for (uint32_t i = threadIdx.x; i < (1u << LOG); i += blockDim.x) {
        keys[i] = kNone;
        vals[i] = 0;
    }
}

// The tiles of one pixel box, row-major, one per call to next(); (cx, cy)
// is the tile next() returned last.
struct TileWalk {
    uint32_t x0, y0, x1, y1;
    uint32_t tx0, tx1, tx, ty, n, k, cx, cy;
    __device__ __forceinline__ void init(uint64_t r, uint32_t sh) {
        n = k = 0;
        tx0 = tx1 = tx = ty = cx = cy = 0;
        x0 = y0 = x1 = y1 = 0;
        if (r == KZ_RECT_EMPTY) return;
        unpack_box(r, x0, y0, x1, y1);
        tx0 = tx = x0 >> sh;
        tx1 = x1 >> sh;
        ty = y0 >> sh;
        n = (tx1 - tx0 + 1) * ((y1 >> sh) - ty + 1);
    }
    __device__ __forceinline__ uint32_t next(uint32_t tiles_x) {
        if (k >= n) return kNone;
        cx = tx;
        cy = ty;
        ++k;
        if (++tx > tx1) {
            tx = tx0;
            ++ty;
        }
        return cy * tiles_x + cx;
    }
    // pixels of the box inside tile (cx, cy)
    __device__ __forceinline__ uint32_t overlap(uint32_t sh) const {
        const uint32_t X0 = cx << sh, Y0 = cy << sh, X1 = X0 + (1u << sh) - 1, Y1 = Y0 + (1u << sh) - 1;
        return (min(x1, X1) - max(x0, X0) + 1) * (min(y1, Y1) - max(y0, Y0) + 1);
    }
};

// Runs of equal keys in adjacent lanes (consecutive triangles usually hit
// the same tile): the first lane of a run is its leader and speaks for
// `len` pairs; `head` is the leader's lane.
struct Run {
    bool leader;
    uint32_t len, head;
};

__device__ __forceinline__ Run run_of(uint32_t key, uint32_t lane) {
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool boundary = lane == 0 || key != prev;
    const uint64_t b = __ballot(boundary);
    Run r;
    r.leader = boundary && key != kNone;
    const uint64_t after = lane == 63 ? 0ull : (b >> (lane + 1));
    r.len = after ? (uint32_t)__builtin_ctzll(after) + 1u : 64u - lane;
    r.head = 63u - (uint32_t)__builtin_clzll(b & ((2ull << lane) - 1ull));
    return r;
}

// Pairs per tile. Workgroup b takes triangles [b * chunk, (b + 1) * chunk).
__global__ __launch_bounds__(kBinThreads) void kz_bin_hist(const uint64_t *rect, uint32_t ntris, uint32_t chunk,
                                                           uint32_t sh, uint32_t tiles_x, uint32_t *counts) {
    __shared__ uint32_t s_key[kHashSize], s_val[kHashSize];
    hash_clear<kHashLog>(s_key, s_val);
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint32_t t0 = blockIdx.x * chunk, t1 = min(t0 + chunk, ntris);
    for (uint32_t b = t0; b < t1; b += kBinThreads) {
        const uint32_t g = b + threadIdx.x;
        TileWalk w;
        w.init(g < t1 ? rect[g] : KZ_RECT_EMPTY, sh);
        const uint32_t steps = wave_max_u32(w.n);
        for (uint32_t s = 0; s < steps; ++s) {
            const uint32_t key = w.next(tiles_x);
            const Run r = run_of(key, lane);
            if (r.leader && hash_add<kHashLog>(s_key, s_val, key, r.len) == kNone) atomicAdd(&counts[key], r.len);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHashSize; i += kBinThreads) {
        const uint32_t v = s_val[i];
        if (v) atomicAdd(&counts[s_key[i]], v);
    }
}

// Adds one pair's statistics to a tile's KZ_ACC counters in HBM.
__device__ __forceinline__ void acc_pair_global(uint32_t *c, uint32_t area, bool shared, uint32_t z) {
    atomicAdd(&c[KZ_ACC_AREA], area);
    const uint32_t bin = area <= 2 ? 0u : area <= 8 ? 1u : area <= 32 ? 2u : area <= 128 ? 3u : 4u;
    atomicAdd(&c[KZ_ACC_HIST + bin], 1u);
    if (shared) atomicAdd(&c[KZ_ACC_SHARED], 1u);
    atomicMax(&c[KZ_ACC_ZMIN], 65535u - (z & 0xFFFFu));
    atomicMax(&c[KZ_ACC_ZMAX], z >> 16);
}

// The same in an LDS slot of kAccWords words with the six small counters
// packed in halves (a workgroup's chunk of at most 65535 triangles gives a
// tile at most that many pairs): [area, h0 | h1 << 16, h2 | h3 << 16,
// h4 | shared << 16, max(65535 - zmin), max(zmax)].
constexpr uint32_t kAccWords = 6;

__device__ __forceinline__ void acc_pair_lds(uint32_t *c, uint32_t stride, uint32_t area, bool shared, uint32_t z) {
    atomicAdd(&c[0], area);
    const uint32_t bin = area <= 2 ? 0u : area <= 8 ? 1u : area <= 32 ? 2u : area <= 128 ? 3u : 4u;
    atomicAdd(&c[(1 + bin / 2) * stride], (1u << (16 * (bin & 1))) + (shared && bin == 4 ? 1u << 16 : 0u));
    if (shared && bin != 4) atomicAdd(&c[3 * stride], 1u << 16);
    atomicMax(&c[4 * stride], 65535u - (z & 0xFFFFu));
    atomicMax(&c[5 * stride], z >> 16);
}

// kz_bin_hist plus the per-tile pair statistics of the cost model
// (KZ_ACC_* in acc[tile * KZ_ACC + j], zeroed by the caller). Large chunks
// (up to 8192 triangles on 1024 threads) keep the HBM flush, one set of
// counters per (workgroup, tile), to a few hundred workgroups.
__global__ __launch_bounds__(kAccThreads) void kz_bin_hist_acc(const uint64_t *rect, const uint32_t *tri_z,
                                                               uint32_t ntris, uint32_t chunk, uint32_t sh,
                                                               uint32_t tiles_x, uint32_t *counts, uint32_t *acc) {
    __shared__ uint32_t s_key[kHashSizeAcc], s_val[kHashSizeAcc], s_acc[kAccWords * kHashSizeAcc];
    hash_clear<kHashLogAcc>(s_key, s_val);
    for (uint32_t i = threadIdx.x; i < kAccWords * kHashSizeAcc; i += kAccThreads) s_acc[i] = 0;
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint32_t t0 = blockIdx.x * chunk, t1 = min(t0 + chunk, ntris);
    for (uint32_t b = t0; b < t1; b += kAccThreads) {
        const uint32_t g = b + threadIdx.x;
        TileWalk w;
        w.init(g < t1 ? rect[g] : KZ_RECT_EMPTY, sh);
        const uint32_t z = w.n ? tri_z[g] : 0u;
        const uint32_t steps = wave_max_u32(w.n);
        for (uint32_t s = 0; s < steps; ++s) {
            const uint32_t key = w.next(tiles_x);
            const Run r = run_of(key, lane);
            uint32_t slot = kNone;
            if (r.leader) {
                slot = hash_add<kHashLogAcc>(s_key, s_val, key, r.len);
                if (slot == kNone) atomicAdd(&counts[key], r.len);
            }
            slot = (uint32_t)__shfl((int)slot, (int)r.head, 64);
            if (key != kNone) {
                if (slot != kNone)
                    acc_pair_lds(&s_acc[slot], kHashSizeAcc, w.overlap(sh), w.n > 1, z);
                else
                    acc_pair_global(&acc[(size_t)key * KZ_ACC], w.overlap(sh), w.n > 1, z);
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHashSizeAcc; i += kAccThreads) {
        const uint32_t v = s_val[i];
        if (!v) continue;
        uint32_t *dst = &acc[(size_t)s_key[i] * KZ_ACC];
        atomicAdd(&counts[s_key[i]], v);
        const uint32_t w[kAccWords] = {s_acc[i], s_acc[kHashSizeAcc + i], s_acc[2 * kHashSizeAcc + i],
                                       s_acc[3 * kHashSizeAcc + i], s_acc[4 * kHashSizeAcc + i],
                                       s_acc[5 * kHashSizeAcc + i]};
        const uint32_t unpacked[KZ_ACC_ZMIN] = {w[0], w[1] & 0xFFFFu, w[1] >> 16, w[2] & 0xFFFFu, w[2] >> 16,
                                               w[3] & 0xFFFFu, w[3] >> 16};
        for (uint32_t j = 0; j < KZ_ACC_ZMIN; ++j)
            if (unpacked[j]) atomicAdd(&dst[j], unpacked[j]);
        atomicMax(&dst[KZ_ACC_ZMIN], w[4]);
        atomicMax(&dst[KZ_ACC_ZMAX], w[5]);
    }
}

// Writes triangle ids into the tile lists (positions within a list are in
// no particular order; kz_sort_lists orders them). Same chunks as
// kz_bin_hist: count per tile in LDS, reserve one range per (workgroup,
// tile), then fill the ranges from LDS cursors.
__global__ __launch_bounds__(kBinThreads) void kz_bin_scatter(const uint64_t *rect, uint32_t ntris, uint32_t chunk,
                                                              uint32_t sh, uint32_t tiles_x, uint32_t *cursor,
                                                              uint32_t *lists, uint32_t capacity) {
    __shared__ uint32_t s_key[kHashSize], s_val[kHashSize];
    hash_clear<kHashLog>(s_key, s_val);
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint32_t t0 = blockIdx.x * chunk, t1 = min(t0 + chunk, ntris);
    for (uint32_t b = t0; b < t1; b += kBinThreads) {
        const uint32_t g = b + threadIdx.x;
        TileWalk w;
        w.init(g < t1 ? rect[g] : KZ_RECT_EMPTY, sh);
        const uint32_t steps = wave_max_u32(w.n);
        for (uint32_t s = 0; s < steps; ++s) {
            const uint32_t key = w.next(tiles_x);
            const Run r = run_of(key, lane);
            if (r.leader) hash_add<kHashLog>(s_key, s_val, key, r.len);  // keys that do not fit reserve in HBM below
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHashSize; i += kBinThreads) {
        const uint32_t v = s_val[i];
        if (v) s_val[i] = atomicAdd(&cursor[s_key[i]], v);
    }
    __syncthreads();
    for (uint32_t b = t0; b < t1; b += kBinThreads) {
        const uint32_t g = b + threadIdx.x;
        TileWalk w;
        w.init(g < t1 ? rect[g] : KZ_RECT_EMPTY, sh);
        const uint32_t steps = wave_max_u32(w.n);
        for (uint32_t s = 0; s < steps; ++s) {
            const uint32_t key = w.next(tiles_x);
            const Run r = run_of(key, lane);
            uint32_t base = 0;
            if (r.leader) {
                const uint32_t slot = hash_find<kHashLog>(s_key, key);
                base = slot != kNone ? atomicAdd(&s_val[slot], r.len) : atomicAdd(&cursor[key], r.len);
            }
            base = (uint32_t)__shfl((int)base, (int)r.head, 64);
            if (key != kNone) {
                const uint32_t pos = base + (lane - r.head);
                if (pos < capacity) lists[pos] = g;
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

// Sorts the lists of at most kShortList ids, one wave per list, in
// registers: a bitonic network over 256 keys, 4 per lane (key 4 * lane + j;
// partners 1 and 2 apart are in the same lane, farther ones in lane ^ d / 4).
constexpr uint32_t kShortList = 256;

__global__ __launch_bounds__(256) void kz_sort_short(const uint32_t *counts, const uint32_t *offsets, uint32_t nl,
                                                     uint32_t *lists, uint32_t capacity) {
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * 4;
    for (uint32_t l = blockIdx.x * 4 + (threadIdx.x >> 6); l < nl; l += waves) {
        const uint32_t n = counts[l];
        const uint32_t off = offsets[l];
        if (n <= 1 || n > kShortList || (uint64_t)off + n > (uint64_t)capacity) continue;
        uint32_t v[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t i = 4 * lane + j;
            v[j] = i < n ? lists[off + i] : kNone;
        }
        for (uint32_t k = 2; k <= kShortList; k <<= 1)
            for (uint32_t d = k >> 1; d > 0; d >>= 1) {
                if (d >= 4) {
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j) {
                        const uint32_t i = 4 * lane + j;
                        const uint32_t o = (uint32_t)__shfl_xor((int)v[j], (int)(d >> 2), 64);
                        const bool keep_min = ((i & d) == 0) == ((i & k) == 0);
                        v[j] = keep_min ? min(v[j], o) : max(v[j], o);
                    }
                } else {
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j) {
                        if (j & d) continue;
                        const uint32_t i = 4 * lane + j;
                        const uint32_t a = v[j], b = v[j | d];
                        if ((a > b) == ((i & k) == 0)) {
                            v[j] = b;
                            v[j | d] = a;
                        }
                    }
                }
            }
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t i = 4 * lane + j;
            if (i < n) lists[off + i] = v[j];
        }
    }
}

// Sorts the lists [0, n) longer than kShortList, grid-stride over lists.
__global__ __launch_bounds__(kSortThreads) void kz_sort_lists(const uint32_t *counts, const uint32_t *offsets,
                                                               uint32_t nl, uint32_t *lists, uint32_t *tmp,
                                                               uint32_t capacity) {
    __shared__ uint32_t s[kSortCap];
    for (uint32_t l = blockIdx.x; l < nl; l += gridDim.x) {
        const uint32_t n = counts[l];
        const uint32_t off = offsets[l];
        if (n <= kShortList || (uint64_t)off + n > (uint64_t)capacity) continue;
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

// --- radix binning ------------------------------------------------------
//
// The batch's (triangle, tile) pairs are generated in primitive order
// (triangle by triangle, each one's tiles row-major in its pixel box) and
// sorted by tile with a stable least-significant-digit radix sort; each
// tile's run of the result is its list, already in primitive order. A pass
// costs the same whatever the skew: a tile holding a million pairs is as
// cheap as a million tiles holding one.

constexpr uint32_t kPairThreads = 1024;
constexpr uint32_t kPairPer = KZ_PAIR_CHUNK / kPairThreads;  // triangles per thread
constexpr uint32_t kBlocksPerChunk = KZ_PAIR_CHUNK / KZ_BIN_BLOCK;
constexpr uint32_t kRadixThreads = 512;
constexpr uint32_t kRadixWaves = kRadixThreads / 64;
constexpr uint32_t kRadixRounds = 8;                            // pairs per lane
constexpr uint32_t kRadixBlock = kRadixThreads * kRadixRounds;  // pairs per workgroup
constexpr uint32_t kRadixMaxBits = 9;
constexpr uint32_t kRadixMaxDigits = 1u << kRadixMaxBits;
static_assert(kRadixMaxDigits <= kRadixThreads, "digit scans take one digit per thread");

__device__ __forceinline__ uint32_t box_tiles(uint64_t r, uint32_t sh) {
    if (r == KZ_RECT_EMPTY) return 0;
    uint32_t x0, y0, x1, y1;
    unpack_box(r, x0, y0, x1, y1);
    return ((x1 >> sh) - (x0 >> sh) + 1) * ((y1 >> sh) - (y0 >> sh) + 1);
}

// The 4 x 4 sub-tiles (bit sy * 4 + sx; sub-tiles of 2^(sh - 2) px, sh >= 2)
// of tile (tx, ty) that the pixel box (x0, y0) - (x1, y1) overlaps, the box
// overlapping the tile: what a split into 16 sub-tiles sends the pair to
// (a split into 4 takes the OR of each 2 x 2 group).
__device__ __forceinline__ uint32_t box_sub_mask16(uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint32_t tx,
                                                   uint32_t ty, uint32_t sh) {
    const uint32_t X0 = tx << sh, Y0 = ty << sh, last = (1u << sh) - 1u, ssh = sh - 2u;
    const uint32_t sx0 = (max(x0, X0) - X0) >> ssh, sx1 = (min(x1, X0 + last) - X0) >> ssh;
    const uint32_t sy0 = (max(y0, Y0) - Y0) >> ssh, sy1 = (min(y1, Y0 + last) - Y0) >> ssh;
    const uint32_t row = (2u << sx1) - (1u << sx0);
    uint32_t m = 0;
    for (uint32_t sy = sy0; sy <= sy1; ++sy) m |= row << (4u * sy);
    return m;
}

// Pair offsets of the KZ_PAIR_CHUNK-triangle chunks (one workgroup; chunk c
// sums blocks [c * kBlocksPerChunk, ...) of kz_bin_count). When the pairs
// exceed `capacity`, the chunks from the first one that does not fit on are
// cut: ctrl[KZ_BIN_CUT] is that chunk and ctrl[KZ_BIN_PAIRS] the pairs
// before it (kz_pair_write then flags the tiles of the cut triangles).
__global__ __launch_bounds__(kScanThreads) void kz_pair_scan(const uint32_t *block_pairs, uint32_t nblocks,
                                                             uint32_t *chunk_offsets, uint32_t nchunks,
                                                             uint32_t capacity, uint32_t *ctrl,
                                                             unsigned long long *stats) {
    __shared__ unsigned long long s_wave[kScanThreads / 64];
    __shared__ unsigned long long s_carry, s_cut;  // s_cut: first cut chunk << 32 | its offset
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        s_carry = 0;
        s_cut = ~0ull;
    }
    __syncthreads();
    for (uint32_t c0 = 0; c0 < nchunks; c0 += kScanThreads) {
        const uint32_t c = c0 + tid;
        unsigned long long v = 0;
#pragma unroll
        for (uint32_t j = 0; j < kBlocksPerChunk; ++j) {
            const uint32_t b = c * kBlocksPerChunk + j;
            v += c < nchunks && b < nblocks ? block_pairs[b] : 0u;
        }
        unsigned long long x = v;
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const unsigned long long o = __shfl_up(x, d, 64);
            if (lane >= d) x += o;
        }
        if (lane == 63) s_wave[wave] = x;
        __syncthreads();
        unsigned long long wbase = 0, total = 0;
        for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
            const unsigned long long t = s_wave[w];
            if (w < wave) wbase += t;
            total += t;
        }
        const unsigned long long carry = s_carry;
        if (c < nchunks) {
            const unsigned long long excl = carry + wbase + x - v;
            chunk_offsets[c] = (uint32_t)min(excl, (unsigned long long)capacity);
            if (excl + v > capacity) atomicMin(&s_cut, ((unsigned long long)c << 32) | (uint32_t)min(excl, (unsigned long long)capacity));
        }
        __syncthreads();
        if (tid == 0) s_carry = carry + total;
        __syncthreads();
    }
    if (tid == 0) {
        const unsigned long long pairs = s_carry, cut = s_cut;
        const bool over = cut != ~0ull;
        ctrl[KZ_BIN_CUT] = over ? (uint32_t)(cut >> 32) : nchunks;
        ctrl[KZ_BIN_PAIRS] = over ? (uint32_t)cut : (uint32_t)pairs;
        ctrl[KZ_BIN_OVER] = over ? 1u : 0u;
        if (stats) {
            atomicAdd(&stats[KZ_STAT_LIST_ENTRIES], pairs);
            if (over) atomicAdd(&stats[KZ_STAT_LIST_OVERFLOW], 1ull);
        }
    }
}

// Writes the pairs of chunk c = blockIdx.x (triangles [c * KZ_PAIR_CHUNK,
// ...)) from its offset on: the chunk's q-th pair is the k-th tile of the
// triangle whose prefix range holds q. A cut chunk instead flags every tile
// its triangles' boxes touch; kz_tile_list reads a flagged tile as one for
// the in-order scan of all pixel boxes. With key_masks (fewer than 65536
// tiles) a key carries the pair's box_sub_mask16 in its upper 16 bits, for
// the split of the tile's list; the sort's digits read only the tile bits.
__global__ __launch_bounds__(kPairThreads) void kz_pair_write(const uint64_t *rect, uint32_t ntris, uint32_t sh,
                                                              uint32_t tiles_x, const uint32_t *chunk_offsets,
                                                              const uint32_t *ctrl, uint32_t *keys, uint32_t *vals,
                                                              uint32_t key_masks, uint32_t *tile_flag) {
    __shared__ uint64_t s_rect[KZ_PAIR_CHUNK];
    __shared__ uint32_t s_pfx[KZ_PAIR_CHUNK];
    __shared__ uint32_t s_wave[kPairThreads / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t t0 = blockIdx.x * KZ_PAIR_CHUNK;
    if (t0 >= ntris) return;
    const uint32_t n = min(KZ_PAIR_CHUNK, ntris - t0);
    if (blockIdx.x >= ctrl[KZ_BIN_CUT]) {
        for (uint32_t i = tid; i < n; i += kPairThreads) {
            const uint64_t r = rect[t0 + i];
            if (r == KZ_RECT_EMPTY) continue;
            uint32_t x0, y0, x1, y1;
            unpack_box(r, x0, y0, x1, y1);
            for (uint32_t ty = y0 >> sh; ty <= (y1 >> sh); ++ty)
                for (uint32_t tx = x0 >> sh; tx <= (x1 >> sh); ++tx) tile_flag[ty * tiles_x + tx] = 1u;
        }
        return;
    }
#pragma unroll
    for (uint32_t j = 0; j < kPairPer; ++j) {
        const uint32_t i = j * kPairThreads + tid;
        s_rect[i] = i < n ? rect[t0 + i] : KZ_RECT_EMPTY;
    }
    __syncthreads();
    // prefix of the pair counts, kPairPer consecutive triangles per thread
    uint32_t cnt[kPairPer], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPairPer; ++j) {
        cnt[j] = box_tiles(s_rect[tid * kPairPer + j], sh);
        sum += cnt[j];
    }
    uint32_t x = sum;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(x, d, 64);
        if (lane >= d) x += o;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t run = x - sum, total = 0;
    for (uint32_t w = 0; w < kPairThreads / 64; ++w) {
        const uint32_t t = s_wave[w];
        if (w < wave) run += t;
        total += t;
    }
#pragma unroll
    for (uint32_t j = 0; j < kPairPer; ++j) {
        s_pfx[tid * kPairPer + j] = run;
        run += cnt[j];
    }
    __syncthreads();
    const uint32_t out = chunk_offsets[blockIdx.x];
    for (uint32_t q = tid; q < total; q += kPairThreads) {
        // the last triangle whose prefix is <= q (s_pfx[0] = 0 <= q)
        uint32_t lo = 0, hi = KZ_PAIR_CHUNK;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pfx[mid] <= q) lo = mid;
            else hi = mid;
        }
        const uint32_t k = q - s_pfx[lo];
        uint32_t x0, y0, x1, y1;
        unpack_box(s_rect[lo], x0, y0, x1, y1);
        const uint32_t tx0 = x0 >> sh, tw = (x1 >> sh) - tx0 + 1;
        const uint32_t ky = k / tw, kx = k - ky * tw;
        const uint32_t tile = ((y0 >> sh) + ky) * tiles_x + tx0 + kx;
        keys[out + q] = key_masks ? tile | (box_sub_mask16(x0, y0, x1, y1, tx0 + kx, (y0 >> sh) + ky, sh) << 16) : tile;
        vals[out + q] = t0 + lo;
    }
}

// The lanes of `valid` whose digit (of `bits` bits) equals this lane's.
__device__ __forceinline__ uint64_t digit_match(uint32_t d, uint32_t bits, uint64_t valid) {
    uint64_t m = valid;
    for (uint32_t i = 0; i < bits; ++i) {
        const bool bit = (d >> i) & 1u;
        const uint64_t b = __ballot(bit);
        m &= bit ? b : ~b;
    }
    return m;
}

// Histogram of the digit (key >> shift) & (2^bits - 1) per block of
// kRadixBlock pairs: table[digit * tstride + block]. A wave counts one
// digit group per ballot match, so a skewed digit costs no LDS contention.
__global__ __launch_bounds__(kRadixThreads) void kz_radix_hist(const uint32_t *keys, const uint32_t *ctrl,
                                                               uint32_t shift, uint32_t bits, uint32_t *table,
                                                               uint32_t tstride) {
    __shared__ uint32_t s_cnt[kRadixMaxDigits];
    const uint32_t npairs = ctrl[KZ_BIN_PAIRS];
    const uint32_t b0 = blockIdx.x * kRadixBlock;
    if (b0 >= npairs) return;
    const uint32_t nd = 1u << bits, tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    for (uint32_t i = tid; i < nd; i += kRadixThreads) s_cnt[i] = 0;
    __syncthreads();
    const uint32_t base = b0 + wave * (kRadixBlock / kRadixWaves);
    uint32_t dg[kRadixRounds];
#pragma unroll
    for (uint32_t r = 0; r < kRadixRounds; ++r) {
        const uint32_t i = base + r * 64 + lane;
        dg[r] = i < npairs ? (keys[i] >> shift) & (nd - 1) : kNone;
    }
#pragma unroll
    for (uint32_t r = 0; r < kRadixRounds; ++r) {
        const bool ok = dg[r] != kNone;
        const uint64_t m = digit_match(dg[r], bits, __ballot(ok));
        if (ok && rank_in(m) == 0) atomicAdd(&s_cnt[dg[r]], (uint32_t)__popcll(m));
    }
    __syncthreads();
    for (uint32_t i = tid; i < nd; i += kRadixThreads) table[(size_t)i * tstride + blockIdx.x] = s_cnt[i];
}

// Exclusive scan of each digit's row over the blocks in use (one workgroup
// per digit); totals[digit] = the digit's pairs.
__global__ __launch_bounds__(kScanThreads) void kz_radix_scan(uint32_t *table, uint32_t tstride,
                                                              const uint32_t *ctrl, uint32_t *totals) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    __shared__ uint32_t s_carry;
    const uint32_t nb = (ctrl[KZ_BIN_PAIRS] + kRadixBlock - 1) / kRadixBlock;
    uint32_t *row = table + (size_t)blockIdx.x * tstride;
    const uint32_t total = block_exclusive_scan(row, row, nullptr, nb, s_wave, &s_carry);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// Exclusive scans of a[0, n) and b[0, n) in place, n <= kRadixThreads.
__device__ void radix_scan2(uint32_t *a, uint32_t *b, uint32_t n, uint32_t *s_wa, uint32_t *s_wb) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t va = tid < n ? a[tid] : 0u, vb = tid < n ? b[tid] : 0u;
    uint32_t xa = va, xb = vb;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t oa = __shfl_up(xa, d, 64), ob = __shfl_up(xb, d, 64);
        if (lane >= d) {
            xa += oa;
            xb += ob;
        }
    }
    if (lane == 63) {
        s_wa[wave] = xa;
        s_wb[wave] = xb;
    }
    __syncthreads();
    uint32_t ba = 0, bb = 0;
    for (uint32_t w = 0; w < wave; ++w) {
        ba += s_wa[w];
        bb += s_wb[w];
    }
    if (tid < n) {
        a[tid] = ba + xa - va;
        b[tid] = bb + xb - vb;
    }
    __syncthreads();
}

// Segmented sum / max over runs of adjacent lanes (same[s]: lane + 2^s is
// in this lane's run): the run's first lane ends with the run's total.
__device__ __forceinline__ uint32_t seg_sum(uint32_t v, const bool *same) {
    for (uint32_t s = 0, d = 1; d < 64; ++s, d <<= 1) {
        const uint32_t o = __shfl_down(v, d, 64);
        if (same[s]) v += o;
    }
    return v;
}
__device__ __forceinline__ uint32_t seg_max(uint32_t v, const bool *same) {
    for (uint32_t s = 0, d = 1; d < 64; ++s, d <<= 1) {
        const uint32_t o = __shfl_down(v, d, 64);
        if (same[s]) v = max(v, o);
    }
    return v;
}

// The cost model's per-tile pair statistics, gathered in the last pass.
struct RadixAcc {
    const uint64_t *rect;
    const uint32_t *tri_z;
    uint32_t *acc;  // KZ_ACC words per tile, or null
    uint32_t sh, tiles_x;
};

// One stable counting-sort pass by the digit (key >> shift) & (2^bits - 1)
// over blocks of kRadixBlock pairs. A wave ranks its 8 rounds of 64 pairs,
// in pair order, by digit-match ballots against a count per digit of its
// own, so equal digits keep their order within the wave; waves are ordered
// by a scan of those counts, blocks by the scanned table, digits by the
// scanned totals. The block is regrouped by digit in LDS so that each
// digit's run is written contiguously. The last pass (lists != null)
// writes only the triangle ids, into the tile lists, adds each tile's run
// lengths to tile_counts and lowers tile_offsets (all ones before) to the
// runs' first positions, so that the lists need no scan of the counts
// (and, with acc, adds the pair statistics). With list_mask (keys that
// carry a sub-tile mask above 16 tile bits) the last pass also writes each
// entry's mask beside the list.
__global__ __launch_bounds__(kRadixThreads) void kz_radix_scatter(const uint32_t *keys_in, const uint32_t *vals_in,
                                                                  uint32_t *keys_out, uint32_t *vals_out,
                                                                  uint32_t *lists, uint16_t *list_mask,
                                                                  uint32_t *tile_counts,
                                                                  uint32_t *tile_offsets,
                                                                  const uint32_t *ctrl, const uint32_t *table,
                                                                  uint32_t tstride, const uint32_t *totals,
                                                                  uint32_t shift, uint32_t bits, RadixAcc acc) {
    __shared__ uint32_t s_wcnt[kRadixWaves * kRadixMaxDigits];
    __shared__ uint32_t s_lstart[kRadixMaxDigits], s_gbase[kRadixMaxDigits];
    __shared__ uint32_t s_key[kRadixBlock], s_val[kRadixBlock];
    __shared__ uint32_t s_wa[kRadixWaves], s_wb[kRadixWaves];
    const uint32_t npairs = ctrl[KZ_BIN_PAIRS];
    const uint32_t b0 = blockIdx.x * kRadixBlock;
    if (b0 >= npairs) return;
    const uint32_t nitems = min(kRadixBlock, npairs - b0);
    const uint32_t nd = 1u << bits, dmask = nd - 1, tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    for (uint32_t i = tid; i < kRadixWaves * nd; i += kRadixThreads) s_wcnt[i] = 0;
    __syncthreads();

    // ranks within the wave
    const uint32_t base = b0 + wave * (kRadixBlock / kRadixWaves);
    uint32_t key[kRadixRounds], val[kRadixRounds], rk[kRadixRounds];
#pragma unroll
    for (uint32_t r = 0; r < kRadixRounds; ++r) {
        const uint32_t i = base + r * 64 + lane;
        const bool ok = i < npairs;
        key[r] = ok ? keys_in[i] : kNone;
        val[r] = ok ? vals_in[i] : 0u;
    }
    uint32_t *wcnt = s_wcnt + wave * nd;
#pragma unroll
    for (uint32_t r = 0; r < kRadixRounds; ++r) {
        const bool ok = key[r] != kNone;
        const uint32_t d = (key[r] >> shift) & dmask;
        const uint64_t m = digit_match(d, bits, __ballot(ok));
        const uint32_t lead = ok ? (uint32_t)__builtin_ctzll(m) : lane;
        uint32_t old = 0;
        if (ok && lead == lane) {
            old = wcnt[d];
            wcnt[d] = old + (uint32_t)__popcll(m);
        }
        old = (uint32_t)__shfl((int)old, (int)lead, 64);
        rk[r] = old + rank_in(m);
    }
    __syncthreads();

    // waves' starts per digit, the block's digit starts, the global starts
    for (uint32_t d = tid; d < nd; d += kRadixThreads) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < kRadixWaves; ++w) {
            const uint32_t c = s_wcnt[w * nd + d];
            s_wcnt[w * nd + d] = s;
            s += c;
        }
        s_lstart[d] = s;
        s_gbase[d] = totals[d];
    }
    __syncthreads();
    radix_scan2(s_lstart, s_gbase, nd, s_wa, s_wb);
    for (uint32_t d = tid; d < nd; d += kRadixThreads) s_gbase[d] += table[(size_t)d * tstride + blockIdx.x];
#pragma unroll
    for (uint32_t r = 0; r < kRadixRounds; ++r) {
        if (key[r] == kNone) continue;
        const uint32_t d = (key[r] >> shift) & dmask;
        const uint32_t p = s_lstart[d] + wcnt[d] + rk[r];  // < nitems by construction
        if (p < nitems) {
            s_key[p] = key[r];
            s_val[p] = val[r];
        }
    }
    __syncthreads();

    // out, digit run by digit run
    for (uint32_t j0 = 0; j0 < nitems; j0 += kRadixThreads) {
        const uint32_t j = j0 + tid;
        const bool ok = j < nitems;
        const uint32_t kk = ok ? s_key[j] : kNone, v = ok ? s_val[j] : 0u;
        uint32_t g = 0;
        if (ok) {
            const uint32_t d = (kk >> shift) & dmask;
            g = s_gbase[d] + (j - s_lstart[d]);  // < npairs by construction
            if (g < npairs) {
                if (lists) {
                    lists[g] = v;
                    if (list_mask) list_mask[g] = (uint16_t)(kk >> 16);
                } else {
                    keys_out[g] = kk;
                    vals_out[g] = v;
                }
            }
        }
        if (!lists) continue;
        const uint32_t k = ok && list_mask ? kk & 0xFFFFu : kk;  // the tile
        // the last pass's block is sorted by tile: one update per run of a tile
        const Run run = run_of(k, lane);
        if (run.leader) {
            atomicAdd(&tile_counts[k], run.len);
            atomicMin(&tile_offsets[k], g);
        }
        if (!acc.acc) continue;
        uint32_t area = 0, lo = 0, hi = 0, zinv = 0, zmax = 0;
        if (ok) {
            uint32_t x0, y0, x1, y1;
            const uint64_t r = acc.rect[v];
            unpack_box(r, x0, y0, x1, y1);
            const uint32_t X0 = (k % acc.tiles_x) << acc.sh, Y0 = (k / acc.tiles_x) << acc.sh;
            const uint32_t X1 = X0 + (1u << acc.sh) - 1, Y1 = Y0 + (1u << acc.sh) - 1;
            area = (min(x1, X1) - max(x0, X0) + 1) * (min(y1, Y1) - max(y0, Y0) + 1);
            const uint32_t bin = area <= 2 ? 0u : area <= 8 ? 1u : area <= 32 ? 2u : area <= 128 ? 3u : 4u;
            const uint32_t shared = box_tiles(r, acc.sh) > 1 ? 1u : 0u;
            // six counters of at most 64 per wave, 10 bits each
            lo = bin < 3 ? 1u << (10 * bin) : 0u;
            hi = (bin >= 3 ? 1u << (10 * (bin - 3)) : 0u) + (shared << 20);
            const uint32_t z = acc.tri_z[v];
            zinv = 65535u - (z & 0xFFFFu);
            zmax = z >> 16;
        }
        bool same[6];
        for (uint32_t s = 0, d = 1; d < 64; ++s, d <<= 1) {
            // every lane takes part in the shuffle (a lane left out reads as 0)
            const uint32_t h = (uint32_t)__shfl_down((int)run.head, d, 64);
            same[s] = lane + d < 64 && h == run.head;
        }
        area = seg_sum(area, same);
        lo = seg_sum(lo, same);
        hi = seg_sum(hi, same);
        zinv = seg_max(zinv, same);
        zmax = seg_max(zmax, same);
        if (run.leader) {
            uint32_t *c = acc.acc + (size_t)k * KZ_ACC;
            atomicAdd(&c[KZ_ACC_AREA], area);
            const uint32_t h[6] = {lo & 1023u, (lo >> 10) & 1023u, lo >> 20, hi & 1023u, (hi >> 10) & 1023u, hi >> 20};
            for (uint32_t b = 0; b < 5; ++b)
                if (h[b]) atomicAdd(&c[KZ_ACC_HIST + b], h[b]);
            if (h[5]) atomicAdd(&c[KZ_ACC_SHARED], h[5]);
            atomicMax(&c[KZ_ACC_ZMIN], zinv);
            atomicMax(&c[KZ_ACC_ZMAX], zmax);
        }
    }
}

// --- load-balanced mode -------------------------------------------------

// A tile's cost in list entries: the binned triangle count, plus the
// model's predicted fragments weighted as entries when a cost model ran.
// The prediction is capped at what n triangles can cover (every sample of
// the tile, at most 4 per pixel, once each), so an input far outside the
// training range cannot make a tile's cost, or the frame's total, infinite.
// (n binned triangles, pred = the model's prediction for the tile)
__device__ __forceinline__ float tile_cost(const KzFrame &f, uint32_t n, float pred) {
    if (!f.tile_pred) return (float)n;
    const float cap = (float)n * (float)(4u << (2 * f.tile_shift));
    return (float)n + f.frag_weight * fminf(pred, cap);
}

// Split level of a non-empty tile: a tile above the fair share `share`
// (the frame's cost over the raster's resident workgroups) would hold up
// the frame's end, so it becomes 4 items, or 16 above 4 shares. A tile
// whose list overflowed stays whole: its triangles come from the in-order
// scan.
__device__ __forceinline__ uint32_t plan_level(const KzFrame &f, uint32_t t, uint32_t n, float pred, float share,
                                               float &cost) {
    cost = tile_cost(f, n, pred);
    uint32_t level = cost > 4.0f * share ? 2u : cost > share ? 1u : 0u;
    level = min(level, f.split_max_level);
    while (level > 0 && f.tile_shift < level + 1) --level;  // sub-tiles at least 2 px wide
    if (level > 0) {
        uint32_t off, cnt;
        kz_tile_list(f, t, off, cnt);
        if ((uint64_t)off + cnt > (uint64_t)f.list_capacity) level = 0;
    }
    return level;
}

// floor(log2(cost + 1)), the LPT bucket of an item
__device__ __forceinline__ uint32_t cost_bucket(float cost) {
    const float c = fminf(fmaxf(cost, 0.0f), 2.0e9f);
    return 31u - __clz((uint32_t)c + 1u);
}

// Reserves `items` slots of LDS counter s[bucket] for every active lane and
// returns the lane's first slot. Lanes with one item and the same bucket
// share one atomic (a frame puts most tiles into a few buckets).
__device__ __forceinline__ uint32_t bucket_reserve(uint32_t *s, uint32_t bucket, uint32_t items, bool active) {
    const uint32_t lane = lane_id();
    const bool single = active && items == 1;
    const uint64_t m = digit_match(bucket, 5, __ballot(single));
    uint32_t base = 0;
    if (single && rank_in(m) == 0) base = atomicAdd(&s[bucket], (uint32_t)__popcll(m));
    const uint32_t lead = single ? (uint32_t)__builtin_ctzll(m) : lane;
    base = (uint32_t)__shfl((int)base, (int)lead, 64);
    if (single) return base + rank_in(m);
    return active ? atomicAdd(&s[bucket], items) : 0u;
}

// The plan runs on many workgroups, in three kernels over the same tile
// partition (workgroup w: tiles [w * kPlanTiles, ...), kPlanPer per thread):
//   kz_plan_cost   each workgroup's part of the frame's total cost (only
//                  with a model or sort binning: otherwise the total is
//                  the binned pairs);
//   kz_plan_count  the fair share (every workgroup sums the parts in the
//                  same order, so all agree on it), then per workgroup the
//                  items of each LPT bucket, its split tiles and their chunks;
//   kz_plan_queue  from those counts, each workgroup's place in every bucket
//                  (buckets most expensive first) and its split slots and
//                  chunks, then the queue, the levels and the split tables.
// table: [workgroups * kPlanCols] counts, then a float part of the frame's
// cost per kPartTiles tiles.
constexpr uint32_t kPlanWg = 256;
constexpr uint32_t kPlanPer = 1;
constexpr uint32_t kPlanTiles = kPlanWg * kPlanPer;
constexpr uint32_t kPartTiles = 64;
static_assert(kPlanPer == 1 && kPartTiles == 64, "a wave's tiles are one part");
constexpr uint32_t kPlanCols = 34;  // 32 buckets, split tiles, chunks
constexpr uint32_t kPlanGroups = kPlanWg / kPlanCols;  // threads per column when summing the table

// the parts, after the counts of `plan_wgs` workgroups
__device__ __forceinline__ float *plan_parts(uint32_t *table, uint32_t plan_wgs) {
    return (float *)(table + plan_wgs * kPlanCols);
}

__device__ __forceinline__ void plan_loads(const KzFrame &f, uint32_t ntiles, uint32_t *t, uint32_t *n, float *pred,
                                           uint32_t *off) {
#pragma unroll
    for (uint32_t j = 0; j < kPlanPer; ++j) {
        t[j] = blockIdx.x * kPlanTiles + j * kPlanWg + threadIdx.x;
        off[j] = 0;
        n[j] = 0;
        if (t[j] < ntiles) kz_tile_list(f, t[j], off[j], n[j]);
        pred[j] = t[j] < ntiles && f.tile_pred ? f.tile_pred[t[j]] : 0.0f;
    }
}

__global__ __launch_bounds__(kPlanWg) void kz_plan_cost(KzFrame f, uint32_t ntiles, uint32_t *table) {
    uint32_t t[kPlanPer], n[kPlanPer], off[kPlanPer];
    float pred[kPlanPer];
    plan_loads(f, ntiles, t, n, pred, off);
    float part = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < kPlanPer; ++j)
        if (n[j]) part += tile_cost(f, n[j], pred[j]);
    for (uint32_t d = 32; d > 0; d >>= 1) part += __shfl_xor(part, d, 64);
    // (wave w of workgroup b: tiles [(4 b + w) * 64, ...))
    if ((threadIdx.x & 63) == 0) plan_parts(table, gridDim.x)[blockIdx.x * (kPlanWg / 64) + (threadIdx.x >> 6)] = part;
}

// Without a model a tile's cost is its list entries, and the frame's total
// is the pairs radix binning wrote: kz_plan_cost is not needed.
__device__ __forceinline__ bool plan_cost_is_pairs(const KzFrame &f) { return !f.tile_pred && f.bin_ctrl; }

// The fair share of one resident raster workgroup. Every workgroup sums
// the parts by the same tree (lane-strided, then waves in order), so all
// of them compute the same float.
__device__ __forceinline__ float plan_share(const KzFrame &f, uint32_t ntiles, const uint32_t *table,
                                            float *s_wave) {
    const float *part = (const float *)(table + gridDim.x * kPlanCols);  // (plan_parts)
    const bool pairs = plan_cost_is_pairs(f);
    const uint32_t nparts = (ntiles + kPartTiles - 1) / kPartTiles;
    float v = 0.0f;
    if (!pairs)
        for (uint32_t w = threadIdx.x; w < nparts; w += kPlanWg) v += part[w];
    for (uint32_t d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = v;
    __syncthreads();
    float total = 0.0f;
    for (uint32_t w = 0; w < kPlanWg / 64; ++w) total += s_wave[w];
    if (pairs) total = (float)f.bin_ctrl[KZ_BIN_PAIRS];
    return fmaxf(total / (float)max(f.raster_slots, 1u), KZ_SPLIT_MIN);
}

__global__ __launch_bounds__(kPlanWg) void kz_plan_count(KzFrame f, uint32_t ntiles, uint32_t *table) {
    __shared__ uint32_t s_cnt[kPlanCols];
    __shared__ float s_wave[kPlanWg / 64];
    if (threadIdx.x < kPlanCols) s_cnt[threadIdx.x] = 0;
    uint32_t t[kPlanPer], n[kPlanPer], off[kPlanPer];
    float pred[kPlanPer];
    plan_loads(f, ntiles, t, n, pred, off);
    const float share = plan_share(f, ntiles, table, s_wave);
#pragma unroll
    for (uint32_t j = 0; j < kPlanPer; ++j) {
        float cost = 0.0f;
        const uint32_t level = n[j] ? plan_level(f, t[j], n[j], pred[j], share, cost) : 0u;
        const uint32_t items = 1u << (2 * level);
        bucket_reserve(s_cnt, cost_bucket(cost / (float)items), items, n[j] != 0);
        if (n[j] && level) {
            atomicAdd(&s_cnt[32], 1u);
            atomicAdd(&s_cnt[33], (n[j] + KZ_SPLIT_CHUNK - 1) / KZ_SPLIT_CHUNK);
        }
    }
    __syncthreads();
    if (threadIdx.x < kPlanCols) table[blockIdx.x * kPlanCols + threadIdx.x] = s_cnt[threadIdx.x];
}

__global__ __launch_bounds__(kPlanWg) void kz_plan_queue(KzFrame f, uint32_t ntiles, const uint32_t *table) {
    __shared__ uint32_t s_cur[kPlanCols], s_tot[kPlanCols];
    __shared__ uint32_t s_gb[kPlanGroups][kPlanCols], s_gt[kPlanGroups][kPlanCols];
    __shared__ float s_wave[kPlanWg / 64];
    uint32_t t[kPlanPer], n[kPlanPer], off[kPlanPer];
    float pred[kPlanPer];
    plan_loads(f, ntiles, t, n, pred, off);
    // this workgroup's start in every column: the columns' sums over the
    // workgroups before it, and their totals over all (kPlanGroups threads
    // per column, each taking every kPlanGroups-th workgroup)
    if (threadIdx.x < kPlanGroups * kPlanCols) {
        const uint32_t col = threadIdx.x % kPlanCols, grp = threadIdx.x / kPlanCols;
        uint32_t before = 0, total = 0;
        for (uint32_t w = grp; w < gridDim.x; w += kPlanGroups) {
            const uint32_t v = table[w * kPlanCols + col];
            if (w < blockIdx.x) before += v;
            total += v;
        }
        s_gb[grp][col] = before;
        s_gt[grp][col] = total;
    }
    __syncthreads();
    if (threadIdx.x < kPlanCols) {
        uint32_t before = 0, total = 0;
        for (uint32_t g = 0; g < kPlanGroups; ++g) {
            before += s_gb[g][threadIdx.x];
            total += s_gt[g][threadIdx.x];
        }
        s_cur[threadIdx.x] = before;
        s_tot[threadIdx.x] = total;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t start = 0;  // buckets most expensive first
        for (int b = 31; b >= 0; --b) {
            s_cur[b] += start;
            start += s_tot[b];
        }
        if (blockIdx.x == 0) {
            for (uint32_t x = 0; x < KZ_QUEUES; ++x) f.ctrl[KZ_CTRL_HEADS + x * KZ_CTRL_HEAD_STRIDE] = 0;
            f.ctrl[KZ_CTRL_SPLIT_CURSOR] = 0;
            f.ctrl[KZ_CTRL_ITEMS] = start;
            f.ctrl[KZ_CTRL_SPLITS] = s_tot[32];
            f.ctrl[KZ_CTRL_CHUNKS] = min(s_tot[33], f.max_chunks);
        }
    }
    const float share = plan_share(f, ntiles, table, s_wave);  // (its barrier also publishes s_cur)
#pragma unroll
    for (uint32_t j = 0; j < kPlanPer; ++j) {
        float cost = 0.0f;
        const uint32_t level = n[j] ? plan_level(f, t[j], n[j], pred[j], share, cost) : 0u;
        const uint32_t items = 1u << (2 * level);
        const uint32_t pos = bucket_reserve(s_cur, cost_bucket(cost / (float)items), items, n[j] != 0);
        if (!n[j]) continue;
        f.tile_level[t[j]] = level;
        for (uint32_t s = 0; s < items; ++s) f.queue[pos + s] = KZ_ITEM_PACK(t[j], level, s);
        if (level) {
            const uint32_t nc = (n[j] + KZ_SPLIT_CHUNK - 1) / KZ_SPLIT_CHUNK;
            const uint32_t slot = atomicAdd(&s_cur[32], 1u);
            const uint32_t c0 = atomicAdd(&s_cur[33], nc);
            f.split_slot[t[j]] = slot;
            f.split_info[2 * slot] = c0;
            f.split_info[2 * slot + 1] = t[j];
            for (uint32_t s = 0; s < 16; ++s) f.split_counts[slot * 16 + s] = 0;  // kz_split_totals writes
            for (uint32_t c = 0; c < nc && c0 + c < f.max_chunks; ++c)
                ((uint4 *)f.chunk_desc)[c0 + c] = make_uint4(off[j] + c * KZ_SPLIT_CHUNK, t[j], (slot << 2) | level,
                                                             min(KZ_SPLIT_CHUNK, n[j] - c * KZ_SPLIT_CHUNK));
        }
    }
}

// A KZ_SPLIT_CHUNK-entry piece of a split tile's (sorted) list, from the
// descriptor kz_plan_queue wrote (one load).
struct Chunk {
    const uint32_t *ids;
    uint32_t n, nsub, level, ssh;
    uint32_t slot;  // split slot
    int32_t X0, Y0;
};

__device__ __forceinline__ Chunk chunk_at(const KzFrame &f, uint32_t c) {
    Chunk k;
    const uint4 d = ((const uint4 *)f.chunk_desc)[c];
    const uint32_t t = d.y;
    k.ids = f.tile_lists + d.x;
    k.n = d.w;
    k.slot = d.z >> 2;
    k.level = d.z & 3u;
    k.nsub = 1u << (2 * k.level);
    k.ssh = f.tile_shift - k.level;
    k.X0 = (int32_t)((t % f.tiles_x) << f.tile_shift);
    k.Y0 = (int32_t)((t / f.tiles_x) << f.tile_shift);
    return k;
}

// Sub-tiles (bit sy << level | sx) of a split tile that a pixel box overlaps.
__device__ __forceinline__ uint32_t sub_mask(uint64_t r, const Chunk &k) {
    uint32_t x0, y0, x1, y1;
    unpack_box(r, x0, y0, x1, y1);
    const int32_t last = (1 << k.level) - 1;
    const int32_t sx0 = max((int32_t)x0 - k.X0, 0) >> k.ssh, sx1 = min(((int32_t)x1 - k.X0) >> k.ssh, last);
    const int32_t sy0 = max((int32_t)y0 - k.Y0, 0) >> k.ssh, sy1 = min(((int32_t)y1 - k.Y0) >> k.ssh, last);
    const uint32_t row = (2u << sx1) - (1u << sx0);
    uint32_t m = 0;
    for (int32_t sy = sy0; sy <= sy1; ++sy) m |= row << (sy << k.level);
    return m;
}

// Entries per sub-tile of every chunk (one wave per chunk), kept per chunk,
// and each entry's sub-tile mask (split_mask), so that the scatter reads
// the masks instead of gathering pixel boxes again.
constexpr uint32_t kSplitRounds = KZ_SPLIT_CHUNK / 64;

// A split tile's sub-tiles of an entry from its 4 x 4 mask (box_sub_mask16):
// at level 2 the mask itself, at level 1 the OR of each 2 x 2 group (bit
// sy * 2 + sx).
__device__ __forceinline__ uint32_t fold_sub_mask(uint32_t m16, uint32_t level) {
    if (level >= 2) return m16;
    const uint32_t c = m16 | (m16 >> 1);  // bit 4 sy + 2 sx: columns 2 sx and 2 sx + 1
    const uint32_t r = c | (c >> 4);      // bit 8 sy + 2 sx: rows 2 sy and 2 sy + 1 as well
    return (r & 1u) | ((r >> 1) & 2u) | ((r >> 6) & 4u) | ((r >> 7) & 8u);
}

// The sub-tile masks of a chunk's entries (lane l: entries r * 64 + l):
// from the masks radix binning wrote beside the list, else from the
// entries' pixel boxes (all ids, then all boxes, are requested before any
// is used, so a wave waits for two loads per chunk rather than two per 64
// entries).
__device__ __forceinline__ void chunk_masks(const KzFrame &f, const Chunk &k, uint32_t lane, uint32_t *m) {
    if (f.list_mask) {
        const uint16_t *lm = f.list_mask + (k.ids - f.tile_lists);
#pragma unroll
        for (uint32_t r = 0; r < kSplitRounds; ++r) {
            const uint32_t e = r * 64 + lane;
            m[r] = e < k.n ? fold_sub_mask(lm[e], k.level) : 0u;
        }
        return;
    }
    uint32_t id[kSplitRounds];
    uint64_t box[kSplitRounds];
#pragma unroll
    for (uint32_t r = 0; r < kSplitRounds; ++r) {
        const uint32_t e = r * 64 + lane;
        id[r] = e < k.n ? k.ids[e] : kNone;
    }
#pragma unroll
    for (uint32_t r = 0; r < kSplitRounds; ++r) box[r] = id[r] != kNone ? f.tri_rect[id[r]] : KZ_RECT_EMPTY;
#pragma unroll
    for (uint32_t r = 0; r < kSplitRounds; ++r) m[r] = box[r] != KZ_RECT_EMPTY ? sub_mask(box[r], k) : 0u;
}

// A chunk's entries per sub-tile (NSUB = 4 or 16 sub-tiles: the counts are
// wave-uniform, held in scalar registers), and its entries' masks.
template <uint32_t NSUB>
__device__ __forceinline__ void split_count_chunk(const KzFrame &f, const Chunk &k, uint32_t c, uint32_t lane,
                                                  const uint32_t *m) {
    uint32_t cnt[NSUB];
#pragma unroll
    for (uint32_t s = 0; s < NSUB; ++s) cnt[s] = 0;
#pragma unroll
    for (uint32_t r = 0; r < kSplitRounds; ++r) {
        if (r * 64 >= k.n) break;
        f.split_mask[(size_t)c * KZ_SPLIT_CHUNK + r * 64 + lane] = (uint16_t)m[r];
#pragma unroll
        for (uint32_t s = 0; s < NSUB; ++s) cnt[s] += (uint32_t)__popcll(__ballot((m[r] >> s) & 1u));
    }
    uint32_t mine = 0;  // lane s: sub-tile s's count
#pragma unroll
    for (uint32_t s = 0; s < NSUB; ++s) mine = lane == s ? cnt[s] : mine;
    if (lane < 16) f.chunk_counts[c * 16 + lane] = mine;
}

__global__ __launch_bounds__(kSplitThreads) void kz_split_count(KzFrame f) {
    const uint32_t lane = lane_id();
    const uint32_t nch = f.ctrl[KZ_CTRL_CHUNKS];
    const uint32_t waves = gridDim.x * (kSplitThreads / 64);
    for (uint32_t c = blockIdx.x * (kSplitThreads / 64) + (threadIdx.x >> 6); c < nch; c += waves) {
        const Chunk k = chunk_at(f, c);
        uint32_t m[kSplitRounds];
        chunk_masks(f, k, lane, m);
        if (k.level == 2) split_count_chunk<16>(f, k, c, lane, m);
        else split_count_chunk<4>(f, k, c, lane, m);
    }
}

// Per split slot (one wave each): its 16 sub-tile totals (split_counts),
// and each of its chunks' counts (chunk_counts) turned into the position of
// the chunk's first entry within each sub-tile list, 64 chunks at a time
// (lane l: chunk j0 + l, its 16 counts loaded at once; per sub-tile a scan
// over the wave; lane s < 16 carries sub-tile s's running total).
__global__ __launch_bounds__(kSplitThreads) void kz_split_totals(KzFrame f) {
    const uint32_t lane = lane_id();
    const uint32_t n_slots = f.ctrl[KZ_CTRL_SPLITS];
    const uint32_t waves = gridDim.x * (kSplitThreads / 64);
    for (uint32_t i = blockIdx.x * (kSplitThreads / 64) + (threadIdx.x >> 6); i < n_slots; i += waves) {
        const uint32_t c0 = f.split_info[2 * i], t = f.split_info[2 * i + 1];
        const uint32_t nch = (f.tile_counts[t] + KZ_SPLIT_CHUNK - 1) / KZ_SPLIT_CHUNK;
        uint32_t run = 0;
        for (uint32_t j0 = 0; j0 < nch; j0 += 64) {
            const uint32_t j = j0 + lane;
            uint4 *row = (uint4 *)(f.chunk_counts + (size_t)(c0 + (j < nch ? j : 0)) * 16);
            uint32_t v[16];
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                const uint4 w = j < nch ? row[q] : make_uint4(0u, 0u, 0u, 0u);
                v[4 * q] = w.x;
                v[4 * q + 1] = w.y;
                v[4 * q + 2] = w.z;
                v[4 * q + 3] = w.w;
            }
#pragma unroll
            for (uint32_t s16 = 0; s16 < 16; ++s16) {
                uint32_t y = v[s16];
                for (uint32_t d = 1; d < 64; d <<= 1) {
                    const uint32_t o = __shfl_up(y, d, 64);
                    if (lane >= d) y += o;
                }
                const uint32_t b = readlane(run, s16);
                const uint32_t tot = readlane(y, 63);
                v[s16] = b + y - v[s16];
                if (lane == s16) run += tot;
            }
            if (j < nch) {
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) row[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
            }
        }
        if (lane < 16) f.split_counts[i * 16 + lane] = run;
    }
}

// The offsets of the sub-tile lists, one thread per split slot (its 16
// lists consecutive): a wave reserves its slots' storage with one atomic on
// the batch's cursor (reset by kz_plan_queue), so slots land in any order.
__global__ __launch_bounds__(kSplitThreads) void kz_split_offsets(KzFrame f) {
    const uint32_t lane = lane_id();
    const uint32_t n_slots = f.ctrl[KZ_CTRL_SPLITS];
    const uint32_t i = blockIdx.x * kSplitThreads + threadIdx.x;
    if (blockIdx.x * kSplitThreads >= n_slots) return;
    uint32_t c[16], v = 0;
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
        c[s] = i < n_slots ? f.split_counts[i * 16 + s] : 0u;
        v += c[s];
    }
    uint32_t x = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(x, d, 64);
        if (lane >= d) x += o;
    }
    const uint32_t total = readlane(x, 63);
    uint32_t base = 0;
    if (lane == 63 && total) {
        base = atomicAdd(&f.ctrl[KZ_CTRL_SPLIT_CURSOR], total);
        atomicMax(&f.ctrl[KZ_CTRL_SPLIT_MAX], base + total);
        atomicAdd(&kz_stats_stripe_of(f)[KZ_STAT_SPLIT_ENTRIES], (unsigned long long)total);
    }
    uint32_t run = readlane(base, 63) + x - v;
    if (i < n_slots) {
#pragma unroll
        for (uint32_t s = 0; s < 16; ++s) {
            f.split_offsets[i * 16 + s] = run;
            run += c[s];
        }
    }
}

// One chunk's entries into the sub-tile lists (NSUB = 4 or 16 sub-tiles).
template <uint32_t NSUB>
__device__ __forceinline__ void split_scatter_chunk(const KzFrame &f, const Chunk &k, uint32_t c, uint32_t lane) {
    // next position in sub-tile s's list (the slot's list offset plus the
    // entries of the tile's earlier chunks), wave-uniform: scalar registers
    const uint32_t start =
        lane < NSUB ? f.split_offsets[k.slot * 16 + lane] + f.chunk_counts[c * 16 + lane] : 0u;
    uint32_t base[NSUB];
#pragma unroll
    for (uint32_t s = 0; s < NSUB; ++s) base[s] = readlane(start, s);
    uint32_t id[kSplitRounds], m[kSplitRounds];
#pragma unroll
    for (uint32_t r = 0; r < kSplitRounds; ++r) {
        const uint32_t e = r * 64 + lane;
        id[r] = e < k.n ? k.ids[e] : kNone;
        m[r] = e < k.n ? f.split_mask[(size_t)c * KZ_SPLIT_CHUNK + e] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < kSplitRounds; ++r) {
        if (r * 64 >= k.n) break;
#pragma unroll
        for (uint32_t s = 0; s < NSUB; ++s) {
            const bool in = (m[r] >> s) & 1u;
            const uint64_t bal = __ballot(in);
            if (in) {
                const uint32_t pos = base[s] + rank_in(bal);
                if (pos < f.split_capacity) f.split_lists[pos] = id[r];
            }
            base[s] += (uint32_t)__popcll(bal);
        }
    }
}

// Writes the sub-tile lists: entries keep their list order, so each
// sub-tile list is in primitive order.
__global__ __launch_bounds__(kSplitThreads) void kz_split_scatter(KzFrame f) {
    const uint32_t lane = lane_id();
    const uint32_t nch = f.ctrl[KZ_CTRL_CHUNKS];
    const uint32_t waves = gridDim.x * (kSplitThreads / 64);
    for (uint32_t c = blockIdx.x * (kSplitThreads / 64) + (threadIdx.x >> 6); c < nch; c += waves) {
        const Chunk k = chunk_at(f, c);
        if (k.level == 2) split_scatter_chunk<16>(f, k, c, lane);
        else split_scatter_chunk<4>(f, k, c, lane);
    }
}

// --- learned cost model -------------------------------------------------

// The 16 features of a tile, from its binned-triangle count, its pair
// statistics and the pipeline (the model normalises them):
//   0     log2(1 + binned triangles)
//   1     log2(1 + summed box-tile overlap pixels)
//   2..6  overlap-area histogram (<= 2, 8, 32, 128, more pixels), as
//         fractions of the pairs
//   7     fraction of pairs whose triangle spans several tiles
//   8     z range of the tile's triangles (max z - min z, [0, 1])
//   9..14 pipeline: log2(samples) / 2, blend, depth test, varyings / 16,
//         textures / 4, uniform buffers / 4
//   15    fraction of the tile inside the framebuffer
__device__ __forceinline__ void tile_features(const KzFrame &f, uint32_t t, float *o) {
    const uint32_t cnt = f.tile_counts[t];
    const float n = (float)cnt;
    const uint32_t *a = f.tile_acc + (size_t)t * KZ_ACC;
    const float inv = cnt ? 1.0f / n : 0.0f;
    const uint32_t ts = 1u << f.tile_shift, p = f.pipe_info;
    const uint32_t X0 = (t % f.tiles_x) * ts, Y0 = (t / f.tiles_x) * ts;
    const float inside = (float)(min(ts, f.width - X0) * min(ts, f.height - Y0)) / (float)(ts * ts);
    o[0] = log2f(1.0f + n);
    o[1] = log2f(1.0f + (float)a[KZ_ACC_AREA]);
    for (uint32_t b = 0; b < 5; ++b) o[2 + b] = (float)a[KZ_ACC_HIST + b] * inv;
    o[7] = (float)a[KZ_ACC_SHARED] * inv;
    const float zmin = (float)(65535u - a[KZ_ACC_ZMIN]), zmax = (float)a[KZ_ACC_ZMAX];
    o[8] = cnt ? fmaxf(zmax - zmin, 0.0f) * (1.0f / 65535.0f) : 0.0f;
    o[9] = 0.5f * (float)KZ_PIPE_LOG2_SAMPLES(p);
    o[10] = (float)KZ_PIPE_BLEND(p);
    o[11] = (float)KZ_PIPE_DEPTH(p);
    o[12] = (float)KZ_PIPE_VARYINGS(p) * (1.0f / 16.0f);
    o[13] = (float)KZ_PIPE_TEXTURES(p) * 0.25f;
    o[14] = (float)KZ_PIPE_UBOS(p) * 0.25f;
    o[15] = inside;
}

__global__ void kz_tile_features(KzFrame f, uint32_t ntiles) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    float o[KZ_FEATURES];
    tile_features(f, t, o);
    float *out = f.tile_feat + (size_t)t * KZ_FEATURES;
#pragma unroll
    for (uint32_t k = 0; k < KZ_FEATURES; ++k) out[k] = o[k];
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8_bits __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// f32 -> bf16 bits, round to nearest even (inputs are finite)
__device__ __forceinline__ short to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

constexpr uint32_t kMlpThreads = 256;

// One 64-unit layer for a wave's 16 rows (tiles or samples): in (bf16
// [row][K], row stride in_stride) x W (bf16 [64][K], transposed) on MFMA
// with fp32 accumulation, + bias, ReLU, to bf16 out ([row][64]; may alias
// `in`: every read precedes every write). K a multiple of 32 (the hidden
// layers) runs on gfx950's v_mfma_f32_16x16x32_bf16: lane l supplies
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j < 8 (one
// 16-byte LDS read each); K = 16 (the input layer) on
// v_mfma_f32_16x16x16bf16_1k: A[l % 16][4 (l / 16) + j], B[4 (l / 16) + j]
// [l % 16], j < 4. Both give D[4 (l / 16) + i][l % 16] to lane l.
template <uint32_t K>
__device__ __forceinline__ void mlp_layer(const short *in, uint32_t in_stride, short *out, const short *w,
                                          const float *bias, uint32_t lane) {
    const uint32_t row = lane & 15, quad = lane >> 4;
    f32x4 d[4];
#pragma unroll
    for (uint32_t nb = 0; nb < 4; ++nb) {
        d[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (K % 32 == 0) {
#pragma unroll
            for (uint32_t kb = 0; kb < K / 32; ++kb) {
                const bf16x8_bits a = *(const bf16x8_bits *)&in[row * in_stride + kb * 32 + quad * 8];
                const bf16x8_bits b = *(const bf16x8_bits *)&w[(nb * 16 + row) * K + kb * 32 + quad * 8];
                d[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                               __builtin_bit_cast(bf16x8, b), d[nb], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (uint32_t kb = 0; kb < K / 16; ++kb) {
                bf16x4 a, b;
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j) {
                    a[j] = in[row * in_stride + kb * 16 + quad * 4 + j];
                    b[j] = w[(nb * 16 + row) * K + kb * 16 + quad * 4 + j];
                }
                d[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d[nb], 0, 0, 0);
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (uint32_t nb = 0; nb < 4; ++nb)
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            const uint32_t unit = nb * 16 + row;
            out[(quad * 4 + i) * KZ_HIDDEN + unit] = to_bf16(fmaxf(d[nb][i] + bias[unit], 0.0f));
        }
    __builtin_amdgcn_wave_barrier();
}

// Predicted fragments per tile: pred[t] = max(exp2(MLP(features)) - 1, 0).
// Each wave evaluates 16 tiles at a time; weights are staged in LDS once per
// workgroup.
// The model's weights in LDS (bf16 as stored, fp32 biases and output row)
// and each wave's activations, for a workgroup of THREADS threads; 16-byte
// aligned: the hidden layers read 8 bf16 at a time.
template <uint32_t THREADS>
struct MlpShared {
    __attribute__((aligned(16))) short w1[KZ_HIDDEN * KZ_FEATURES], w2[KZ_HIDDEN * KZ_HIDDEN], w3[KZ_HIDDEN * KZ_HIDDEN];
    float b[3][KZ_HIDDEN], w4[KZ_HIDDEN];
    __attribute__((aligned(16))) short act[THREADS / 64][16 * KZ_HIDDEN];
};

template <uint32_t THREADS>
__device__ __forceinline__ void mlp_stage(const KzCostModel *m, MlpShared<THREADS> &sh) {
    for (uint32_t i = threadIdx.x; i < KZ_HIDDEN * KZ_FEATURES; i += THREADS) sh.w1[i] = (short)m->w1[i];
    for (uint32_t i = threadIdx.x; i < KZ_HIDDEN * KZ_HIDDEN; i += THREADS) {
        sh.w2[i] = (short)m->w2[i];
        sh.w3[i] = (short)m->w3[i];
    }
    if (threadIdx.x < KZ_HIDDEN) {
        sh.b[0][threadIdx.x] = m->b1[threadIdx.x];
        sh.b[1][threadIdx.x] = m->b2[threadIdx.x];
        sh.b[2][threadIdx.x] = m->b3[threadIdx.x];
        sh.w4[threadIdx.x] = m->w4[threadIdx.x];
    }
    __syncthreads();
}

// The prediction for tile `row` of a wave's 16 (lane: row = lane % 16, its
// normalised input features quad * 4 .. + 3 in x, quad = lane / 16); every
// lane of the wave takes part, lanes row + 16 q return the same value.
template <uint32_t THREADS>
__device__ __forceinline__ float mlp_rows(const KzCostModel *m, MlpShared<THREADS> &sh, const float *x, uint32_t lane) {
    const uint32_t row = lane & 15, quad = lane >> 4;
    short *act = sh.act[threadIdx.x >> 6];
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) act[row * KZ_HIDDEN + quad * 4 + j] = to_bf16(x[j]);
    __builtin_amdgcn_wave_barrier();
    mlp_layer<KZ_FEATURES>(act, KZ_HIDDEN, act, sh.w1, sh.b[0], lane);
    mlp_layer<KZ_HIDDEN>(act, KZ_HIDDEN, act, sh.w2, sh.b[1], lane);
    mlp_layer<KZ_HIDDEN>(act, KZ_HIDDEN, act, sh.w3, sh.b[2], lane);
    // output unit: lanes row, row + 16, row + 32, row + 48 share tile `row`
    float y = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
        const uint32_t u = quad * 16 + j;
        y += __uint_as_float((uint32_t)(unsigned short)act[row * KZ_HIDDEN + u] << 16) * sh.w4[u];
    }
    y += __shfl_xor(y, 16, 64);
    y += __shfl_xor(y, 32, 64);
    __builtin_amdgcn_wave_barrier();
    return fmaxf(exp2f(y + m->b4) - 1.0f, 0.0f);
}

__global__ __launch_bounds__(kMlpThreads) void kz_cost_mlp(const KzCostModel *m, const float *feat, uint32_t ntiles,
                                                           float *pred) {
    __shared__ MlpShared<kMlpThreads> sh;
    mlp_stage(m, sh);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t row = lane & 15, quad = lane >> 4;
    const uint32_t stride = gridDim.x * (kMlpThreads / 64) * 16;
    for (uint32_t base = (blockIdx.x * (kMlpThreads / 64) + wave) * 16; base < ntiles; base += stride) {
        // normalised features of tile base + row, 4 per lane, as layer-1 input
        const uint32_t t = base + row;
        float x[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t k = quad * 4 + j;
            x[j] = t < ntiles ? (feat[(size_t)t * KZ_FEATURES + k] - m->feat_mean[k]) * m->feat_scale[k] : 0.0f;
        }
        const float p = mlp_rows(m, sh, x, lane);
        if (quad == 0 && t < ntiles) pred[t] = p;
    }
}

// The balanced form's cost pass with a model, in the plan's tile partition
// (workgroup w: tiles [w * kPlanTiles, ...)): each tile's features (also
// kept in tile_feat for the tile profile), the model's prediction
// (tile_pred) and the workgroup's part of the frame's cost for
// kz_plan_count / kz_plan_queue, summed in a fixed order. Replaces
// kz_tile_features, kz_cost_mlp and kz_plan_cost. A workgroup takes one
// part's tiles, a wave 16 of them.
constexpr uint32_t kCostPlanThreads = kPartTiles / 16 * 64;
__global__ __launch_bounds__(kCostPlanThreads) void kz_cost_plan(const KzCostModel *m, KzFrame f, uint32_t ntiles,
                                                                 uint32_t *table) {
    __shared__ MlpShared<kCostPlanThreads> sh;
    __shared__ float s_part[kCostPlanThreads / 64];
    mlp_stage(m, sh);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t row = lane & 15, quad = lane >> 4;
    float part = 0.0f;  // (lanes of quad 0: their tile's cost)
    {
        const uint32_t t = blockIdx.x * kPartTiles + wave * 16 + row;
        float o[KZ_FEATURES];
        if (t < ntiles) tile_features(f, t, o);
        else
            for (uint32_t k = 0; k < KZ_FEATURES; ++k) o[k] = 0.0f;
        float x[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            // o[quad * 4 + j] without indexing a register array by a lane value
            float v = o[j];
            v = quad == 1 ? o[4 + j] : v;
            v = quad == 2 ? o[8 + j] : v;
            v = quad == 3 ? o[12 + j] : v;
            const uint32_t k = quad * 4 + j;
            x[j] = t < ntiles ? (v - m->feat_mean[k]) * m->feat_scale[k] : 0.0f;
            if (t < ntiles) f.tile_feat[(size_t)t * KZ_FEATURES + k] = v;
        }
        const float p = mlp_rows(m, sh, x, lane);
        if (quad == 0 && t < ntiles) {
            f.tile_pred[t] = p;
            uint32_t off, n;
            kz_tile_list(f, t, off, n);
            if (n) part += tile_cost(f, n, p);
        }
    }
    for (uint32_t d = 32; d > 0; d >>= 1) part += __shfl_xor(part, d, 64);
    if (lane == 0) s_part[wave] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (uint32_t w = 0; w < kCostPlanThreads / 64; ++w) s += s_part[w];
        plan_parts(table, (ntiles + kPlanTiles - 1) / kPlanTiles)[blockIdx.x] = s;
    }
}

// --- cost model training (one workgroup, MFMA forward and backward) ------

constexpr uint32_t kTrainThreads = 256;

__device__ __forceinline__ float from_bf16(short h) { return __uint_as_float((uint32_t)(unsigned short)h << 16); }

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// g[mb][nb] += dZ^T x H over the wave's 16 rows: dW[u][i] = sum_s dz[s][u]
// h[s][i] for 64 output units u and KIN inputs i. Block (mb, nb) lane l
// holds dW[mb * 16 + 4 (l / 16) + r][nb * 16 + l % 16], r = 0..3.
template <uint32_t KIN>
__device__ __forceinline__ void accum_wgrad(const short *dz, const short *h, f32x4 *g, uint32_t lane) {
    const uint32_t row = lane & 15, quad = lane >> 4;
#pragma unroll
    for (uint32_t mb = 0; mb < 4; ++mb)
#pragma unroll
        for (uint32_t nb = 0; nb < KIN / 16; ++nb) {
            bf16x4 a, b;
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                a[j] = dz[(quad * 4 + j) * KZ_HIDDEN + mb * 16 + row];
                b[j] = h[(quad * 4 + j) * KIN + nb * 16 + row];
            }
            g[mb * (KIN / 16) + nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, g[mb * (KIN / 16) + nb], 0, 0, 0);
        }
}

// dH = dZ x W for the wave's 16 rows (W bf16 [64][64]): dh[nb] lane l holds
// dH[4 (l / 16) + r][nb * 16 + l % 16].
__device__ __forceinline__ void back_act(const short *dz, const short *w, f32x4 *dh, uint32_t lane) {
    const uint32_t row = lane & 15, quad = lane >> 4;
#pragma unroll
    for (uint32_t nb = 0; nb < 4; ++nb) {
        dh[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (uint32_t kb = 0; kb < 4; ++kb) {
            bf16x4 a, b;
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                a[j] = dz[row * KZ_HIDDEN + kb * 16 + quad * 4 + j];
                b[j] = w[(kb * 16 + quad * 4 + j) * KZ_HIDDEN + nb * 16 + row];
            }
            dh[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, dh[nb], 0, 0, 0);
        }
    }
}

// dZ = dH * relu'(H) in the D layout of back_act, to bf16 dz [row][64];
// adds the bias gradient partials (unit nb * 16 + l % 16) to gb[nb].
__device__ __forceinline__ void relu_back(const f32x4 *dh, const short *h, short *dz, float *gb, uint32_t lane) {
    const uint32_t row = lane & 15, quad = lane >> 4;
#pragma unroll
    for (uint32_t nb = 0; nb < 4; ++nb)
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
            const uint32_t s = quad * 4 + r, u = nb * 16 + row;
            const float d = h[s * KZ_HIDDEN + u] > 0 ? dh[nb][r] : 0.0f;
            gb[nb] += d;
            dz[s * KZ_HIDDEN + u] = to_bf16(d);
        }
}

// Adds a weight-gradient register block set to the LDS gradient sums.
template <uint32_t KIN>
__device__ __forceinline__ void flush_wgrad(const f32x4 *g, float *s_g, uint32_t lane) {
    const uint32_t row = lane & 15, quad = lane >> 4;
#pragma unroll
    for (uint32_t mb = 0; mb < 4; ++mb)
#pragma unroll
        for (uint32_t nb = 0; nb < KIN / 16; ++nb)
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r)
                atomicAdd(&s_g[(mb * 16 + quad * 4 + r) * KIN + nb * 16 + row], g[mb * (KIN / 16) + nb][r]);
}

// Minibatch Adam on the squared error of the MLP's output against y, the
// forward pass in the same bf16/fp32 arithmetic as kz_cost_mlp. state holds
// [W | M | V] (KZ_PARAMS each, layout KZ_P_*); x: normalised features
// [n][16]; loss_out[step] = the step's mean squared error.
__global__ __launch_bounds__(kTrainThreads) void kz_cost_train(float *state, const float *x, const float *y,
                                                               KzTrainParams p, float *loss_out) {
    float *W = state, *M = state + KZ_PARAMS, *V = state + 2 * KZ_PARAMS;
    // (16-byte aligned: the hidden layers read 8 bf16 at a time)
    __shared__ __attribute__((aligned(16))) short s_w1[KZ_HIDDEN * KZ_FEATURES], s_w2[KZ_HIDDEN * KZ_HIDDEN],
        s_w3[KZ_HIDDEN * KZ_HIDDEN];
    __shared__ float s_b[3][KZ_HIDDEN], s_w4[KZ_HIDDEN], s_b4, s_loss;
    __shared__ float s_g[KZ_PARAMS];
    __shared__ short s_h0[kTrainThreads / 64][16 * KZ_FEATURES];
    __shared__ __attribute__((aligned(16))) short s_h[kTrainThreads / 64][3][16 * KZ_HIDDEN];
    __shared__ short s_dz[kTrainThreads / 64][16 * KZ_HIDDEN];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t row = lane & 15, quad = lane >> 4;
    for (uint32_t i = tid; i < KZ_PARAMS; i += kTrainThreads) {
        const float w = W[i];
        if (i < KZ_P_W2) s_w1[i - KZ_P_W1] = to_bf16(w);
        else if (i < KZ_P_W3) s_w2[i - KZ_P_W2] = to_bf16(w);
        else if (i < KZ_P_B1) s_w3[i - KZ_P_W3] = to_bf16(w);
        else if (i < KZ_P_W4) s_b[(i - KZ_P_B1) / KZ_HIDDEN][(i - KZ_P_B1) % KZ_HIDDEN] = w;
        else if (i < KZ_P_B4) s_w4[i - KZ_P_W4] = w;
        else s_b4 = w;
        s_g[i] = 0.0f;
    }
    if (tid == 0) s_loss = 0.0f;
    __syncthreads();
    short *h0 = s_h0[wave], *h1 = s_h[wave][0], *h2 = s_h[wave][1], *h3 = s_h[wave][2], *dz = s_dz[wave];

    for (uint32_t step = 0; step < p.steps; ++step) {
        f32x4 g3[16], g2[16], g1[4];
        for (uint32_t i = 0; i < 16; ++i) g3[i] = g2[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t i = 0; i < 4; ++i) g1[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        float gb3[16], gw4[16], gb2[4] = {0, 0, 0, 0}, gb1[4] = {0, 0, 0, 0}, gb4 = 0.0f, loss = 0.0f;
        for (uint32_t j = 0; j < 16; ++j) gb3[j] = gw4[j] = 0.0f;

        for (uint32_t tile = 0; tile < p.batch / 64; ++tile) {  // 16 samples per wave per tile
            // this lane's sample (row), drawn with replacement
            const uint64_t k = ((uint64_t)(p.step0 + step) * p.batch) + (tile * 4 + wave) * 16 + row;
            const uint32_t idx = (uint32_t)(splitmix64(p.seed ^ (k * 0x632BE59BD9B4E019ull)) % p.n);
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j)
                h0[row * KZ_FEATURES + quad * 4 + j] = to_bf16(x[(size_t)idx * KZ_FEATURES + quad * 4 + j]);
            const float target = y[idx];
            __builtin_amdgcn_wave_barrier();
            mlp_layer<KZ_FEATURES>(h0, KZ_FEATURES, h1, s_w1, s_b[0], lane);
            mlp_layer<KZ_HIDDEN>(h1, KZ_HIDDEN, h2, s_w2, s_b[1], lane);
            mlp_layer<KZ_HIDDEN>(h2, KZ_HIDDEN, h3, s_w3, s_b[2], lane);
            // output, error and dL/dy for sample `row` (lanes row + 16 q share it)
            float part = 0.0f;
#pragma unroll
            for (uint32_t j = 0; j < 16; ++j) part += from_bf16(h3[row * KZ_HIDDEN + quad * 16 + j]) * s_w4[quad * 16 + j];
            part += __shfl_xor(part, 16, 64);
            part += __shfl_xor(part, 32, 64);
            const float err = part + s_b4 - target;
            const float dy = 2.0f * err / (float)p.batch;
            if (quad == 0) {
                loss += err * err;
                gb4 += dy;
            }
            // layer 4 and dZ3 (lane: sample row, units quad * 16 + j)
#pragma unroll
            for (uint32_t j = 0; j < 16; ++j) {
                const uint32_t u = quad * 16 + j;
                const float h = from_bf16(h3[row * KZ_HIDDEN + u]);
                gw4[j] += dy * h;
                const float d = h > 0.0f ? dy * s_w4[u] : 0.0f;
                gb3[j] += d;
                dz[row * KZ_HIDDEN + u] = to_bf16(d);
            }
            __builtin_amdgcn_wave_barrier();
            f32x4 dh[4];
            accum_wgrad<KZ_HIDDEN>(dz, h2, g3, lane);
            back_act(dz, s_w3, dh, lane);
            __builtin_amdgcn_wave_barrier();
            relu_back(dh, h2, dz, gb2, lane);
            __builtin_amdgcn_wave_barrier();
            accum_wgrad<KZ_HIDDEN>(dz, h1, g2, lane);
            back_act(dz, s_w2, dh, lane);
            __builtin_amdgcn_wave_barrier();
            relu_back(dh, h1, dz, gb1, lane);
            __builtin_amdgcn_wave_barrier();
            accum_wgrad<KZ_FEATURES>(dz, h0, g1, lane);
            __builtin_amdgcn_wave_barrier();
        }

        // the wave's gradient sums into LDS
        flush_wgrad<KZ_HIDDEN>(g3, s_g + KZ_P_W3, lane);
        flush_wgrad<KZ_HIDDEN>(g2, s_g + KZ_P_W2, lane);
        flush_wgrad<KZ_FEATURES>(g1, s_g + KZ_P_W1, lane);
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j) {
            float a = gb3[j], b = gw4[j];
            for (uint32_t d = 1; d < 16; d <<= 1) {
                a += __shfl_xor(a, d, 64);
                b += __shfl_xor(b, d, 64);
            }
            if (row == 0) {
                atomicAdd(&s_g[KZ_P_B3 + quad * 16 + j], a);
                atomicAdd(&s_g[KZ_P_W4 + quad * 16 + j], b);
            }
        }
#pragma unroll
        for (uint32_t nb = 0; nb < 4; ++nb) {
            float a = gb2[nb], b = gb1[nb];
            a += __shfl_xor(a, 16, 64);
            a += __shfl_xor(a, 32, 64);
            b += __shfl_xor(b, 16, 64);
            b += __shfl_xor(b, 32, 64);
            if (quad == 0) {
                atomicAdd(&s_g[KZ_P_B2 + nb * 16 + row], a);
                atomicAdd(&s_g[KZ_P_B1 + nb * 16 + row], b);
            }
        }
        for (uint32_t d = 1; d < 64; d <<= 1) {
            gb4 += __shfl_xor(gb4, d, 64);
            loss += __shfl_xor(loss, d, 64);
        }
        if (lane == 0) {
            atomicAdd(&s_g[KZ_P_B4], gb4);
            atomicAdd(&s_loss, loss);
        }
        __syncthreads();

        // Adam, then the bf16 / fp32 copies the next step computes with
        const float t = (float)(p.step0 + step + 1);
        const float c1 = 1.0f / (1.0f - powf(p.beta1, t)), c2 = 1.0f / (1.0f - powf(p.beta2, t));
        for (uint32_t i = tid; i < KZ_PARAMS; i += kTrainThreads) {
            const float g = s_g[i];
            const float m = p.beta1 * M[i] + (1.0f - p.beta1) * g;
            const float v = p.beta2 * V[i] + (1.0f - p.beta2) * g * g;
            const float w = W[i] - p.lr * (m * c1) / (sqrtf(v * c2) + p.eps);
            M[i] = m;
            V[i] = v;
            W[i] = w;
            s_g[i] = 0.0f;
            if (i < KZ_P_W2) s_w1[i - KZ_P_W1] = to_bf16(w);
            else if (i < KZ_P_W3) s_w2[i - KZ_P_W2] = to_bf16(w);
            else if (i < KZ_P_B1) s_w3[i - KZ_P_W3] = to_bf16(w);
            else if (i < KZ_P_W4) s_b[(i - KZ_P_B1) / KZ_HIDDEN][(i - KZ_P_B1) % KZ_HIDDEN] = w;
            else if (i < KZ_P_B4) s_w4[i - KZ_P_W4] = w;
            else s_b4 = w;
        }
        if (tid == 0) {
            if (loss_out) loss_out[step] = s_loss / (float)p.batch;
            s_loss = 0.0f;
        }
        __syncthreads();
    }
}

// --- training-data generator ----------------------------------------------

// Output k of the seeded SplitMix64 stream `seed` (state seed + (k + 1) *
// gamma), as a float in [0, 1): (u >> 40) * 2^-24.
__device__ __forceinline__ double stream_f32(uint64_t seed, uint64_t k) {
    const uint64_t u = splitmix64(seed + k * 0x9E3779B97F4A7C15ull);  // splitmix64 adds one more gamma
    return (double)(float)((double)(u >> 40) * (1.0 / 16777216.0));
}

// The Gaussian-cluster scene (triangles 0 .. nc-1 around the centre with
// sigma = sigma_frac * (W, H), the rest uniform; vertices at offsets
// U[-radius, radius]^2 px; z and RGBA uniform), as pos4 + colour4 vertices,
// in double precision like the host generator of the same seed.
__global__ void kz_gen_cluster(float *v, uint32_t n, uint32_t nc, uint32_t width, uint32_t height, double sigma_frac,
                               double radius, uint64_t seed) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double W = (double)width, H = (double)height;
    const double u0 = stream_f32(seed, 4ull * t), u1 = stream_f32(seed, 4ull * t + 1);
    const double u2 = stream_f32(seed, 4ull * t + 2), u3 = stream_f32(seed, 4ull * t + 3);
    double cx, cy;
    if (t < nc) {
        const double r = sqrt(-2.0 * log(fmax(u0, 1e-12)));
        const double a = 2.0 * 3.141592653589793 * u1;
        cx = W / 2 + sigma_frac * W * r * cos(a);
        cy = H / 2 + sigma_frac * H * r * sin(a);
    } else {
        cx = u2 * W;
        cy = u3 * H;
    }
    for (uint32_t j = 0; j < 3; ++j) {
        const uint64_t vi = 3ull * t + j;
        const double wx = stream_f32(seed + 3, 6ull * t + 2 * j) * 2 - 1;
        const double wy = stream_f32(seed + 3, 6ull * t + 2 * j + 1) * 2 - 1;
        float *o = v + vi * 8;
        o[0] = (float)((cx + wx * radius) * (2.0 / W) - 1.0);
        o[1] = (float)((cy + wy * radius) * (2.0 / H) - 1.0);
        o[2] = (float)stream_f32(seed + 5, vi);
        o[3] = 1.0f;
        for (uint32_t c = 0; c < 4; ++c) o[4 + c] = (float)stream_f32(seed + 6, 4 * vi + c);
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

// Sums the KZ_STAT_STRIPES partial counters into the totals that follow them:
// thread t reads stripe t's KZ_STAT_COUNT counters (one load each, all in
// flight at once), the waves sum them by butterfly and the first
// KZ_STAT_COUNT threads add up the waves.
constexpr uint32_t kStatsThreads = KZ_STAT_STRIPES;
static_assert(kStatsThreads % 64 == 0 && kStatsThreads <= 1024, "one stripe per thread");
__global__ __launch_bounds__(kStatsThreads) void kz_stats_reduce(unsigned long long *stats) {
    __shared__ unsigned long long s_part[kStatsThreads / 64][KZ_STAT_COUNT];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    unsigned long long v[KZ_STAT_COUNT];
#pragma unroll
    for (uint32_t c = 0; c < KZ_STAT_COUNT; ++c) v[c] = stats[(size_t)tid * KZ_STAT_COUNT + c];
#pragma unroll
    for (uint32_t c = 0; c < KZ_STAT_COUNT; ++c) {
        for (uint32_t d = 32; d > 0; d >>= 1) v[c] += __shfl_xor(v[c], (int)d, 64);
        if (lane == 0) s_part[wave][c] = v[c];
    }
    __syncthreads();
    if (tid < KZ_STAT_COUNT) {
        unsigned long long s = 0;
        for (uint32_t w = 0; w < kStatsThreads / 64; ++w) s += s_part[w][tid];
        stats[(size_t)KZ_STAT_STRIPES * KZ_STAT_COUNT + tid] = s;
    }
}

uint32_t grid_for(uint64_t n, uint32_t block) {
    uint64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    return (uint32_t)(g == 0 ? 1 : g);
}

// Triangles per binning workgroup: enough workgroups to fill the device
// twice over, at most 4096 triangles each (LDS table occupancy), at least
// one full round of the workgroup.
uint32_t bin_chunk(uint32_t ntris) {
    uint32_t c = (ntris / 1024 + kBinThreads - 1) / kBinThreads * kBinThreads;
    if (c < kBinThreads) c = kBinThreads;
    if (c > 4096) c = 4096;
    return c;
}

}  // namespace

hipError_t kz_launch_fill_u32(hipStream_t st, uint32_t *dst, uint64_t n, uint32_t value) {
    if (n == 0) return hipSuccess;
    kz_fill_u32<<<grid_for(n, 256), 256, 0, st>>>(dst, n, value);
    return hipGetLastError();
}

hipError_t kz_launch_fill_tiles(hipStream_t st, uint32_t *counts, uint32_t *offsets, uint32_t *flags, uint32_t n) {
    if (n == 0) return hipSuccess;
    kz_fill_tiles<<<grid_for(n, 256), 256, 0, st>>>(counts, offsets, flags, n);
    return hipGetLastError();
}

hipError_t kz_launch_fill_rect_u32(hipStream_t st, uint32_t *dst, uint32_t width, int32_t x0, int32_t y0, uint32_t w,
                                   uint32_t h, uint32_t spp, uint32_t value) {
    const uint64_t n = (uint64_t)w * h * spp;
    if (n == 0) return hipSuccess;
    kz_fill_rect_u32<<<grid_for(n, 256), 256, 0, st>>>(dst, width, x0, y0, w, h, spp, value);
    return hipGetLastError();
}

hipError_t kz_launch_scan(hipStream_t st, uint32_t *counts, uint32_t *offsets, uint32_t *cursor, uint32_t n,
                          uint32_t capacity, unsigned long long *stats) {
    kz_scan<<<1, kScanThreads, 0, st>>>(counts, offsets, cursor, n, capacity, stats);
    return hipGetLastError();
}

uint32_t kz_radix_blocks(uint64_t capacity) {
    return (uint32_t)std::max<uint64_t>((capacity + kRadixBlock - 1) / kRadixBlock, 1);
}

uint32_t kz_radix_digits_max() { return kRadixMaxDigits; }

uint32_t kz_radix_passes(uint32_t ntiles) {
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < ntiles) ++bits;
    return (bits + kRadixMaxBits - 1) / kRadixMaxBits;
}

hipError_t kz_launch_bin_pairs(hipStream_t st, const KzFrame &f, const KzRadixScratch &s) {
    if (f.total_tris == 0) return hipSuccess;
    const uint32_t nblocks = (f.total_tris + KZ_BIN_BLOCK - 1) / KZ_BIN_BLOCK;
    const uint32_t nchunks = (f.total_tris + KZ_PAIR_CHUNK - 1) / KZ_PAIR_CHUNK;
    kz_pair_scan<<<1, kScanThreads, 0, st>>>(f.block_pairs, nblocks, s.chunk_offsets, nchunks, f.list_capacity,
                                             s.ctrl, f.stats);
    kz_pair_write<<<nchunks, kPairThreads, 0, st>>>(f.tri_rect, f.total_tris, f.tile_shift, f.tiles_x,
                                                    s.chunk_offsets, s.ctrl, s.keys[0], s.vals[0],
                                                    f.list_mask ? 1u : 0u, s.tile_flag);
    return hipGetLastError();
}

hipError_t kz_launch_bin_radix(hipStream_t st, const KzFrame &f, const KzRadixScratch &s, uint32_t ntiles) {
    if (f.total_tris == 0) return hipSuccess;
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < ntiles) ++bits;
    const uint32_t passes = kz_radix_passes(ntiles);
    const uint32_t d = (bits + passes - 1) / passes;
    RadixAcc acc{f.tri_rect, f.tri_z, f.tile_acc, f.tile_shift, f.tiles_x};
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t in = p & 1u, out = in ^ 1u;
        const uint32_t shift = p * d, pb = std::min(d, bits - shift);
        const bool last = p + 1 == passes;
        kz_radix_hist<<<s.tstride, kRadixThreads, 0, st>>>(s.keys[in], s.ctrl, shift, pb, s.table, s.tstride);
        kz_radix_scan<<<1u << pb, kScanThreads, 0, st>>>(s.table, s.tstride, s.ctrl, s.totals);
        kz_radix_scatter<<<s.tstride, kRadixThreads, 0, st>>>(
            s.keys[in], s.vals[in], last ? nullptr : s.keys[out], last ? nullptr : s.vals[out],
            last ? f.tile_lists : nullptr, last ? (uint16_t *)f.list_mask : nullptr, last ? f.tile_counts : nullptr,
            last ? f.tile_offsets : nullptr, s.ctrl,
            s.table, s.tstride, s.totals,
            shift, pb, last ? acc : RadixAcc{nullptr, nullptr, nullptr, 0, 0});
    }
    return hipGetLastError();
}

hipError_t kz_launch_bin_hist(hipStream_t st, const KzFrame &f) {
    if (f.total_tris == 0) return hipSuccess;
    const uint32_t chunk = bin_chunk(f.total_tris);
    kz_bin_hist<<<(f.total_tris + chunk - 1) / chunk, kBinThreads, 0, st>>>(f.tri_rect, f.total_tris, chunk,
                                                                             f.tile_shift, f.tiles_x, f.tile_counts);
    return hipGetLastError();
}

hipError_t kz_launch_bin_hist_acc(hipStream_t st, const KzFrame &f) {
    if (f.total_tris == 0) return hipSuccess;
    // a few hundred workgroups: 1024..8192 triangles each
    uint32_t chunk = (f.total_tris / 512 + kAccThreads - 1) / kAccThreads * kAccThreads;
    chunk = std::min(std::max(chunk, kAccThreads), 8192u);
    kz_bin_hist_acc<<<(f.total_tris + chunk - 1) / chunk, kAccThreads, 0, st>>>(
        f.tri_rect, f.tri_z, f.total_tris, chunk, f.tile_shift, f.tiles_x, f.tile_counts, f.tile_acc);
    return hipGetLastError();
}

hipError_t kz_launch_tile_features(hipStream_t st, const KzFrame &f, uint32_t ntiles) {
    kz_tile_features<<<(ntiles + 255) / 256, 256, 0, st>>>(f, ntiles);
    return hipGetLastError();
}

hipError_t kz_launch_cost_mlp(hipStream_t st, const KzCostModel *model, const float *feat, uint32_t ntiles,
                              float *pred) {
    if (ntiles == 0) return hipSuccess;
    const uint32_t per_block = (kMlpThreads / 64) * 16;
    uint32_t grid = (ntiles + per_block - 1) / per_block;
    if (grid > 1024) grid = 1024;
    kz_cost_mlp<<<grid, kMlpThreads, 0, st>>>(model, feat, ntiles, pred);
    return hipGetLastError();
}

hipError_t kz_launch_gen_cluster(hipStream_t st, float *vertices, uint32_t n, uint32_t nc, uint32_t width,
                                 uint32_t height, double sigma_frac, double radius, uint64_t seed) {
    if (n == 0) return hipSuccess;
    kz_gen_cluster<<<(n + 255) / 256, 256, 0, st>>>(vertices, n, nc, width, height, sigma_frac, radius, seed);
    return hipGetLastError();
}

hipError_t kz_launch_cost_train(hipStream_t st, float *state, const float *x, const float *y, const KzTrainParams &p,
                                float *loss_out) {
    if (p.n == 0 || p.batch == 0 || p.batch % 64 != 0) return hipErrorInvalidValue;
    kz_cost_train<<<1, kTrainThreads, 0, st>>>(state, x, y, p, loss_out);
    return hipGetLastError();
}

hipError_t kz_launch_bin_scatter(hipStream_t st, const KzFrame &f) {
    if (f.total_tris == 0) return hipSuccess;
    const uint32_t chunk = bin_chunk(f.total_tris);
    kz_bin_scatter<<<(f.total_tris + chunk - 1) / chunk, kBinThreads, 0, st>>>(
        f.tri_rect, f.total_tris, chunk, f.tile_shift, f.tiles_x, f.tile_cursor, f.tile_lists, f.list_capacity);
    return hipGetLastError();
}

hipError_t kz_launch_sort_lists(hipStream_t st, const KzFrame &f, uint32_t ntiles) {
    if (ntiles == 0) return hipSuccess;
    // short lists: one wave each; long lists (few, skewed scenes): a workgroup each
    const uint32_t short_grid = std::min((ntiles + 3) / 4, 16384u);
    kz_sort_short<<<short_grid, 256, 0, st>>>(f.tile_counts, f.tile_offsets, ntiles, f.tile_lists, f.list_capacity);
    const uint32_t long_grid = std::min(ntiles, 1024u);
    kz_sort_lists<<<long_grid, kSortThreads, 0, st>>>(f.tile_counts, f.tile_offsets, ntiles, f.tile_lists,
                                                      f.sort_tmp, f.list_capacity);
    return hipGetLastError();
}

uint32_t kz_plan_table_words(uint32_t ntiles) {
    return ((ntiles + kPlanTiles - 1) / kPlanTiles) * (kPlanCols + kPlanTiles / kPartTiles);
}

hipError_t kz_launch_cost_plan(hipStream_t st, const KzCostModel *model, const KzFrame &f, uint32_t ntiles,
                               uint32_t *table) {
    const uint32_t grid = std::max<uint32_t>((ntiles + kPartTiles - 1) / kPartTiles, 1);
    kz_cost_plan<<<grid, kCostPlanThreads, 0, st>>>(model, f, ntiles, table);
    return hipGetLastError();
}

hipError_t kz_launch_plan(hipStream_t st, const KzFrame &f, uint32_t ntiles, uint32_t *table, bool parts_ready) {
    const uint32_t grid = std::max<uint32_t>((ntiles + kPlanTiles - 1) / kPlanTiles, 1);
    if (!parts_ready && (f.tile_pred || !f.bin_ctrl)) kz_plan_cost<<<grid, kPlanWg, 0, st>>>(f, ntiles, table);
    kz_plan_count<<<grid, kPlanWg, 0, st>>>(f, ntiles, table);
    kz_plan_queue<<<grid, kPlanWg, 0, st>>>(f, ntiles, table);
    return hipGetLastError();
}

hipError_t kz_launch_split(hipStream_t st, const KzFrame &f, uint32_t grid, uint32_t ntiles) {
    kz_split_count<<<grid, kSplitThreads, 0, st>>>(f);
    kz_split_totals<<<grid, kSplitThreads, 0, st>>>(f);
    kz_split_offsets<<<(ntiles + kSplitThreads - 1) / kSplitThreads, kSplitThreads, 0, st>>>(f);
    kz_split_scatter<<<grid, kSplitThreads, 0, st>>>(f);
    return hipGetLastError();
}

hipError_t kz_launch_stats_reduce(hipStream_t st, unsigned long long *stats) {
    kz_stats_reduce<<<1, kStatsThreads, 0, st>>>(stats);
    return hipGetLastError();
}

hipError_t kz_launch_resolve4(hipStream_t st, const uint32_t *src, uint32_t *dst, uint32_t width, int32_t x0,
                              int32_t y0, uint32_t w, uint32_t h) {
    if ((uint64_t)w * h == 0) return hipSuccess;
    kz_resolve4<<<grid_for((uint64_t)w * h, 256), 256, 0, st>>>(src, dst, width, x0, y0, w, h);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// kazan_cost_model_eval_noalloc: kazan_cost_model_eval (tools.cpp) without
// device allocations. The model and the feature and prediction staging live
// in `scratch`, a file-scope device array each GPU holds for the life of
// the process. Its device address is looked up once per device with
// hipGetSymbolAddress (in host code the array's name is only a stand-in);
// every use goes through that address. Calls on one device are serialised
// by a mutex; the array keeps the last call's contents, and nothing in it
// needs to start at zero (the kernel writes every prediction it is asked
// for).

namespace {

constexpr size_t kNoAllocModelBytes = (sizeof(KzCostModel) + 255) / 256 * 256;
constexpr size_t kNoAllocFeatBytes = (size_t)KAZAN_NOALLOC_MAX_ROWS * KZ_FEATURES * sizeof(float);
constexpr size_t kNoAllocPredBytes = (size_t)KAZAN_NOALLOC_MAX_ROWS * sizeof(float);
static_assert(kNoAllocFeatBytes % 256 == 0, "prediction staging stays 256-byte aligned");
static_assert(kNoAllocModelBytes + kNoAllocFeatBytes + kNoAllocPredBytes <= KAZAN_NOALLOC_SCRATCH_BYTES,
              "scratch layout exceeds its stated size");

__device__ __align__(256) static unsigned char scratch[KAZAN_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu[kMaxDevices];
unsigned char *g_noalloc_ptr[kMaxDevices];  // device address of `scratch`, per device (under its mutex)

}  // namespace

extern "C" int kazan_cost_model_eval_noalloc(const void *blob, size_t bytes, const float *features, uint32_t n,
                                             float *out) {
    KzCostModel m;
    if (!blob || bytes != sizeof m || (!features && n) || (!out && n)) return -1;
    std::memcpy(&m, blob, sizeof m);
    if (m.magic != KZ_COST_MODEL_MAGIC || m.version != 1 || m.bytes != sizeof m) return -2;
    if (n == 0) return 0;
    if (n > KAZAN_NOALLOC_MAX_ROWS) return -5;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -4;
    std::lock_guard<std::mutex> lock(g_noalloc_mu[dev]);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -4;
        g_noalloc_ptr[dev] = (unsigned char *)p;
    }
    unsigned char *base = g_noalloc_ptr[dev];
    KzCostModel *dm = (KzCostModel *)base;
    float *df = (float *)(base + kNoAllocModelBytes);
    float *dp = (float *)(base + kNoAllocModelBytes + kNoAllocFeatBytes);
    if (hipMemcpy(dm, &m, sizeof m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(df, features, (size_t)n * KZ_FEATURES * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        kz_launch_cost_mlp(nullptr, dm, df, n, dp) != hipSuccess ||
        hipMemcpy(out, dp, (size_t)n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return -4;
    return 0;
}
