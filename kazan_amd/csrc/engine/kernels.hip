// kernels.hip — the pipeline-independent kernels of the draw path (gfx950).
//
//   kz_fill_u32      attachment clears and counter resets (a fill kernel
//                    rather than a memset, so command buffers can be captured
//                    into hipGraphs)
//   kz_bin_scan      exclusive prefix sum of per-tile triangle counts
//   kz_bin_scatter   writes triangle ids into the per-tile lists
//   kz_bin_sort      sorts each tile's list into primitive order (bitonic in
//                    LDS; lists longer than one LDS load are sorted in
//                    LDS-sized runs and merged with merge-path in HBM)
//   kz_resolve4      4x -> 1x box resolve, (s0+s1+s2+s3+2) >> 2 per channel
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace {

constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortCap = 4096;

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

__global__ __launch_bounds__(kScanThreads) void kz_bin_scan(KzFrame f, uint32_t ntiles) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    __shared__ uint32_t s_carry;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < ntiles; b += kScanThreads) {
        const uint32_t i = b + tid;
        const uint32_t v = i < ntiles ? f.tile_counts[i] : 0u;
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
        if (i < ntiles) {
            const uint32_t excl = carry + wbase + x - v;
            f.tile_offsets[i] = excl;
            f.tile_cursor[i] = excl;
        }
        __syncthreads();
        if (tid == 0) s_carry = carry + total;
        __syncthreads();
    }
    if (tid == 0) {
        atomicAdd(&f.stats[KZ_STAT_LIST_ENTRIES], (unsigned long long)s_carry);
        if (s_carry > f.list_capacity) atomicAdd(&f.stats[KZ_STAT_LIST_OVERFLOW], 1ull);
    }
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Reserves list slots for every lane with a tile this step, one atomic per
// distinct tile per wave; returns this lane's slot (or ~0u).
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *cursor, uint32_t tile, uint32_t lane) {
    uint64_t pending = __ballot(tile != 0xFFFFFFFFu);
    uint32_t my_slot = 0xFFFFFFFFu;
    while (pending) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
        const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)tile, leader);
        const uint64_t peers = __ballot(tile == t) & pending;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&cursor[t], (uint32_t)__popcll(peers));
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
        if ((peers >> lane) & 1ull)
            my_slot = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
        pending &= ~peers;
    }
    return my_slot;
}

__global__ __launch_bounds__(256) void kz_bin_scatter(KzFrame f) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = lane_id();
    uint64_t r = g < f.total_tris ? f.tri_rect[g] : KZ_RECT_EMPTY;
    uint32_t tx0 = 1, tx1 = 0, ty0 = 1, ty1 = 0;
    if (r != KZ_RECT_EMPTY) {
        tx0 = (uint32_t)(r & 0xFFFF);
        ty0 = (uint32_t)((r >> 16) & 0xFFFF);
        tx1 = (uint32_t)((r >> 32) & 0xFFFF);
        ty1 = (uint32_t)(r >> 48);
    }
    const uint32_t ntiles = tx0 <= tx1 ? (tx1 - tx0 + 1) * (ty1 - ty0 + 1) : 0u;
    uint32_t steps = ntiles;
    for (uint32_t d = 32; d > 0; d >>= 1) {
        const uint32_t o = __shfl_xor(steps, d, 64);
        steps = steps > o ? steps : o;
    }
    uint32_t tx = tx0, ty = ty0;
    for (uint32_t k = 0; k < steps; ++k) {
        const bool act = k < ntiles;
        const uint32_t slot = wave_reserve(f.tile_cursor, act ? ty * f.tiles_x + tx : 0xFFFFFFFFu, lane);
        if (act) {
            if (slot < f.list_capacity) f.tile_lists[slot] = g;
            if (++tx > tx1) {
                tx = tx0;
                ++ty;
            }
        }
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

// Sorts src[0, n) in place when it fits one LDS load.
__device__ void sort_run_lds(uint32_t *g, uint32_t n, uint32_t *s) {
    const uint32_t P = next_pow2(n);
    for (uint32_t i = threadIdx.x; i < P; i += kSortThreads) s[i] = i < n ? g[i] : 0xFFFFFFFFu;
    __syncthreads();
    bitonic_lds(s, P);
    for (uint32_t i = threadIdx.x; i < n; i += kSortThreads) g[i] = s[i];
    __syncthreads();
}

// Merges adjacent sorted runs of width w from src into dst (merge path: each
// output element finds its split point by binary search; keys are unique).
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
        uint32_t v;
        if (i < a && (j >= b || A[i] < B[j])) v = A[i];
        else v = B[j];
        dst[o] = v;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kSortThreads) void kz_bin_sort(KzFrame f) {
    __shared__ uint32_t s[kSortCap];
    const uint32_t tile = blockIdx.x;
    const uint32_t n = f.tile_counts[tile];
    const uint32_t off = f.tile_offsets[tile];
    if (n <= 1 || (uint64_t)off + n > (uint64_t)f.list_capacity) return;
    uint32_t *list = f.tile_lists + off;
    if (n <= kSortCap) {
        sort_run_lds(list, n, s);
        return;
    }
    for (uint32_t r = 0; r < n; r += kSortCap) sort_run_lds(list + r, min(kSortCap, n - r), s);
    uint32_t *src = list, *dst = f.sort_tmp + off;
    for (uint32_t w = kSortCap; w < n; w <<= 1) {
        merge_pass(src, dst, n, w);
        uint32_t *t = src;
        src = dst;
        dst = t;
    }
    if (src != list)
        for (uint32_t i = threadIdx.x; i < n; i += kSortThreads) list[i] = src[i];
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

hipError_t kz_launch_bin_scan(hipStream_t st, const KzFrame &f, uint32_t ntiles) {
    kz_bin_scan<<<1, kScanThreads, 0, st>>>(f, ntiles);
    return hipGetLastError();
}

hipError_t kz_launch_bin_scatter(hipStream_t st, const KzFrame &f) {
    if (f.total_tris == 0) return hipSuccess;
    kz_bin_scatter<<<(f.total_tris + 255) / 256, 256, 0, st>>>(f);
    return hipGetLastError();
}

hipError_t kz_launch_stats_reduce(hipStream_t st, unsigned long long *stats) {
    kz_stats_reduce<<<1, 64, 0, st>>>(stats);
    return hipGetLastError();
}

hipError_t kz_launch_bin_sort(hipStream_t st, const KzFrame &f, uint32_t ntiles) {
    kz_bin_sort<<<ntiles, kSortThreads, 0, st>>>(f);
    return hipGetLastError();
}

hipError_t kz_launch_resolve4(hipStream_t st, const uint32_t *src, uint32_t *dst, uint32_t width, int32_t x0,
                              int32_t y0, uint32_t w, uint32_t h) {
    if ((uint64_t)w * h == 0) return hipSuccess;
    kz_resolve4<<<grid_for((uint64_t)w * h, 256), 256, 0, st>>>(src, dst, width, x0, y0, w, h);
    return hipGetLastError();
}
