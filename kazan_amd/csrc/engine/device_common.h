// device_common.h — data shared by the host engine, the prebuilt binning
// kernels (kernels.hip) and the per-pipeline kernels that pipeline creation
// compiles from raster_template.hip. Plain structs only: they are passed by
// value as kernel arguments and copied into device memory as-is, so the
// layout must be identical in the host compiler and in device code.
#pragma once

#ifndef __HIPCC_RTC__
#include <stdint.h>
#else
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::int64_t int64_t;
#endif

#define KZ_MAX_BINDINGS 8
#define KZ_MAX_TEXTURES 4
#define KZ_MAX_UBOS 4
#define KZ_PUSH_WORDS 32
#define KZ_MAX_VARYINGS 16
#define KZ_MAX_SSBOS 8

// Per-draw record. One record per (vkCmdDraw*, instance).
struct KzDraw {
    const uint8_t *vb[KZ_MAX_BINDINGS];  // binding base = buffer address + bind offset
    const uint8_t *ib;                   // index buffer base (+ bind offset)
    uint32_t index_type;                 // 0 = u16, 1 = u32, 2 = non-indexed
    uint32_t first;                      // firstIndex (indexed) or firstVertex
    int32_t vertex_offset;               // vertexOffset (indexed only)
    uint32_t instance;                   // gl_InstanceIndex of this record
    uint32_t tri_count;                  // triangles in this record
    uint32_t tri_begin;                  // index of its first triangle in the batch
    uint32_t pad[2];
    uint32_t push[KZ_PUSH_WORDS];        // push-constant bytes at draw time
};

// Combined image sampler as the fragment kernel sees it. Images are linear
// RGBA8, row pitch = w * 4 bytes.
struct KzTexture {
    const uint8_t *data;
    uint32_t w, h;
    uint32_t filter;  // VkFilter
    uint32_t addr_u, addr_v;
    uint32_t pad;
};

// The argument of a compute pipeline's kernel (vkCmdDispatch): the bound
// descriptors and push constants at dispatch time, and the dispatch's base
// workgroup (vkCmdDispatchBase).
struct KzComputeFrame {
    KzTexture tex[KZ_MAX_TEXTURES];
    const uint8_t *ubo[KZ_MAX_UBOS];
    uint8_t *ssbo[KZ_MAX_SSBOS];
    uint32_t ssbo_bytes[KZ_MAX_SSBOS];  // the bound range of each storage buffer (OpArrayLength)
    uint32_t push[KZ_PUSH_WORDS];
    uint32_t base[3];
    uint32_t pad;
};

// One batch of draws that share a pipeline, dynamic state and descriptors.
struct KzFrame {
    const KzDraw *draws;
    const uint32_t *draw_begin;  // [ndraws] draws[i].tri_begin, densely
    const uint32_t *block_draw;  // [ceil(total_tris / KZ_BIN_BLOCK) + 1] draw of each block's first triangle, the last draw
    uint32_t ndraws;
    uint32_t total_tris;

    // attachments: colour RGBA8 and depth D32, W*H*samples, samples
    // interleaved per pixel
    uint8_t *color;
    float *depth;
    uint32_t width, height;
    uint32_t tiles_x, tiles_y;
    uint32_t tile_shift;
    uint32_t pad1;

    // viewport transform constants, precomputed on the host in the same
    // order the oracle uses: hw = w*0.5f, ox = x + hw, ...
    float vp_hw, vp_hh, vp_ox, vp_oy, vp_dz, vp_zmin;
    float z_lo, z_hi;  // depth clamp range
    // scissor ∩ framebuffer, inclusive pixel rectangle
    int32_t sc_x0, sc_y0, sc_x1, sc_y1;

    KzTexture tex[KZ_MAX_TEXTURES];
    const uint8_t *ubo[KZ_MAX_UBOS];

    // binning scratch (device memory owned by the engine)
    uint32_t *tile_counts;   // [tiles]
    uint32_t *tile_offsets;  // [tiles] first entry of the tile's list (any value when empty)
    uint32_t *tile_cursor;   // [tiles] scatter cursors
    uint32_t *tile_flag;     // [tiles] nonzero: cut by the list capacity (radix binning), or null
    const uint32_t *bin_ctrl;  // KZ_BIN_* words of radix binning (the first cut chunk), or null
    uint32_t *tile_lists;    // [list_capacity] triangle ids, per tile, sorted
    uint32_t *sort_tmp;      // [list_capacity] merge scratch for long lists
    uint32_t list_capacity;
    uint32_t pad0;
    uint64_t *tri_rect;      // [total_tris] packed pixel bounding box or ~0
    uint32_t *tri_rec;       // [total_tris * KZ_REC_WORDS_FOR(varyings)] setup records, or null
    uint32_t *block_pairs;   // [ceil(total_tris / 256)] (triangle, tile) pairs per binning block, or null
    uint32_t *tri_draw;      // [total_tris] draw of each triangle (batches of several draws), or null
    unsigned long long *stats;  // KZ_STAT_* counters

    // Load-balanced mode. A work item is a square sub-tile of a screen
    // tile: level L splits the tile into 4^L squares of side TS >> L (L = 0
    // is the tile itself). Queue entries pack an item as
    // tile << 6 | L << 4 | sub. Level-0 items rasterise their tile's list;
    // the sub-tiles of a split tile ("split slot" r = split_slot[tile]) get
    // lists of their own, r * 16 + sub, cut from the tile's list in order.
    uint32_t balanced;          // 1: persistent raster over `queue`
    uint32_t split_capacity;    // entries of split_lists
    uint32_t *queue;            // [16 * tiles] items, most expensive first
    uint32_t *ctrl;             // KZ_CTRL_* counters
    uint32_t *tile_level;       // [tiles] split level
    uint32_t *split_slot;       // [tiles] split slot of a split tile
    uint32_t *split_info;       // [2 * split slots] (first chunk, tile) of a slot
    uint32_t *chunk_desc;       // [max_chunks * 4] per KZ_SPLIT_CHUNK-entry chunk of a split tile: the list
                                // position of its first entry, the tile, slot << 2 | level, its entries
    uint32_t *chunk_counts;     // [chunks * 16] entries per sub-tile, then bases within the slot's lists
    uint16_t *split_mask;       // [max_chunks * KZ_SPLIT_CHUNK] sub-tiles each chunk entry overlaps
    const uint16_t *list_mask;  // [list capacity] the 4 x 4 sub-tiles (bit sy * 4 + sx) each list entry's
                                // pixel box overlaps in its tile (radix binning wrote them), or null
    uint32_t *split_counts;     // [split slots * 16]
    uint32_t *split_offsets;    // [split slots * 16]
    uint32_t *split_lists;      // [split_capacity]
    uint32_t max_chunks;
    uint32_t pipe_info;         // KZ_PIPE_* bits of the pipeline (cost-model features)

    // Cost model inputs and outputs (null when unused): per-triangle depth
    // range and per-tile pair statistics gathered while binning, the 16
    // features derived from them, the model's predicted fragments per tile
    // (the plan splits by cost = binned triangles + frag_weight * predicted
    // fragments, in list entries), and, in the static form with tile
    // profiling, each tile's fragments (training labels) and raster cycles.
    uint32_t *tri_z;            // [total_tris] z range, min | max << 16 (16-bit unorm)
    uint32_t *tile_acc;         // [tiles * KZ_ACC]
    float *tile_feat;           // [tiles * KZ_FEATURES]
    float *tile_pred;           // [tiles]
    uint32_t *tile_frags;       // [tiles]
    uint32_t *tile_cycles;      // [tiles]
    float frag_weight;
    uint32_t raster_slots;      // workgroups of the persistent raster (split threshold)
    uint32_t split_max_level;   // 0..2: deepest split the plan may choose
    // The render pass's clears folded into this batch (bit 0: colour, bit
    // 1: depth; recorded just before it, over the whole framebuffer): the
    // raster starts every tile from these values instead of loading it, and
    // writes every tile, also those without triangles.
    uint32_t clear_mask;
    uint32_t clear_color;       // packed RGBA8
    float clear_depth;
};

// Per-tile pair statistics (tile_acc) over the (triangle, tile) pairs; the
// overlap of a pair is the area of the triangle's pixel box inside the tile.
#define KZ_ACC 9
#define KZ_ACC_AREA 0   // sum of overlaps, pixels
#define KZ_ACC_HIST 1   // 5 bins: overlap <= 2, <= 8, <= 32, <= 128, > 128 pixels
#define KZ_ACC_SHARED 6 // pairs whose triangle's box spans more than one tile
#define KZ_ACC_ZMIN 7   // max of (65535 - 16-bit min z)
#define KZ_ACC_ZMAX 8   // max of 16-bit max z

// pipe_info bits
#define KZ_PIPE_LOG2_SAMPLES(p) ((p) & 7u)
#define KZ_PIPE_BLEND(p) (((p) >> 3) & 1u)
#define KZ_PIPE_DEPTH(p) (((p) >> 4) & 1u)
#define KZ_PIPE_VARYINGS(p) (((p) >> 8) & 31u)
#define KZ_PIPE_TEXTURES(p) (((p) >> 16) & 7u)
#define KZ_PIPE_UBOS(p) (((p) >> 20) & 7u)
#define KZ_PIPE_PACK(log2s, blend, depth, nv, ntex, nubo)                                                   \
    ((uint32_t)(log2s) | ((uint32_t)(blend) << 3) | ((uint32_t)(depth) << 4) | ((uint32_t)(nv) << 8) | \
     ((uint32_t)(ntex) << 16) | ((uint32_t)(nubo) << 20))

// Learned per-tile cost model: a bf16 MLP 16 -> 64 -> 64 -> 64 -> 1 with
// fp32 accumulation and ReLU, on features normalised as
// (x - feat_mean) * feat_scale. The output y predicts log2(1 + fragments)
// of the tile. frag_weight converts fragments to list-entry equivalents of
// raster time (fitted from measured tile cycles). Weights are bf16 bit
// patterns, stored [out][in] (the layout the MFMA B operand reads).
#define KZ_FEATURES 16
#define KZ_HIDDEN 64
struct KzCostModel {
    uint32_t magic;    // KZ_COST_MODEL_MAGIC
    uint32_t version;  // 1
    uint32_t bytes;    // sizeof(KzCostModel)
    float frag_weight;
    float feat_mean[KZ_FEATURES];
    float feat_scale[KZ_FEATURES];
    uint16_t w1[KZ_HIDDEN * KZ_FEATURES];
    uint16_t w2[KZ_HIDDEN * KZ_HIDDEN];
    uint16_t w3[KZ_HIDDEN * KZ_HIDDEN];
    float b1[KZ_HIDDEN], b2[KZ_HIDDEN], b3[KZ_HIDDEN];
    float w4[KZ_HIDDEN];
    float b4;
    uint32_t pad[3];
};
#define KZ_COST_MODEL_MAGIC 0x4D435A4Bu  // "KZCM"

// Training: the fp32 parameter vector (master weights; Adam keeps two more
// vectors of the same layout after it) and the hyper-parameters of one run
// of the training kernel.
#define KZ_P_W1 0
#define KZ_P_W2 (KZ_P_W1 + KZ_HIDDEN * KZ_FEATURES)
#define KZ_P_W3 (KZ_P_W2 + KZ_HIDDEN * KZ_HIDDEN)
#define KZ_P_B1 (KZ_P_W3 + KZ_HIDDEN * KZ_HIDDEN)
#define KZ_P_B2 (KZ_P_B1 + KZ_HIDDEN)
#define KZ_P_B3 (KZ_P_B2 + KZ_HIDDEN)
#define KZ_P_W4 (KZ_P_B3 + KZ_HIDDEN)
#define KZ_P_B4 (KZ_P_W4 + KZ_HIDDEN)
#define KZ_PARAMS (KZ_P_B4 + 1)
struct KzTrainParams {
    float lr, beta1, beta2, eps;
    uint32_t steps;   // Adam steps in this run
    uint32_t step0;   // steps taken before (bias correction, sample stream)
    uint32_t batch;   // samples per step, a multiple of 64
    uint32_t n;       // samples in the data set
    uint64_t seed;    // minibatch sampling stream
};

enum {
    KZ_STAT_TRIANGLES_IN = 0,
    KZ_STAT_TRIANGLES_SETUP = 1,
    KZ_STAT_TRIANGLES_CLIPPED = 2,
    KZ_STAT_FRAGMENTS = 3,
    KZ_STAT_DEPTH_PASSED = 4,
    KZ_STAT_LIST_ENTRIES = 5,
    KZ_STAT_LIST_OVERFLOW = 6,
    KZ_STAT_SPLIT_ENTRIES = 7,
    KZ_STAT_COUNT = 8
};

// ctrl words of the balanced mode
enum {
    KZ_CTRL_ITEMS = 1,     // queue length
    KZ_CTRL_SPLITS = 2,    // split slots used
    KZ_CTRL_CHUNKS = 3,    // chunks of split tiles' lists
    KZ_CTRL_SPLIT_MAX = 4, // largest sub-tile list total of any batch so far (sizing)
    KZ_CTRL_SPLIT_CURSOR = 5, // sub-tile list storage reserved so far this batch
    // persistent raster: per sub-queue, the items taken beyond the first
    // round (KZ_QUEUES counters, KZ_CTRL_HEAD_STRIDE words apart: 128 bytes)
    KZ_CTRL_HEADS = 64,
    KZ_CTRL_HEAD_STRIDE = 32,
    KZ_CTRL_WORDS = 64 + 8 * 32
};
// sub-queues of the persistent raster (one per XCD)
#define KZ_QUEUES 8u

// Entries of a split tile's list that one wave cuts into sub-tile lists
// (small enough that a frame's split tiles keep every SIMD busy).
#define KZ_SPLIT_CHUNK 256u

// Radix binning: kz_bin_count adds the (triangle, tile) pairs of each block
// of KZ_BIN_BLOCK triangles; pairs are generated KZ_PAIR_CHUNK triangles
// (KZ_PAIR_CHUNK / KZ_BIN_BLOCK blocks) per workgroup.
#define KZ_BIN_BLOCK 256u
#define KZ_PAIR_CHUNK 4096u
enum {
    KZ_BIN_PAIRS = 0,  // pairs generated and sorted this batch (<= list capacity)
    KZ_BIN_CUT = 1,    // first pair chunk whose pairs did not fit (chunks when all fit)
    KZ_BIN_OVER = 2,   // 1 when chunks were cut
    KZ_BIN_WORDS = 4
};

// Statistics are accumulated in KZ_STAT_STRIPES copies (a workgroup adds to
// copy blockIdx % KZ_STAT_STRIPES) so that no single address takes every
// workgroup's atomics; kz_stats_reduce folds them when the host waits.
#define KZ_STAT_STRIPES 256

// Per-triangle pixel bounding box (inclusive, clipped to the scissor),
// packed as px0 | py0 << 16 | px1 << 32 | py1 << 48; ~0 = no pixels.
#define KZ_RECT_EMPTY (~0ULL)

// 32-bit words of a setup record with nv varyings: 16 words of edge and
// plane constants, 3 per varying (base, two deltas), rounded up to 16 B.
#define KZ_REC_WORDS_FOR(nv) ((16u + 3u * (nv) + 3u) & ~3u)

#define KZ_ITEM_PACK(tile, level, sub) (((tile) << 6) | ((level) << 4) | (sub))
#define KZ_ITEM_TILE(it) ((it) >> 6)
#define KZ_ITEM_LEVEL(it) (((it) >> 4) & 3u)
#define KZ_ITEM_SUB(it) ((it) & 15u)

// Balanced mode: a tile is split when its cost (list entries) exceeds the
// frame's cost per resident raster workgroup, and never below this.
#define KZ_SPLIT_MIN 512.0f

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
// Fragment-shader arithmetic at Vulkan's precision, not IEEE's: a division
// within 2.5 ULP (reciprocal then product) and a square root within 1 ULP
// (v_rcp_f32, v_sqrt_f32), where correctly rounded ones take ~10
// instructions each. Fragment shaders only feed colour (RGBA8); vertex
// shaders, which place the samples and depths, keep the IEEE operations.
__device__ __forceinline__ float kz_fs_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float kz_fs_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Shader integer and float helpers with SPIR-V's definitions where C++
// leaves the result undefined: division by zero gives 0 (SPIR-V leaves it
// undefined; this keeps the GPU from computing garbage slowly), INT_MIN / -1
// wraps, shift counts use their low 5 bits, out-of-range float -> unsigned
// conversions saturate.
__device__ __forceinline__ int32_t kz_sdiv(int32_t a, int32_t b) {
    return b == 0 ? 0 : (b == -1 ? (int32_t)(0u - (uint32_t)a) : a / b);
}
__device__ __forceinline__ int32_t kz_udiv(int32_t a, int32_t b) {
    return b == 0 ? 0 : (int32_t)((uint32_t)a / (uint32_t)b);
}
__device__ __forceinline__ int32_t kz_srem(int32_t a, int32_t b) { return (b == 0 || b == -1) ? 0 : a % b; }
__device__ __forceinline__ int32_t kz_smod(int32_t a, int32_t b) {
    const int32_t r = kz_srem(a, b);
    return (r != 0 && ((r < 0) != (b < 0))) ? r + b : r;
}
__device__ __forceinline__ int32_t kz_umod(int32_t a, int32_t b) {
    return b == 0 ? 0 : (int32_t)((uint32_t)a % (uint32_t)b);
}
__device__ __forceinline__ int32_t kz_shl(int32_t a, int32_t b) { return (int32_t)((uint32_t)a << (b & 31)); }
__device__ __forceinline__ int32_t kz_lshr(int32_t a, int32_t b) { return (int32_t)((uint32_t)a >> (b & 31)); }
__device__ __forceinline__ int32_t kz_ashr(int32_t a, int32_t b) { return a >> (b & 31); }
__device__ __forceinline__ int32_t kz_smin(int32_t a, int32_t b) { return a < b ? a : b; }
__device__ __forceinline__ int32_t kz_smax(int32_t a, int32_t b) { return a > b ? a : b; }
__device__ __forceinline__ int32_t kz_umin(int32_t a, int32_t b) { return (uint32_t)a < (uint32_t)b ? a : b; }
__device__ __forceinline__ int32_t kz_umax(int32_t a, int32_t b) { return (uint32_t)a > (uint32_t)b ? a : b; }
__device__ __forceinline__ int32_t kz_sclamp(int32_t x, int32_t lo, int32_t hi) { return kz_smin(kz_smax(x, lo), hi); }
__device__ __forceinline__ int32_t kz_uclamp(int32_t x, int32_t lo, int32_t hi) { return kz_umin(kz_umax(x, lo), hi); }
__device__ __forceinline__ int32_t kz_sabs(int32_t a) { return a < 0 ? (int32_t)(0u - (uint32_t)a) : a; }
__device__ __forceinline__ float kz_fsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : x); }
__device__ __forceinline__ float kz_fmod_glsl(float x, float y) { return x - y * __builtin_floorf(x / y); }
__device__ __forceinline__ int32_t kz_f32_to_f16_bits(float x) {
    return (int32_t)__builtin_bit_cast(unsigned short, (_Float16)x);
}
__device__ __forceinline__ float kz_f16_bits_to_f32(int32_t v) {
    return (float)__builtin_bit_cast(_Float16, (unsigned short)(uint32_t)v);
}
__device__ __forceinline__ float kz_ldexp(float x, int32_t e) { return __builtin_ldexpf(x, e); }
__device__ __forceinline__ float kz_frexp_mant(float x) {
    int e;
    return __builtin_frexpf(x, &e);
}
__device__ __forceinline__ int32_t kz_frexp_exp(float x) {
    int e;
    __builtin_frexpf(x, &e);
    return e;
}
__device__ __forceinline__ int32_t kz_bit_count(int32_t x) { return __builtin_popcount((uint32_t)x); }
__device__ __forceinline__ int32_t kz_bit_reverse(int32_t x) { return (int32_t)__builtin_bitreverse32((uint32_t)x); }
__device__ __forceinline__ int32_t kz_find_lsb(int32_t x) { return x == 0 ? -1 : __builtin_ctz((uint32_t)x); }
__device__ __forceinline__ int32_t kz_find_umsb(int32_t x) { return x == 0 ? -1 : 31 - __builtin_clz((uint32_t)x); }
__device__ __forceinline__ int32_t kz_find_smsb(int32_t x) { return kz_find_umsb(x < 0 ? ~x : x); }
// Bit fields of offset + count <= 32 bits (SPIR-V leaves the rest undefined;
// here a field is cut at bit 31 and an offset of 32 or more selects nothing).
__device__ __forceinline__ uint32_t kz_field_mask(int32_t count) {
    return (uint32_t)count >= 32u ? 0xFFFFFFFFu : (1u << (uint32_t)count) - 1u;
}
__device__ __forceinline__ int32_t kz_bitfield_insert(int32_t base, int32_t ins, int32_t off, int32_t count) {
    if ((uint32_t)off >= 32u || count == 0) return base;
    const uint32_t m = kz_field_mask(count) << (uint32_t)off;
    return (int32_t)(((uint32_t)base & ~m) | (((uint32_t)ins << (uint32_t)off) & m));
}
__device__ __forceinline__ int32_t kz_bitfield_uextract(int32_t base, int32_t off, int32_t count) {
    if ((uint32_t)off >= 32u || count == 0) return 0;
    return (int32_t)(((uint32_t)base >> (uint32_t)off) & kz_field_mask(count));
}
__device__ __forceinline__ int32_t kz_bitfield_sextract(int32_t base, int32_t off, int32_t count) {
    if ((uint32_t)off >= 32u || count == 0) return 0;
    const uint32_t c = (uint32_t)count >= 32u ? 32u : (uint32_t)count;
    return (int32_t)(((uint32_t)base >> (uint32_t)off) << (32u - c)) >> (32u - c);
}
__device__ __forceinline__ uint32_t kz_f2u(float x) {
    return !(x > 0.0f) ? 0u : (x >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)x);
}
// A shader array's run-time word index, clamped into its n words.
__device__ __forceinline__ int32_t kz_word_index(int32_t i, int32_t n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// Tile t's list: its offset and count. A tile flagged by a cut (binning
// dropped the pairs of triangles past the list capacity, some of them in
// this tile) reads as offset list_capacity with at least one entry: the
// raster then takes its triangles from the in-order scan of all pixel
// boxes, and the plan never splits it.
__device__ __forceinline__ void kz_tile_list(const KzFrame &f, uint32_t t, uint32_t &off, uint32_t &n) {
    off = f.tile_offsets[t];
    n = f.tile_counts[t];
    if (f.tile_flag && f.tile_flag[t]) {
        off = f.list_capacity;
        n = n > 1u ? n : 1u;
    }
}
#endif
