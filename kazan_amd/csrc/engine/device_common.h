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

// One batch of draws that share a pipeline, dynamic state and descriptors.
struct KzFrame {
    const KzDraw *draws;
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
    uint32_t *tile_offsets;  // [tiles] exclusive prefix of counts
    uint32_t *tile_cursor;   // [tiles] scatter cursors
    uint32_t *tile_lists;    // [list_capacity] triangle ids, per tile, sorted
    uint32_t *sort_tmp;      // [list_capacity] merge scratch for long lists
    uint32_t list_capacity;
    uint32_t pad0;
    uint64_t *tri_rect;      // [total_tris] packed pixel bounding box or ~0
    unsigned long long *stats;  // KZ_STAT_* counters

    // Work items (load-balanced mode). A work item is a square sub-tile of
    // a screen tile: level L splits the tile into 4^L squares of side
    // TS >> L. Item geometry is packed as tile << 6 | L << 4 | sub.
    // In the static mode every tile is one item and item i is tile i.
    uint32_t num_items;
    uint32_t balanced;        // 1: persistent raster over `queue`
    uint32_t *tile_level;     // [tiles] split level
    uint32_t *tile_item_base; // [tiles] first item of the tile
    uint32_t *items;          // [num_items] packed geometry
    uint32_t *item_counts;    // [num_items]
    uint32_t *item_offsets;   // [num_items]
    uint32_t *item_cursor;    // [num_items]
    uint32_t *queue;          // [num_items] item ids, most expensive first
    uint32_t *queue_head;     // persistent raster work counter
    float *tile_cost;         // [tiles] predicted cost (triangle-equivalents)
    uint32_t *tile_frags;     // [tiles] measured fragments (training labels) or null
};

enum {
    KZ_STAT_TRIANGLES_IN = 0,
    KZ_STAT_TRIANGLES_SETUP = 1,
    KZ_STAT_TRIANGLES_UNSUPPORTED = 2,
    KZ_STAT_FRAGMENTS = 3,
    KZ_STAT_DEPTH_PASSED = 4,
    KZ_STAT_LIST_ENTRIES = 5,
    KZ_STAT_LIST_OVERFLOW = 6,
    KZ_STAT_COUNT = 8
};

// Statistics are accumulated in KZ_STAT_STRIPES copies (a workgroup adds to
// copy blockIdx % KZ_STAT_STRIPES) so that no single address takes every
// workgroup's atomics; kz_stats_reduce folds them at the end of a submission.
#define KZ_STAT_STRIPES 256

// Per-triangle pixel bounding box (inclusive, clipped to the scissor),
// packed as px0 | py0 << 16 | px1 << 32 | py1 << 48; ~0 = no pixels.
#define KZ_RECT_EMPTY (~0ULL)

#define KZ_ITEM_PACK(tile, level, sub) (((tile) << 6) | ((level) << 4) | (sub))
#define KZ_ITEM_TILE(it) ((it) >> 6)
#define KZ_ITEM_LEVEL(it) (((it) >> 4) & 3u)
#define KZ_ITEM_SUB(it) ((it) & 15u)

// Split thresholds of the balanced mode, in predicted cost units
// (triangle-equivalents of raster work): a tile becomes 4 items above
// KZ_SPLIT_4 and 16 items above KZ_SPLIT_16.
#define KZ_SPLIT_4 1536.0f
#define KZ_SPLIT_16 6144.0f
