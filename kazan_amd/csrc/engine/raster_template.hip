// raster_template.hip — per-pipeline kernels of the draw path for gfx950.
//
// vkCreateGraphicsPipelines compiles this file (through hiprtc) after a
// generated prelude that defines the pipeline's fixed-function state as
// macros and its shaders as device functions produced by the HIP shader
// backend from the shader IR:
//
//   KZ_TILE_SHIFT, KZ_SAMPLES, KZ_NV (packed varying floats),
//   KZ_CULL_MODE, KZ_FRONT_FACE, KZ_DEPTH_TEST, KZ_DEPTH_WRITE, KZ_DEPTH_OP,
//   KZ_BLEND, KZ_SRC_COLOR, KZ_DST_COLOR, KZ_COLOR_OP, KZ_SRC_ALPHA,
//   KZ_DST_ALPHA, KZ_ALPHA_OP, KZ_WRITE_MASK
//   void kz_vertex(const KzFrame&, const KzDraw&, uint32_t vertex_index,
//                  float pos[4], float vary[KZ_NV]);
//   bool kz_fragment(const KzFrame&, const float vary[KZ_NV], float out[4]);  (false: discarded)
//
// The rasterisation rules restated here are the Vulkan 1.1 polygon rules at
// Kazan's advertised precision (vulkan-driver/src/api_impl.rs:1559-1568,
// :1584, :1609): 1/16-px snapping with round-half-even, int64 edge
// functions, top-left tie-break with y pointing down, standard 4x sample
// positions, f32 barycentrics with a fixed operation order. The code is
// compiled with -ffp-contract=off and IEEE division/square root; the
// fragment shaders' own division and square root take Vulkan's precision
// (kz_fs_div, kz_fs_sqrt).
//
// Kernels:
//   kz_bin_count  one thread per triangle: vertex fetch + position part of
//                 the vertex shader + setup; writes the triangle's pixel box
//                 (the box of the pixels it covers when the box is small,
//                 empty when it covers no sample) and the pair count of its
//                 256-triangle block, from which the prebuilt kernels build
//                 the per-tile lists in primitive order.
//   kz_raster     one wave64 per screen tile: walks the tile's triangle list
//                 in batches of 64, one triangle per lane; coverage is
//                 compacted with ballot + prefix sums into an LDS fragment
//                 queue; depth test and blending are applied in primitive
//                 order per pixel against the tile's colour and depth held
//                 in LDS; the tile is written back once.
//   kz_raster_persistent  the same per work item (a tile or a sub-tile of a
//                 split tile), resident workgroups taking items from the
//                 plan's queue, most expensive first.

#define KZ_TS (1 << KZ_TILE_SHIFT)
#define KZ_TP (KZ_TS * KZ_TS)
#define KZ_FRAG_CAP (KZ_TP * 4)
// A triangle with at most this many candidate pixels in the square (and
// short edges) is walked by its own lane; the others take the whole wave,
// one triangle at a time, which costs a fixed pass over the wave per
// triangle: for the 1-8 px triangles of dense meshes the lane walk is the
// cheaper of the two well past 16 candidates (a 1080p grid of 2 px
// triangles: raster 89 -> 68 us). The candidates' sample masks must fit
// a 64-bit mask: 16 pixels at 4 samples.
#define KZ_SMALL_CAND (KZ_SAMPLES == 1 ? 24 : 16)
// binning tests coverage of triangles whose pixel box has at most this many
// pixels (and bins them by the box of the pixels they cover)
#define KZ_TIGHT_CAND 16
#define KZ_NO_CLAIM 0xFFFFFFFFu

#if KZ_NV > 0
#define KZ_NVA KZ_NV
#else
#define KZ_NVA 1
#endif

// ---------------------------------------------------------------------------
// wave64 helpers

__device__ __forceinline__ uint32_t kz_lane() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint64_t kz_ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t kz_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ int32_t kz_bperm_i(int32_t v, uint32_t src_lane) {
    return __builtin_amdgcn_ds_bpermute((int)(src_lane << 2), v);
}
__device__ __forceinline__ float kz_bperm_f(float v, uint32_t src_lane) {
    return __int_as_float(kz_bperm_i(__float_as_int(v), src_lane));
}
__device__ __forceinline__ int32_t kz_readlane_i(int32_t v, uint32_t lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ float kz_readlane_f(float v, uint32_t lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ uint32_t kz_wave_incl_scan(uint32_t v, uint32_t lane) {
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint32_t o = (uint32_t)kz_bperm_i((int32_t)v, lane >= d ? lane - d : lane);
        if (lane >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ int32_t kz_wave_max(int32_t v, uint32_t lane) {
    for (uint32_t d = 32; d > 0; d >>= 1) {
        int32_t o = kz_bperm_i(v, lane ^ d);
        v = v > o ? v : o;
    }
    return v;
}

__device__ __forceinline__ unsigned long long kz_wave_sum_u64(unsigned long long v, uint32_t lane) {
    for (uint32_t d = 32; d > 0; d >>= 1) {
        uint32_t lo = (uint32_t)kz_bperm_i((int32_t)(uint32_t)v, lane ^ d);
        uint32_t hi = (uint32_t)kz_bperm_i((int32_t)(uint32_t)(v >> 32), lane ^ d);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

// ---------------------------------------------------------------------------
// primitive assembly

// The draw record of batch triangle g: binary search over the draws' first
// triangles (a dense 4-byte array, so the probes of a wave share cache
// lines), between the draws of the first triangles of g's 256-triangle
// block and of the next block.
__device__ __forceinline__ uint32_t kz_find_draw(const KzFrame &f, uint32_t g) {
    uint32_t lo = 0, hi = f.ndraws - 1;
    if (f.block_draw) {
        lo = f.block_draw[g / KZ_BIN_BLOCK];
        hi = f.block_draw[g / KZ_BIN_BLOCK + 1];
    }
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (f.draw_begin[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t kz_vertex_index(const KzDraw &d, uint32_t t, uint32_t j) {
    const uint32_t k = d.first + 3u * t + j;
    if (d.index_type == 2u) return k;
    const uint32_t idx = d.index_type == 1u ? ((const uint32_t *)d.ib)[k] : (uint32_t)((const uint16_t *)d.ib)[k];
    return (uint32_t)((int32_t)idx + d.vertex_offset);
}

// ---------------------------------------------------------------------------
// triangle setup (identical decisions in kz_bin_count and kz_raster)

struct KzTri {
    int32_t x0, y0, d1x, d1y, d2x, d2y;
    int64_t S;
    float invS, z0, dz1, dz2, iw0, iw1, iw2;
    int32_t px0, py0, px1, py1;
};

// 0 = dropped (clip-rejected, degenerate or culled), 1 = set up,
// 2 = needs clipping (kz_clip_polygon; the pieces are set up with `piece`).
// `piece`: p is a piece of a clipped triangle: no clip checks, and a
// vertex with w <= 0 or outside the guard band drops the piece.
// Straight-line: every field of t is written whatever the outcome (the
// checks pick the status at the end), so that a wave whose lanes end in
// different cases takes no branches.
__device__ __forceinline__ int kz_setup_rt(const KzFrame &f, const float p[3][4], KzTri &t, bool piece) {
    // trivially rejected: all three vertices outside one clip plane
    bool reject = false;
#pragma unroll
    for (int plane = 0; plane < 6; ++plane) {
        int out = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bool o;
            switch (plane) {
            case 0: o = p[k][0] < -p[k][3]; break;
            case 1: o = p[k][0] > p[k][3]; break;
            case 2: o = p[k][1] < -p[k][3]; break;
            case 3: o = p[k][1] > p[k][3]; break;
            case 4: o = p[k][2] < 0.0f; break;
            default: o = p[k][2] > p[k][3]; break;
            }
            out += o ? 1 : 0;
        }
        reject |= out == 3;
    }
    // needs clipping: w <= 0, z outside [0, w], or outside the guard band
    // (a piece: w <= 0 or outside the band drops it)
    bool unsup = false;
    int32_t sx[3], sy[3];
    float zf[3], iw[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        unsup |= !(p[k][3] > 0.0f);
        unsup |= !piece && (!(p[k][2] >= 0.0f) || !(p[k][2] <= p[k][3]));
        // x / 1 == x exactly: skip the IEEE divisions when w is exactly one
        const bool w1 = p[k][3] == 1.0f;
        const float xd = w1 ? p[k][0] : p[k][0] / p[k][3];
        const float yd = w1 ? p[k][1] : p[k][1] / p[k][3];
        const float zd = w1 ? p[k][2] : p[k][2] / p[k][3];
        const float xf = xd * f.vp_hw + f.vp_ox;
        const float yf = yd * f.vp_hh + f.vp_oy;
        zf[k] = zd * f.vp_dz + f.vp_zmin;
        const bool band = xf >= -8192.0f && xf <= 8191.0f && yf >= -8192.0f && yf <= 8191.0f;
        unsup |= !band;
        // (outside the band, or NaN: a stand-in that converts; the status discards it)
        sx[k] = (int32_t)__builtin_rintf(band ? xf * 16.0f : 0.0f);
        sy[k] = (int32_t)__builtin_rintf(band ? yf * 16.0f : 0.0f);
        iw[k] = w1 ? 1.0f : 1.0f / p[k][3];
    }
    t.x0 = sx[0];
    t.y0 = sy[0];
    t.d1x = sx[1] - sx[0];
    t.d1y = sy[1] - sy[0];
    t.d2x = sx[2] - sx[0];
    t.d2y = sy[2] - sy[0];
    const int64_t S = (int64_t)t.d1x * t.d2y - (int64_t)t.d2x * t.d1y;
    const bool front = (KZ_FRONT_FACE == 0) ? (S < 0) : (S > 0);
    const bool culled = ((KZ_CULL_MODE & 1) && front) || ((KZ_CULL_MODE & 2) && !front);
    t.S = S;
    t.invS = 1.0f / (float)S;
    t.z0 = zf[0];
    t.dz1 = zf[1] - zf[0];
    t.dz2 = zf[2] - zf[0];
    t.iw0 = iw[0];
    t.iw1 = iw[1];
    t.iw2 = iw[2];
    const int32_t xmin = min(sx[0], min(sx[1], sx[2])), xmax = max(sx[0], max(sx[1], sx[2]));
    const int32_t ymin = min(sy[0], min(sy[1], sy[2])), ymax = max(sy[0], max(sy[1], sy[2]));
    t.px0 = max(xmin >> 4, f.sc_x0);
    t.px1 = min(xmax >> 4, f.sc_x1);
    t.py0 = max(ymin >> 4, f.sc_y0);
    t.py1 = min(ymax >> 4, f.sc_y1);
    if (piece) return (unsup || S == 0 || culled) ? 0 : 1;
    return reject ? 0 : unsup ? 2 : (S == 0 || culled) ? 0 : 1;
}

__device__ __forceinline__ int kz_setup(const KzFrame &f, const float p[3][4], KzTri &t) {
    return kz_setup_rt(f, p, t, false);
}

// ---------------------------------------------------------------------------
// clipping (a triangle kz_setup gives status 2), one wave per triangle
//
// Homogeneous Sutherland-Hodgman against z >= 0, z <= w and the guard
// band's planes x = gx * w, y = gy * w (screen x, y = -8000 and 8000 px,
// inside viewportBoundsRange [-8192, 8191], vulkan-driver/src/
// api_impl.rs:1568), in that order. A polygon vertex is a record of NC
// floats: its clip position, then its varyings. Per plane, vertex i is
// kept when its distance d >= 0, and the edge (i, i + 1) adds its crossing
// point when its ends lie on different sides, interpolated from the inside
// end, every component alike: t = d_in / (d_in - d_out),
// v = v_in + t * (v_out - v_in) (so the shared edge of two triangles gets
// the same point from both). The emitted vertices, in that order, are the
// next polygon (at most KZ_CLIP_MAX; fewer than 3 drops the triangle).
// Lane i holds vertex i: a plane costs a ballot pair for the output
// positions and a gather of each output from the lane that emitted it.
// The pieces are the fan (0, j + 1, j + 2), j = 0 .. n - 3, drawn in that
// order in the triangle's place.
#define KZ_CLIP_MAX 12

struct KzClipBand {
    float xl, xh, yl, yh;
};

__device__ __forceinline__ KzClipBand kz_clip_band(const KzFrame &f) {
    const float ax = (-8000.0f - f.vp_ox) / f.vp_hw, bx = (8000.0f - f.vp_ox) / f.vp_hw;
    const float ay = (-8000.0f - f.vp_oy) / f.vp_hh, by = (8000.0f - f.vp_oy) / f.vp_hh;
    return {fminf(ax, bx), fmaxf(ax, bx), fminf(ay, by), fmaxf(ay, by)};
}

__device__ __forceinline__ float kz_clip_distance(int plane, const KzClipBand &g, const float *q) {
    switch (plane) {
    case 0: return q[2];
    case 1: return q[3] - q[2];
    case 2: return q[0] - g.xl * q[3];
    case 3: return g.xh * q[3] - q[0];
    case 4: return q[1] - g.yl * q[3];
    default: return g.yh * q[3] - q[1];
    }
}

// The whole wave calls this with lane `src`'s triangle (p, vary: that
// lane's values count). Returns the polygon's vertex count (0: clipped
// away); lane i < count holds vertex i in v (NC = 4 + varyings carried).
template <int NC>
__device__ __forceinline__ int kz_clip_polygon(const KzFrame &f, const float p[3][4], const float vary[3][KZ_NVA],
                                               uint32_t src, uint32_t lane, float v[NC]) {
    const KzClipBand g = kz_clip_band(f);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float ra = kz_readlane_f(p[0][c], src), rb = kz_readlane_f(p[1][c], src),
                    rd = kz_readlane_f(p[2][c], src);
        v[c] = lane == 0 ? ra : lane == 1 ? rb : rd;
    }
#pragma unroll
    for (int i = 0; i < NC - 4; ++i) {
        const float ra = kz_readlane_f(vary[0][i], src), rb = kz_readlane_f(vary[1][i], src),
                    rd = kz_readlane_f(vary[2][i], src);
        v[4 + i] = lane == 0 ? ra : lane == 1 ? rb : rd;
    }
    int n = 3;
#pragma unroll 1
    for (int plane = 0; plane < 6; ++plane) {
        const bool act = lane < (uint32_t)n;
        const uint32_t nxt = lane + 1 >= (uint32_t)n ? 0u : lane + 1;
        float u[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) u[c] = kz_bperm_f(v[c], nxt);
        const float di = kz_clip_distance(plane, g, v), dj = kz_clip_distance(plane, g, u);
        const bool ini = di >= 0.0f, inj = dj >= 0.0f;
        const bool keep = act && ini, cross = act && ini != inj;
        const uint64_t mk = kz_ballot(keep), mc = kz_ballot(cross);
        const int total = __popcll(mk) + __popcll(mc);
        if (total < 3) return 0;
        const uint32_t off = kz_rank(mk) + kz_rank(mc);
        const uint32_t end = off + (keep ? 1u : 0u) + (cross ? 1u : 0u);
        // output vertex `lane` comes from the first lane whose emissions end after it
        uint32_t from = 0;
        for (int i = 0; i < n; ++i) from += (uint32_t)kz_readlane_i((int32_t)end, (uint32_t)i) <= lane ? 1u : 0u;
        from = min(from, (uint32_t)(n - 1));
        // (both reads in full EXEC: a bpermute under a branch returns
        // nothing from a source lane the branch switched off, so no
        // short-circuit here)
        const int32_t keep_from = kz_bperm_i(keep ? 1 : 0, from), off_from = kz_bperm_i((int32_t)off, from);
        const bool kept = (keep_from != 0) & (lane == (uint32_t)off_from);
        // the edge's crossing point, from its inside end
        const float din = ini ? di : dj, dout = ini ? dj : di;
        const float t = din / (din - dout);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const float vin = ini ? v[c] : u[c], vout = ini ? u[c] : v[c];
            const float x = vin + t * (vout - vin);
            const float kv = kz_bperm_f(v[c], from), xv = kz_bperm_f(x, from);
            v[c] = kept ? kv : xv;
        }
        n = min(total, KZ_CLIP_MAX);
    }
    return n;
}

// Piece j's vertices (polygon vertices 0, j + 1, j + 2) from the lanes of
// kz_clip_polygon: clip positions into p, the NC - 4 varyings into vary.
template <int NC>
__device__ __forceinline__ void kz_clip_piece(const float v[NC], uint32_t j, float p[3][4], float vary[3][KZ_NVA]) {
    const uint32_t src[3] = {0u, (j + 1) & 63u, (j + 2) & 63u};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int c = 0; c < 4; ++c) p[q][c] = kz_bperm_f(v[c], src[q]);
#pragma unroll
        for (int i = 0; i < NC - 4; ++i) vary[q][i] = kz_bperm_f(v[4 + i], src[q]);
    }
}

// Coverage constants: the interior-positive orientation and the top-left
// thresholds (0 on a top-left edge, 1 otherwise) of E01, E12, E20; `small`:
// every edge delta is under 8192 (512 px), so that for any sample in the
// triangle's pixel box (|r| < 2^14) the edge products stay under 2^27 and
// E01, E20, E12 and S fit 32 bits (24-bit multiplies, same integers).
struct KzCov {
    int32_t flip, th01, th12, th20, small;
};

__device__ __forceinline__ int32_t kz_tl_threshold(int64_t A, int64_t B) { return (A > 0 || (A == 0 && B > 0)) ? 0 : 1; }

__device__ __forceinline__ bool kz_small_deltas(int32_t d1x, int32_t d1y, int32_t d2x, int32_t d2y) {
    return max(max(abs(d1x), abs(d1y)), max(abs(d2x), abs(d2y))) < 8192;
}

__device__ __forceinline__ KzCov kz_cov_setup(const KzTri &t) {
    KzCov c;
    const int64_t sg = t.S < 0 ? -1 : 1;
    c.flip = t.S < 0 ? 1 : 0;
    c.th01 = kz_tl_threshold(-(int64_t)t.d1y * sg, (int64_t)t.d1x * sg);
    c.th20 = kz_tl_threshold((int64_t)t.d2y * sg, -(int64_t)t.d2x * sg);
    c.th12 = kz_tl_threshold(-((int64_t)t.d2y - t.d1y) * sg, ((int64_t)t.d2x - t.d1x) * sg);
    c.small = kz_small_deltas(t.d1x, t.d1y, t.d2x, t.d2y) ? 1 : 0;
    return c;
}

// Signed (unflipped) edge values at absolute sample position (sx, sy) in
// 1/16 px: E01, E20 and E12 = S - E01 - E20 (the three sum to S). The
// sample lies in the triangle's pixel box.
__device__ __forceinline__ void kz_edges(int32_t x0, int32_t y0, int32_t d1x, int32_t d1y, int32_t d2x, int32_t d2y,
                                         int64_t S, int32_t sx, int32_t sy, int64_t &e01, int64_t &e20, int64_t &e12) {
    const int32_t rx = sx - x0, ry = sy - y0;
    if (kz_small_deltas(d1x, d1y, d2x, d2y)) {
        const int32_t a = __mul24(d1x, ry) - __mul24(d1y, rx);
        const int32_t b = __mul24(d2y, rx) - __mul24(d2x, ry);
        e01 = a;
        e20 = b;
        e12 = (int64_t)((int32_t)S - a - b);
        return;
    }
    e01 = (int64_t)d1x * ry - (int64_t)d1y * rx;
    e20 = (int64_t)d2y * rx - (int64_t)d2x * ry;
    e12 = S - e01 - e20;
}

// kz_edges' values converted to float. `narrow` (wave-uniform): the triangle
// has short edges (kz_small_deltas), so the values fit 32 bits and the
// int32 -> float conversion rounds them as the int64 one does (to nearest
// even), without the 64-bit products and conversions.
__device__ __forceinline__ void kz_edges_f(bool narrow, int32_t x0, int32_t y0, int32_t d1x, int32_t d1y,
                                           int32_t d2x, int32_t d2y, int64_t S, int32_t sx, int32_t sy, float &f01,
                                           float &f20, float &f12) {
    if (narrow) {
        const int32_t rx = sx - x0, ry = sy - y0;
        const int32_t a = __mul24(d1x, ry) - __mul24(d1y, rx);
        const int32_t b = __mul24(d2y, rx) - __mul24(d2x, ry);
        f01 = (float)a;
        f20 = (float)b;
        f12 = (float)((int32_t)S - a - b);
        return;
    }
    int64_t e01, e20, e12;
    kz_edges(x0, y0, d1x, d1y, d2x, d2y, S, sx, sy, e01, e20, e12);
    f01 = (float)e01;
    f20 = (float)e20;
    f12 = (float)e12;
}

__device__ __forceinline__ bool kz_inside(int64_t e01, int64_t e20, int64_t e12, int32_t flip, int32_t th01,
                                          int32_t th12, int32_t th20) {
    if (flip) {
        e01 = -e01;
        e20 = -e20;
        e12 = -e12;
    }
    return e01 >= th01 && e12 >= th12 && e20 >= th20;
}

__constant__ const int32_t kz_sample_x[4] = {KZ_SAMPLES == 4 ? 6 : 8, 14, 2, 10};
__constant__ const int32_t kz_sample_y[4] = {KZ_SAMPLES == 4 ? 2 : 8, 6, 10, 14};

__device__ __forceinline__ uint32_t kz_cover_mask(const KzTri &t, const KzCov &c, int32_t px, int32_t py) {
    uint32_t m = 0;
    if (c.small) {
        const int32_t S = (int32_t)t.S;
#pragma unroll
        for (int s = 0; s < KZ_SAMPLES; ++s) {
            const int32_t rx = px * 16 + kz_sample_x[s] - t.x0, ry = py * 16 + kz_sample_y[s] - t.y0;
            int32_t e01 = __mul24(t.d1x, ry) - __mul24(t.d1y, rx);
            int32_t e20 = __mul24(t.d2y, rx) - __mul24(t.d2x, ry);
            int32_t e12 = S - e01 - e20;
            if (c.flip) {
                e01 = -e01;
                e20 = -e20;
                e12 = -e12;
            }
            if (e01 >= c.th01 && e12 >= c.th12 && e20 >= c.th20) m |= 1u << s;
        }
        return m;
    }
#pragma unroll
    for (int s = 0; s < KZ_SAMPLES; ++s) {
        int64_t e01, e20, e12;
        kz_edges(t.x0, t.y0, t.d1x, t.d1y, t.d2x, t.d2y, t.S, px * 16 + kz_sample_x[s], py * 16 + kz_sample_y[s], e01,
                 e20, e12);
        if (kz_inside(e01, e20, e12, c.flip, c.th01, c.th12, c.th20)) m |= 1u << s;
    }
    return m;
}

// Incremental coverage of a triangle with c.small, walked pixel by pixel
// (row-major) over a box of its pixel box. Per edge E01, E12, E20 it keeps
// the value at the current pixel's first sample with the orientation applied
// and the top-left threshold subtracted, so a sample is inside iff its three
// values are >= 0, i.e. iff their bitwise OR is. E = A * rx + B * ry + C with
// (A, B, C) = (-d1y, d1x, 0) for E01, (d2y, -d2x, 0) for E20 and
// (d1y - d2y, d2x - d1x, S) for E12; a step to the next pixel of a row adds
// 16 * A, from a row's last pixel (box width w) to the next row's first
// 16 * (B - (w - 1) * A); another sample of the pixel adds a constant
// offset. Every value is an edge value at a point of the pixel box, so it
// fits 32 bits (c.small) and equals the one kz_cover_mask computes.
struct KzWalk {
    int32_t e[3];                // at the current pixel
    int32_t ax[3], wy[3];        // step to the next pixel, to the next row
    int32_t o[3][KZ_SAMPLES];    // sample s relative to sample 0
};

// the walk of the box of width w whose first pixel is (px, py)
__device__ __forceinline__ void kz_walk_init(KzWalk &w, const KzTri &t, const KzCov &c, int32_t px, int32_t py,
                                             int32_t width) {
    // (wrapping 32-bit arithmetic: a lane may run this on a triangle that
    // is not small, and then never uses the values)
    const uint32_t sg = c.flip ? ~0u : 1u;
    const uint32_t d1x = (uint32_t)t.d1x, d1y = (uint32_t)t.d1y, d2x = (uint32_t)t.d2x, d2y = (uint32_t)t.d2y;
    const int32_t A[3] = {(int32_t)((0u - d1y) * sg), (int32_t)((d1y - d2y) * sg), (int32_t)(d2y * sg)};
    const int32_t B[3] = {(int32_t)(d1x * sg), (int32_t)((d2x - d1x) * sg), (int32_t)((0u - d2x) * sg)};
    const uint32_t C[3] = {0u, (uint32_t)t.S * sg, 0u};
    const int32_t th[3] = {c.th01, c.th12, c.th20};
    const int32_t rx = px * 16 + kz_sample_x[0] - t.x0, ry = py * 16 + kz_sample_y[0] - t.y0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w.e[k] = (int32_t)((uint32_t)__mul24(A[k], rx) + (uint32_t)__mul24(B[k], ry) + C[k] - (uint32_t)th[k]);
        w.ax[k] = (int32_t)((uint32_t)A[k] << 4);
        w.wy[k] = (int32_t)(((uint32_t)B[k] - (uint32_t)__mul24(width - 1, A[k])) << 4);
#pragma unroll
        for (int s = 0; s < KZ_SAMPLES; ++s)
            w.o[k][s] = __mul24(A[k], kz_sample_x[s] - kz_sample_x[0]) + __mul24(B[k], kz_sample_y[s] - kz_sample_y[0]);
    }
}

// at one sample per pixel: bit 31 set iff the current pixel is outside
__device__ __forceinline__ uint32_t kz_walk_or(const KzWalk &w) {
    return (uint32_t)(w.e[0] | w.e[1] | w.e[2]);
}

// at four samples per pixel: bit s set iff sample s of the current pixel is
// outside (wrapping 32-bit sums, as the walk's steps)
__device__ __forceinline__ uint32_t kz_walk_out4(const KzWalk &w) {
    uint32_t m = 0;
#pragma unroll
    for (int s = 0; s < KZ_SAMPLES && s < 4; ++s) {
        const uint32_t v = ((uint32_t)w.e[0] + (uint32_t)w.o[0][s]) | ((uint32_t)w.e[1] + (uint32_t)w.o[1][s]) |
                           ((uint32_t)w.e[2] + (uint32_t)w.o[2][s]);
        m |= (v >> 31) << s;
    }
    return m;
}

__device__ __forceinline__ uint32_t kz_walk_mask(const KzWalk &w) {
    uint32_t m = 0;
#pragma unroll
    for (int s = 0; s < KZ_SAMPLES; ++s) {
        const int32_t v = (w.e[0] + w.o[0][s]) | (w.e[1] + w.o[1][s]) | (w.e[2] + w.o[2][s]);
        m |= (v >= 0 ? 1u : 0u) << s;
    }
    return m;
}

// to the next pixel; `wrap`: the current one ends its row
__device__ __forceinline__ void kz_walk_step(KzWalk &w, bool wrap) {
#pragma unroll
    for (int k = 0; k < 3; ++k) w.e[k] = (int32_t)((uint32_t)w.e[k] + (uint32_t)(wrap ? w.wy[k] : w.ax[k]));
}

// The draw of batch triangle g: the per-triangle table binning wrote, the
// only draw, or a search.
__device__ __forceinline__ uint32_t kz_draw_of(const KzFrame &f, uint32_t g) {
    if (f.ndraws == 1) return 0u;
    return f.tri_draw ? f.tri_draw[g] : kz_find_draw(f, g);
}

// Vertex indices of batch triangle g of draw `draw`.
__device__ __forceinline__ void kz_triangle_indices(const KzFrame &f, uint32_t g, uint32_t draw, uint32_t idx[3]) {
    const KzDraw &d = f.draws[draw];
    const uint32_t t = g - d.tri_begin;
#pragma unroll
    for (int j = 0; j < 3; ++j) idx[j] = kz_vertex_index(d, t, j);
}

// Vertex shader outputs of the vertices idx of draw `draw`.
__device__ __forceinline__ void kz_fetch_vertices(const KzFrame &f, uint32_t draw, const uint32_t idx[3],
                                                  float p[3][4], float vary[3][KZ_NVA]) {
    const KzDraw &d = f.draws[draw];
#pragma unroll
    for (int j = 0; j < 3; ++j) kz_vertex(f, d, idx[j], p[j], vary[j]);
}

// Vertex shader outputs of batch triangle g of draw `draw`.
__device__ __forceinline__ void kz_fetch_triangle(const KzFrame &f, uint32_t g, uint32_t draw, float p[3][4],
                                                  float vary[3][KZ_NVA]) {
    const KzDraw &d = f.draws[draw];
    const uint32_t t = g - d.tri_begin;
#pragma unroll
    for (int j = 0; j < 3; ++j) kz_vertex(f, d, kz_vertex_index(d, t, j), p[j], vary[j]);
}

// ---------------------------------------------------------------------------
// binning, pass 1: vertex shader + setup of every triangle, to its pixel box
// (the tile lists are built from the boxes by the prebuilt binning kernels)

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2. Renumbering block b (on XCD b % 8) into a contiguous range per XCD
// gives every XCD a compact band of tiles, so that triangles and texels
// shared by neighbouring tiles are fetched into one L2, not eight. A
// bijection of [0, n) for any n.
__device__ __forceinline__ uint32_t kz_xcd_remap(uint32_t b, uint32_t n) {
    const uint32_t xcd = b & 7u, k = b >> 3, q = n >> 3, r = n & 7u;
    return xcd * q + min(xcd, r) + k;
}

__device__ __forceinline__ unsigned long long *kz_stats_stripe(const KzFrame &f) {
    return f.stats + (size_t)(blockIdx.x & (KZ_STAT_STRIPES - 1)) * KZ_STAT_COUNT;
}

// Setup records: what the raster needs of a set-up triangle (edge and
// plane constants, varyings' base and deltas), KZ_REC_WORDS 32-bit words,
// written once per triangle so that a tile's raster does one gather per
// list entry instead of index -> vertex -> setup again.
#define KZ_REC_WORDS KZ_REC_WORDS_FOR(KZ_NV)

__device__ __forceinline__ void kz_store_rec(const KzFrame &f, uint32_t g, const KzTri &t,
                                             const float vary[3][KZ_NVA]) {
    uint32_t w[KZ_REC_WORDS];
    w[0] = (uint32_t)t.x0;
    w[1] = (uint32_t)t.y0;
    w[2] = (uint32_t)t.d1x;
    w[3] = (uint32_t)t.d1y;
    w[4] = (uint32_t)t.d2x;
    w[5] = (uint32_t)t.d2y;
    w[6] = (uint32_t)(uint64_t)t.S;
    w[7] = (uint32_t)((uint64_t)t.S >> 32);
    w[8] = __float_as_uint(t.invS);
    w[9] = __float_as_uint(t.z0);
    w[10] = __float_as_uint(t.dz1);
    w[11] = __float_as_uint(t.dz2);
    w[12] = __float_as_uint(t.iw0);
    w[13] = __float_as_uint(t.iw1);
    w[14] = __float_as_uint(t.iw2);
    w[15] = 0u;
#pragma unroll
    for (int i = 0; i < KZ_NV; ++i) {
        w[16 + 3 * i] = __float_as_uint(vary[0][i]);
        w[17 + 3 * i] = __float_as_uint(vary[1][i] - vary[0][i]);
        w[18 + 3 * i] = __float_as_uint(vary[2][i] - vary[0][i]);
    }
#pragma unroll
    for (int i = 16 + 3 * KZ_NV; i < KZ_REC_WORDS; ++i) w[i] = 0u;
    uint4 *dst = (uint4 *)(f.tri_rec + (size_t)g * KZ_REC_WORDS);
#pragma unroll
    for (int q = 0; q < KZ_REC_WORDS / 4; ++q) dst[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

// The record of a triangle that needs clipping: word 15 = 1, its vertex
// shader outputs instead of a setup (clip position P[v][c] at word 4v + c,
// varying A[v][i] at 16 + 3i + v), from which the raster clips it.
__device__ __forceinline__ void kz_store_clip_rec(const KzFrame &f, uint32_t g, const float p[3][4],
                                                  const float vary[3][KZ_NVA]) {
    uint32_t w[KZ_REC_WORDS];
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int c = 0; c < 4; ++c) w[4 * v + c] = __float_as_uint(p[v][c]);
    w[12] = w[13] = w[14] = 0u;
    w[15] = 1u;
#pragma unroll
    for (int i = 0; i < KZ_NV; ++i)
#pragma unroll
        for (int v = 0; v < 3; ++v) w[16 + 3 * i + v] = __float_as_uint(vary[v][i]);
#pragma unroll
    for (int i = 16 + 3 * KZ_NV; i < KZ_REC_WORDS; ++i) w[i] = 0u;
    uint4 *dst = (uint4 *)(f.tri_rec + (size_t)g * KZ_REC_WORDS);
#pragma unroll
    for (int q = 0; q < KZ_REC_WORDS / 4; ++q) dst[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

__device__ __forceinline__ void kz_load_clip_rec(const KzFrame &f, uint32_t g, float p[3][4], float vary[3][KZ_NVA]) {
    const uint32_t *src = f.tri_rec + (size_t)g * KZ_REC_WORDS;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
#pragma unroll
        for (int c = 0; c < 4; ++c) p[v][c] = __uint_as_float(src[4 * v + c]);
#pragma unroll
        for (int i = 0; i < KZ_NVA; ++i) vary[v][i] = i < KZ_NV ? __uint_as_float(src[16 + 3 * i + v]) : 0.0f;
    }
}

// The record of triangle g and its pixel box: the values kz_setup and the
// varying differences gave when the record was written. Returns whether the
// triangle needs clipping (its other values are then meaningless).
__device__ __forceinline__ bool kz_load_rec(const KzFrame &f, uint32_t g, KzTri &t, float *a0, float *da1,
                                            float *da2) {
    const uint4 *src = (const uint4 *)(f.tri_rec + (size_t)g * KZ_REC_WORDS);
    uint32_t w[KZ_REC_WORDS];
#pragma unroll
    for (int q = 0; q < KZ_REC_WORDS / 4; ++q) {
        const uint4 v = src[q];
        w[4 * q] = v.x;
        w[4 * q + 1] = v.y;
        w[4 * q + 2] = v.z;
        w[4 * q + 3] = v.w;
    }
    const uint64_t r = f.tri_rect[g];
    t.x0 = (int32_t)w[0];
    t.y0 = (int32_t)w[1];
    t.d1x = (int32_t)w[2];
    t.d1y = (int32_t)w[3];
    t.d2x = (int32_t)w[4];
    t.d2y = (int32_t)w[5];
    t.S = (int64_t)((uint64_t)w[6] | ((uint64_t)w[7] << 32));
    t.invS = __uint_as_float(w[8]);
    t.z0 = __uint_as_float(w[9]);
    t.dz1 = __uint_as_float(w[10]);
    t.dz2 = __uint_as_float(w[11]);
    t.iw0 = __uint_as_float(w[12]);
    t.iw1 = __uint_as_float(w[13]);
    t.iw2 = __uint_as_float(w[14]);
    t.px0 = (int32_t)(r & 0xFFFF);
    t.py0 = (int32_t)((r >> 16) & 0xFFFF);
    t.px1 = (int32_t)((r >> 32) & 0xFFFF);
    t.py1 = (int32_t)(r >> 48);
#pragma unroll
    for (int i = 0; i < KZ_NV; ++i) {
        a0[i] = __uint_as_float(w[16 + 3 * i]);
        da1[i] = __uint_as_float(w[17 + 3 * i]);
        da2[i] = __uint_as_float(w[18 + 3 * i]);
    }
    return w[15] != 0u;
}

extern "C" __global__ __launch_bounds__(KZ_BIN_BLOCK) void kz_bin_count(KzFrame f) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = kz_lane();
    uint32_t n_setup = 0, n_clip = 0, n_pairs = 0;
    float p[3][4], vary[3][KZ_NVA];
    KzTri t;
    int st = 0;
    uint64_t rect = KZ_RECT_EMPTY;
    if (g < f.total_tris) {
        const uint32_t draw = kz_find_draw(f, g);
        if (f.tri_draw) f.tri_draw[g] = draw;  // the raster's lookup, one load instead of a search
        kz_fetch_triangle(f, g, draw, p, vary);
        st = kz_setup(f, p, t);
        if (st == 1) {
            n_setup = 1;
            // A small triangle's box shrinks to the pixels it covers, and a
            // triangle that covers no sample is not binned at all: pixels
            // cannot depend on it (no fragment, no depth or colour write).
            if (t.px0 <= t.px1 && t.py0 <= t.py1 && (t.px1 - t.px0 + 1) * (t.py1 - t.py0 + 1) <= KZ_TIGHT_CAND) {
                const KzCov c = kz_cov_setup(t);
                int32_t cx0 = t.px1 + 1, cy0 = t.py1 + 1, cx1 = t.px0 - 1, cy1 = t.py0 - 1;
                if (c.small) {
                    KzWalk wk;
                    kz_walk_init(wk, t, c, t.px0, t.py0, t.px1 - t.px0 + 1);
                    int32_t px = t.px0, py = t.py0;
                    const int32_t nc = (t.px1 - t.px0 + 1) * (t.py1 - t.py0 + 1);
                    for (int32_t j = 0; j < nc; ++j) {
                        if (kz_walk_mask(wk)) {
                            cx0 = min(cx0, px);
                            cx1 = max(cx1, px);
                            cy0 = min(cy0, py);
                            cy1 = max(cy1, py);
                        }
                        const bool wrap = px == t.px1;
                        kz_walk_step(wk, wrap);
                        py += wrap ? 1 : 0;
                        px = wrap ? t.px0 : px + 1;
                    }
                } else {
                    for (int32_t py = t.py0; py <= t.py1; ++py)
                        for (int32_t px = t.px0; px <= t.px1; ++px)
                            if (kz_cover_mask(t, c, px, py)) {
                                cx0 = min(cx0, px);
                                cx1 = max(cx1, px);
                                cy0 = min(cy0, py);
                                cy1 = max(cy1, py);
                            }
                }
                t.px0 = cx0;
                t.px1 = cx1;
                t.py0 = cy0;
                t.py1 = cy1;
            }
            if (t.px0 <= t.px1 && t.py0 <= t.py1) {
                rect = (uint64_t)(uint32_t)t.px0 | ((uint64_t)(uint32_t)t.py0 << 16) |
                       ((uint64_t)(uint32_t)t.px1 << 32) | ((uint64_t)(uint32_t)t.py1 << 48);
                if (KZ_SETUP_RECORDS && f.tri_rec) kz_store_rec(f, g, t, vary);
            }
            if (f.tri_z) {  // cost-model feature: the triangle's depth range
                const float z1 = t.z0 + t.dz1, z2 = t.z0 + t.dz2;
                const float lo = fminf(fmaxf(fminf(t.z0, fminf(z1, z2)), 0.0f), 1.0f);
                const float hi = fminf(fmaxf(fmaxf(t.z0, fmaxf(z1, z2)), 0.0f), 1.0f);
                f.tri_z[g] = (uint32_t)(lo * 65535.0f + 0.5f) | ((uint32_t)(hi * 65535.0f + 0.5f) << 16);
            }
        }
    }
    // Triangles that need clipping: the whole wave clips each in turn (lane
    // j sets up piece j); the triangle is binned by the union of its
    // pieces' pixel boxes, and counts as set up when one of them is.
    uint64_t cm = kz_ballot(st == 2);
    while (cm) {
        const uint32_t b = (uint32_t)__builtin_ctzll(cm);
        cm &= cm - 1;
        float v[4], q[3][4], qv[3][KZ_NVA];
        const int np = kz_clip_polygon<4>(f, p, vary, b, lane, v);  // (the box needs positions only)
        kz_clip_piece<4>(v, lane, q, qv);
        KzTri pt;
        const bool piece = (int)lane + 2 < np && kz_setup_rt(f, q, pt, true) == 1;
        const bool boxed = piece && pt.px0 <= pt.px1 && pt.py0 <= pt.py1;
        const int32_t none = -0x40000000;  // below any pixel coordinate (an empty box when no piece has one)
        const int32_t bx0 = -kz_wave_max(boxed ? -pt.px0 : none, lane);
        const int32_t by0 = -kz_wave_max(boxed ? -pt.py0 : none, lane);
        const int32_t bx1 = kz_wave_max(boxed ? pt.px1 : none, lane);
        const int32_t by1 = kz_wave_max(boxed ? pt.py1 : none, lane);
        const bool any = kz_ballot(piece) != 0;
        if (lane == b) {
            n_clip = 1;
            n_setup = any ? 1 : 0;
            if (bx0 <= bx1 && by0 <= by1) {
                t.px0 = bx0;
                t.py0 = by0;
                t.px1 = bx1;
                t.py1 = by1;
                rect = (uint64_t)(uint32_t)bx0 | ((uint64_t)(uint32_t)by0 << 16) | ((uint64_t)(uint32_t)bx1 << 32) |
                       ((uint64_t)(uint32_t)by1 << 48);
            }
        }
    }
    if (g < f.total_tris) {
        if (st == 2) {
            if (KZ_SETUP_RECORDS && f.tri_rec && rect != KZ_RECT_EMPTY) kz_store_clip_rec(f, g, p, vary);
            if (f.tri_z) f.tri_z[g] = 0xFFFF0000u;  // depth range [0, 1]: the pieces are not kept
        }
        f.tri_rect[g] = rect;
        // (triangle, tile) pairs: the tiles of the pixel box
        if (rect != KZ_RECT_EMPTY)
            n_pairs = (uint32_t)(((t.px1 >> KZ_TILE_SHIFT) - (t.px0 >> KZ_TILE_SHIFT) + 1) *
                                 ((t.py1 >> KZ_TILE_SHIFT) - (t.py0 >> KZ_TILE_SHIFT) + 1));
    }
    const uint64_t ms = kz_ballot(n_setup != 0), mc = kz_ballot(n_clip != 0), mi = kz_ballot(g < f.total_tris);
    if (lane == 0) {
        unsigned long long *sts = kz_stats_stripe(f);
        if (mi) atomicAdd(&sts[KZ_STAT_TRIANGLES_IN], (unsigned long long)__popcll(mi));
        if (ms) atomicAdd(&sts[KZ_STAT_TRIANGLES_SETUP], (unsigned long long)__popcll(ms));
        if (mc) atomicAdd(&sts[KZ_STAT_TRIANGLES_CLIPPED], (unsigned long long)__popcll(mc));
    }
    if (f.block_pairs) {  // the block's pair total, for the pair offsets of radix binning
        __shared__ uint32_t s_pairs[KZ_BIN_BLOCK / 64];
        for (uint32_t d = 32; d > 0; d >>= 1) n_pairs += (uint32_t)__shfl_xor((int)n_pairs, (int)d, 64);
        if (lane == 0) s_pairs[threadIdx.x >> 6] = n_pairs;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t s = 0;
            for (uint32_t w = 0; w < KZ_BIN_BLOCK / 64; ++w) s += s_pairs[w];
            f.block_pairs[blockIdx.x] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// fixed-function fragment operations

__device__ __forceinline__ uint32_t kz_wrap(int32_t i, uint32_t n, uint32_t mode) {
    if (mode == 0) {  // REPEAT
        // a power-of-two size (wave-uniform): the low bits are the residue
        if ((n & (n - 1)) == 0) return (uint32_t)i & (n - 1);
        const int32_t r = i % (int32_t)n;
        return (uint32_t)(r < 0 ? r + (int32_t)n : r);
    }
    return (uint32_t)(i < 0 ? 0 : (i >= (int32_t)n ? (int32_t)n - 1 : i));  // CLAMP_TO_EDGE
}

__device__ __forceinline__ float kz_clamp_coord(float v) {
    const float lim = 1.0e8f;
    v = fminf(fmaxf(v, -lim), lim);
    return v != v ? 0.0f : v;
}

// Combined image sampler lookup on a linear RGBA8 image. LINEAR uses 4-bit
// weights (subTexelPrecisionBits = 4, vulkan-driver/src/api_impl.rs:1560):
// f = u*(W*16) - 8, i = floor(f), texel = i >> 4, weight = i & 15, integer
// lerp over 16ths, result = sum * (1/65280). NEAREST takes floor(u*W).
// Texels are addressed by a 32-bit byte offset from the image's base (an
// image is at most 16384 x 16384, maxImageDimension2D, so rows and offsets
// fit 24-bit multiplies and 32 bits): one vector offset per load against a
// wave-uniform base, no 64-bit address arithmetic.
__device__ __forceinline__ uint32_t kz_texel(const KzTexture &t, uint32_t row, uint32_t x) {
    return *(const uint32_t *)(t.data + ((row + x) << 2));
}

__device__ __forceinline__ void kz_sample(const KzTexture &t, float u, float v, float out[4]) {
    if (t.filter == 0) {
        const int32_t x = (int32_t)floorf(kz_clamp_coord(u * (float)t.w));
        const int32_t y = (int32_t)floorf(kz_clamp_coord(v * (float)t.h));
        const uint32_t texel = kz_texel(t, __umul24(kz_wrap(y, t.h, t.addr_v), t.w), kz_wrap(x, t.w, t.addr_u));
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] = (float)((texel >> (8 * c)) & 255u) * (1.0f / 255.0f);
        return;
    }
    const float fu = kz_clamp_coord(u * (float)(t.w * 16u) - 8.0f);
    const float fv = kz_clamp_coord(v * (float)(t.h * 16u) - 8.0f);
    const int32_t iu = (int32_t)floorf(fu), iv = (int32_t)floorf(fv);
    const int32_t x0 = iu >> 4, y0 = iv >> 4;
    const uint32_t ax = (uint32_t)(iu & 15), ay = (uint32_t)(iv & 15);
    const uint32_t xa = kz_wrap(x0, t.w, t.addr_u), xb = kz_wrap(x0 + 1, t.w, t.addr_u);
    const uint32_t ya = kz_wrap(y0, t.h, t.addr_v), yb = kz_wrap(y0 + 1, t.h, t.addr_v);
    const uint32_t row_a = __umul24(ya, t.w), row_b = __umul24(yb, t.w);
    const uint32_t t00 = kz_texel(t, row_a, xa), t10 = kz_texel(t, row_a, xb), t01 = kz_texel(t, row_b, xa),
                   t11 = kz_texel(t, row_b, xb);
    // two channels per 32-bit lerp: R and B (then G and A) in 16-bit fields;
    // the weights of each lerp sum to 16, so a field stays <= 255 * 16 * 16
    // = 65280 and never carries into the next (the same integers per channel)
    const uint32_t m = 0x00FF00FFu, wx = 16u - ax, wy = 16u - ay;
    const uint32_t rb = ((t00 & m) * wx + (t10 & m) * ax) * wy + ((t01 & m) * wx + (t11 & m) * ax) * ay;
    const uint32_t ga = (((t00 >> 8) & m) * wx + ((t10 >> 8) & m) * ax) * wy +
                        (((t01 >> 8) & m) * wx + ((t11 >> 8) & m) * ax) * ay;
    out[0] = (float)(rb & 0xFFFFu) * (1.0f / 65280.0f);
    out[1] = (float)(ga & 0xFFFFu) * (1.0f / 65280.0f);
    out[2] = (float)(rb >> 16) * (1.0f / 65280.0f);
    out[3] = (float)(ga >> 16) * (1.0f / 65280.0f);
}

__device__ __forceinline__ float kz_clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

__device__ __forceinline__ bool kz_depth_cmp(float z, float d) {
#if KZ_DEPTH_OP == 0
    return false;
#elif KZ_DEPTH_OP == 1
    return z < d;
#elif KZ_DEPTH_OP == 2
    return z == d;
#elif KZ_DEPTH_OP == 3
    return z <= d;
#elif KZ_DEPTH_OP == 4
    return z > d;
#elif KZ_DEPTH_OP == 5
    return z != d;
#elif KZ_DEPTH_OP == 6
    return z >= d;
#else
    return true;
#endif
}

__device__ __forceinline__ float kz_blend_factor(int fct, const float s[4], const float d[4], int c) {
    switch (fct) {
    case 0: return 0.0f;
    case 1: return 1.0f;
    case 2: return s[c];
    case 3: return 1.0f - s[c];
    case 4: return d[c];
    case 5: return 1.0f - d[c];
    case 6: return s[3];
    case 7: return 1.0f - s[3];
    case 8: return d[3];
    case 9: return 1.0f - d[3];
    case 14: return c == 3 ? 1.0f : fminf(s[3], 1.0f - d[3]);
    default: return 0.0f;
    }
}

__device__ __forceinline__ float kz_blend_op(int op, float s, float sf, float d, float df) {
    switch (op) {
    case 0: return s * sf + d * df;
    case 1: return s * sf - d * df;
    case 2: return d * df - s * sf;
    case 3: return fminf(s, d);
    default: return fmaxf(s, d);
    }
}

__device__ __forceinline__ uint32_t kz_encode8(float v) { return (uint32_t)(kz_clamp01(v) * 255.0f + 0.5f); }

__device__ __forceinline__ uint32_t kz_rop(uint32_t dst, const float src[4]) {
#if KZ_BLEND
    float d[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) d[c] = (float)((dst >> (8 * c)) & 255u) * (1.0f / 255.0f);
    float r[4];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        r[c] = kz_blend_op(KZ_COLOR_OP, src[c], kz_blend_factor(KZ_SRC_COLOR, src, d, c), d[c],
                           kz_blend_factor(KZ_DST_COLOR, src, d, c));
    r[3] = kz_blend_op(KZ_ALPHA_OP, src[3], kz_blend_factor(KZ_SRC_ALPHA, src, d, 3), d[3],
                       kz_blend_factor(KZ_DST_ALPHA, src, d, 3));
#else
    const float *r = src;
#endif
    uint32_t out = dst;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if ((KZ_WRITE_MASK >> c) & 1) out = (out & ~(255u << (8 * c))) | (kz_encode8(r[c]) << (8 * c));
    return out;
}

// ---------------------------------------------------------------------------
// the tile rasteriser

// Sample masks of a small triangle's candidate pixels (KZ_SAMPLES bits for
// each of at most KZ_SMALL_CAND pixels)
#if KZ_SAMPLES * KZ_SMALL_CAND <= 32
typedef uint32_t KzCandMask;
#else
typedef uint64_t KzCandMask;
#endif

struct KzLaneTri {
    KzTri t;
    KzCov c;
    float a0[KZ_NVA], da1[KZ_NVA], da2[KZ_NVA];
    int32_t cx0, cy0, cx1, cy1;  // bbox ∩ tile
    bool live;
};

// Does pixel box r overlap the square [X0, X0 + RW) x [Y0, Y0 + RW)?
__device__ __forceinline__ bool kz_box_overlaps(uint64_t r, int32_t X0, int32_t Y0, int32_t RW) {
    if (r == KZ_RECT_EMPTY) return false;
    const int32_t x0 = (int32_t)(r & 0xFFFF), y0 = (int32_t)((r >> 16) & 0xFFFF);
    const int32_t x1 = (int32_t)((r >> 32) & 0xFFFF), y1 = (int32_t)(r >> 48);
    return x1 >= X0 && x0 < X0 + RW && y1 >= Y0 && y0 < Y0 + RW;
}

__device__ __forceinline__ uint32_t kz_pack_frag(uint32_t pix, uint32_t seq, uint32_t mask) {
    return (pix << 10) | (seq << 4) | mask;
}

// A batch's triangles, acquired one batch ahead: the ids (list entries k ..
// k + 63) are requested while the batch before is set up, their draws and
// vertex indices while it resolves its fragments, so that a batch starts
// with only its vertex loads outstanding.
struct KzAhead {
    uint32_t k;       // first list entry
    uint32_t g;       // the lane's triangle
    uint32_t draw;
    uint32_t idx[3];  // its vertex indices (not with setup records)
    bool valid;
};

__device__ __forceinline__ void kz_ahead_ids(KzAhead &a, const uint32_t *ids, uint32_t k, uint32_t n,
                                             uint32_t lane) {
    a.k = k;
    a.valid = k + lane < n;
    a.g = a.valid ? (ids ? ids[k + lane] : k + lane) : 0u;
}

__device__ __forceinline__ void kz_ahead_indices(KzAhead &a, const KzFrame &f) {
    a.draw = 0;
    a.idx[0] = a.idx[1] = a.idx[2] = 0;
    if (!a.valid || (KZ_SETUP_RECORDS && f.tri_rec)) return;
    a.draw = kz_draw_of(f, a.g);
    kz_triangle_indices(f, a.g, a.draw, a.idx);
}

// Resolves the fragment queue s_frag[0, total) of the batch held by the
// wave's lanes (L), 64 fragments at a time: depth test and write, varyings,
// fragment shader and blend, in primitive order per pixel. NARROW: every
// triangle of the batch has short edges (32-bit edge values).
template <bool NARROW>
__device__ __forceinline__ void kz_resolve(const KzFrame &f, const KzLaneTri &L, int32_t X0, int32_t Y0,
                                           uint32_t total, uint32_t lane, uint32_t *s_color, float *s_depth,
                                           uint32_t *s_claim, const uint32_t *s_frag, unsigned long long &nfrag,
                                           unsigned long long &npass) {
    for (uint32_t c0 = 0; c0 < total; c0 += 64) {
        const uint32_t qi = c0 + lane;
        const bool active = qi < total;
        const uint32_t e = active ? s_frag[qi] : 0u;
        const uint32_t pix = e >> 10, seq = (e >> 4) & 63u, mask = e & 15u;
        // the owning triangle's data, from lane `seq`
        const int32_t x0 = kz_bperm_i(L.t.x0, seq), y0 = kz_bperm_i(L.t.y0, seq);
        const int32_t d1x = kz_bperm_i(L.t.d1x, seq), d1y = kz_bperm_i(L.t.d1y, seq);
        const int32_t d2x = kz_bperm_i(L.t.d2x, seq), d2y = kz_bperm_i(L.t.d2y, seq);
        const int32_t s_lo = kz_bperm_i((int32_t)(uint32_t)(uint64_t)L.t.S, seq);
        const int32_t s_hi = NARROW ? (s_lo >> 31) : kz_bperm_i((int32_t)((uint64_t)L.t.S >> 32), seq);
        const int64_t S = (int64_t)(((uint64_t)(uint32_t)s_lo) | ((uint64_t)(uint32_t)s_hi << 32));
        const float invS = kz_bperm_f(L.t.invS, seq);
        const float z0 = kz_bperm_f(L.t.z0, seq), dz1 = kz_bperm_f(L.t.dz1, seq), dz2 = kz_bperm_f(L.t.dz2, seq);
        const int32_t px = X0 + (int32_t)(pix & (KZ_TS - 1)), py = Y0 + (int32_t)(pix >> KZ_TILE_SHIFT);

        float zs[KZ_SAMPLES];
        float c01 = 0.0f, c20 = 0.0f, c12 = 0.0f;  // one sample per pixel: it is the pixel centre
        uint32_t nf = 0;
#pragma unroll
        for (int s = 0; s < KZ_SAMPLES; ++s) {
            zs[s] = 0.0f;
            if ((mask >> s) & 1u) {
                ++nf;
                float e01, e20, e12;
                kz_edges_f(NARROW, x0, y0, d1x, d1y, d2x, d2y, S, px * 16 + kz_sample_x[s],
                           py * 16 + kz_sample_y[s], e01, e20, e12);
                if (KZ_SAMPLES == 1) {
                    c01 = e01;
                    c20 = e20;
                    c12 = e12;
                }
                const float l1 = e20 * invS, l2 = e01 * invS;
                const float z = (z0 + l1 * dz1) + l2 * dz2;
                zs[s] = fminf(fmaxf(z, f.z_lo), f.z_hi);
            }
        }
        nfrag += nf;

        // The samples that can still pass the depth test: those that pass
        // against the depth the tile holds now. The round's earlier
        // fragments only move a pixel's depth away from passing (or leave
        // it), so the others fail after them too; only NOT_EQUAL with depth
        // writes is not monotone. The fragment shader runs for the rest.
        uint32_t maybe = active ? mask : 0u;
#if KZ_DEPTH_TEST && !(KZ_DEPTH_OP == 5 && KZ_DEPTH_WRITE)
        {
            uint32_t m = 0;
#pragma unroll
            for (int s = 0; s < KZ_SAMPLES; ++s)
                if (((maybe >> s) & 1u) && kz_depth_cmp(zs[s], s_depth[pix * KZ_SAMPLES + s])) m |= 1u << s;
            maybe = m;
        }
#endif

        // varyings at the pixel centre and the fragment shader
        float col[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        {
            float a0[KZ_NVA], da1[KZ_NVA], da2[KZ_NVA];
#pragma unroll
            for (int i = 0; i < KZ_NV; ++i) {
                a0[i] = kz_bperm_f(L.a0[i], seq);
                da1[i] = kz_bperm_f(L.da1[i], seq);
                da2[i] = kz_bperm_f(L.da2[i], seq);
            }
            const float iw0 = kz_bperm_f(L.t.iw0, seq), iw1 = kz_bperm_f(L.t.iw1, seq),
                        iw2 = kz_bperm_f(L.t.iw2, seq);
            if (maybe) {
                float e01 = c01, e20 = c20, e12 = c12;
                if (KZ_SAMPLES != 1)
                    kz_edges_f(NARROW, x0, y0, d1x, d1y, d2x, d2y, S, px * 16 + 8, py * 16 + 8, e01, e20, e12);
                const float l0 = e12 * invS, l1 = e20 * invS, l2 = e01 * invS;
                const float q = ((l0 * iw0) + (l1 * iw1)) + (l2 * iw2);
                const float rq = 1.0f / q;
                const float b1 = (l1 * iw1) * rq, b2 = (l2 * iw2) * rq;
                float vary[KZ_NVA];
#pragma unroll
                for (int i = 0; i < KZ_NV; ++i) vary[i] = (a0[i] + b1 * da1[i]) + b2 * da2[i];
                const float zc = fminf(fmaxf((z0 + l1 * dz1) + l2 * dz2, f.z_lo), f.z_hi);
                const float frag_coord[4] = {(float)px + 0.5f, (float)py + 0.5f, zc, q};
                // a discarded fragment takes no part in the depth test,
                // depth write or blend
                if (!kz_fragment(f, vary, frag_coord, col)) maybe = 0;
#pragma unroll
                for (int c = 0; c < 4; ++c) col[c] = kz_clamp01(col[c]);
            }
        }

        // depth test and write, blend and write, in primitive order per
        // pixel: the lowest triangle of each pixel claims it, then the next
        uint32_t pass = 0;
        bool act = maybe != 0;
        while (kz_ballot(act)) {
            if (act) atomicMin(&s_claim[pix], seq);
            __syncthreads();
            const bool win = act && s_claim[pix] == seq;
            if (win) {
#pragma unroll
                for (int s = 0; s < KZ_SAMPLES; ++s) {
                    if (!((maybe >> s) & 1u)) continue;
#if KZ_DEPTH_TEST
                    float &dd = s_depth[pix * KZ_SAMPLES + s];
                    if (!kz_depth_cmp(zs[s], dd)) continue;
#if KZ_DEPTH_WRITE
                    dd = zs[s];
#endif
#endif
                    pass |= 1u << s;
                    s_color[pix * KZ_SAMPLES + s] = kz_rop(s_color[pix * KZ_SAMPLES + s], col);
                }
                s_claim[pix] = KZ_NO_CLAIM;
                act = false;
            }
            __syncthreads();
        }
        npass += (uint32_t)__popc(pass);
    }
}

// The pieces ks + lane of the clip polygon of triangle g0 (one that needs
// clipping), set up, with their varyings' base and deltas; ok: the lane's
// piece exists and set up. The whole wave calls it, and only for a batch
// that starts with such a triangle, so it stays out of line: its
// temporaries do not weigh on the raster's registers.
struct KzClipBatch {
    KzTri t;
    float a0[KZ_NVA], da1[KZ_NVA], da2[KZ_NVA];
    uint32_t ok, npieces;
};

__device__ __forceinline__ KzClipBatch kz_clip_batch(const KzFrame &f, uint32_t g0, uint32_t ks, uint32_t lane) {
    float p[3][4], vary[3][KZ_NVA];
    if (KZ_SETUP_RECORDS && f.tri_rec) {  // its record holds the vertex shader outputs
        kz_load_clip_rec(f, g0, p, vary);
    } else {
        const uint32_t draw = kz_draw_of(f, g0);
        uint32_t idx[3];
        kz_triangle_indices(f, g0, draw, idx);
        kz_fetch_vertices(f, draw, idx, p, vary);
    }
    float v[4 + KZ_NV];
    const int np = kz_clip_polygon<4 + KZ_NV>(f, p, vary, 0, lane, v);
    KzClipBatch o;
    o.npieces = np >= 3 ? (uint32_t)(np - 2) : 0u;
    kz_clip_piece<4 + KZ_NV>(v, ks + lane, p, vary);
    const int st = kz_setup_rt(f, p, o.t, true);
    o.ok = ks + lane < o.npieces && st == 1 ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < KZ_NVA; ++i) {
        o.a0[i] = i < KZ_NV ? vary[0][i] : 0.0f;
        o.da1[i] = i < KZ_NV ? vary[1][i] - vary[0][i] : 0.0f;
        o.da2[i] = i < KZ_NV ? vary[2][i] - vary[0][i] : 0.0f;
    }
    return o;
}

// Rasterises n triangles into the square of side 1 << RS at (X0, Y0) (a
// whole tile, or a sub-tile of the balanced mode), in primitive order. The
// triangles are ids[0, n) (ids = null: triangles 0 .. n-1); with `filter`,
// only those whose pixel box overlaps the square (the fallback when a list
// did not fit its storage: the parent tile's list, or every triangle).
// Returns the square's fragment count (to every lane).
__device__ unsigned long long kz_raster_rect(const KzFrame &f, int32_t X0, int32_t Y0, uint32_t RS,
                                             const uint32_t *ids, uint32_t n, bool filter, uint32_t clear_mask,
                                             uint32_t *s_color, float *s_depth, uint32_t *s_claim,
                                             uint32_t *s_frag) {
    const uint32_t lane = kz_lane();
    if (n == 0 && !clear_mask) return 0;  // (with folded clears the square is written regardless)
    const int32_t RW = 1 << RS;
    const uint32_t *gcolor = (const uint32_t *)f.color;

    // the first batch's ids, requested before the tile's colour and depth
    KzAhead nx;
    if (!filter) kz_ahead_ids(nx, ids, 0, n, lane);

    for (uint32_t j = lane; j < (uint32_t)(RW * RW); j += 64) {
        const int32_t x = X0 + (int32_t)(j & (RW - 1)), y = Y0 + (int32_t)(j >> RS);
        const uint32_t i = (uint32_t)((y - Y0) * KZ_TS + (x - X0));
        if (x < (int32_t)f.width && y < (int32_t)f.height) {
            const size_t gp = ((size_t)y * f.width + x) * KZ_SAMPLES;
#pragma unroll
            for (int s = 0; s < KZ_SAMPLES; ++s) {
                s_color[i * KZ_SAMPLES + s] = (clear_mask & 1u) ? f.clear_color : gcolor[gp + s];
                s_depth[i * KZ_SAMPLES + s] = (clear_mask & 2u) ? f.clear_depth : f.depth[gp + s];
            }
        }
        s_claim[i] = KZ_NO_CLAIM;
    }
    __syncthreads();

    unsigned long long nfrag = 0, npass = 0;
    // a scan of the triangles (filter without ids) starts at the first cut
    // chunk's first triangle: the ones before it were in the tile's list
    uint32_t k = filter && !ids && f.bin_ctrl ? min(f.bin_ctrl[KZ_BIN_CUT] * KZ_PAIR_CHUNK, n) : 0u;
    uint32_t ks = 0;  // first piece still to draw of a clipped triangle at entry k
    if (!filter) kz_ahead_indices(nx, f);

    while (k < n) {
        uint32_t g = 0, draw = 0, idx[3] = {0u, 0u, 0u};
        bool valid = k + lane < n;
        if (filter) {
            if (valid) g = ids ? ids[k + lane] : k + lane;
            valid = valid && kz_box_overlaps(f.tri_rect[g], X0, Y0, RW);
            if (kz_ballot(valid) == 0) {
                k += 64;
                continue;
            }
            if (valid && !(KZ_SETUP_RECORDS && f.tri_rec)) {
                draw = kz_draw_of(f, g);
                kz_triangle_indices(f, g, draw, idx);
            }
        } else {
            g = nx.g;
            valid = nx.valid;
            draw = nx.draw;
            idx[0] = nx.idx[0];
            idx[1] = nx.idx[1];
            idx[2] = nx.idx[2];
            // the next batch's ids, assuming it starts 64 entries on (the
            // batch cut below corrects a shorter batch)
            kz_ahead_ids(nx, ids, k + 64, n, lane);
        }

        // --- per-lane triangle: vertex shader + setup -------------------
        KzLaneTri L;
        // every listed triangle was set up by kz_bin_count: its record, or
        // (without records) the vertex shader and setup once more. Without
        // records every lane runs them, a lane without a triangle on vertex
        // indices 0 of draw 0 (discarded), so that the code is straight-line.
        bool ok, clip = false;
        float p[3][4], vary[3][KZ_NVA];  // (without records) the lane's vertex shader outputs
        if (KZ_SETUP_RECORDS && f.tri_rec) {
            L.t.x0 = L.t.y0 = L.t.d1x = L.t.d1y = L.t.d2x = L.t.d2y = 0;
            L.t.S = 0;
            L.t.invS = L.t.z0 = L.t.dz1 = L.t.dz2 = 0.0f;
            L.t.iw0 = L.t.iw1 = L.t.iw2 = 0.0f;
            L.t.px0 = L.t.py0 = 0;
            L.t.px1 = L.t.py1 = -1;
#pragma unroll
            for (int i = 0; i < KZ_NVA; ++i) L.a0[i] = L.da1[i] = L.da2[i] = 0.0f;
            if (valid) clip = kz_load_rec(f, g, L.t, L.a0, L.da1, L.da2);
            ok = valid && !clip;
        } else {
            kz_fetch_vertices(f, draw, idx, p, vary);
        }
        if (!(KZ_SETUP_RECORDS && f.tri_rec)) {
            const int st = kz_setup_rt(f, p, L.t, false);
            ok = valid && st == 1;
            clip = valid && st == 2;
#pragma unroll
            for (int i = 0; i < KZ_NVA; ++i) {
                L.a0[i] = i < KZ_NV ? vary[0][i] : 0.0f;
                L.da1[i] = i < KZ_NV ? vary[1][i] - vary[0][i] : 0.0f;
                L.da2[i] = i < KZ_NV ? vary[2][i] - vary[0][i] : 0.0f;
            }
        }
        // A triangle that needs clipping is drawn as its own batch of
        // pieces (lane j: piece ks + j), in its place: a batch whose first
        // triangle needs it draws that triangle's pieces (out of line), any
        // other batch ends before its first such triangle.
        const uint64_t clipm = kz_ballot(clip);
        const bool clip_batch = (clipm & 1ull) != 0;
        const uint32_t first_clip = clipm ? (uint32_t)__builtin_ctzll(clipm) : 64u;
        uint32_t npieces = 0;
        if (clip_batch) {
            const KzClipBatch cb = kz_clip_batch(f, (uint32_t)kz_readlane_i((int32_t)g, 0), ks, lane);
            L.t = cb.t;
#pragma unroll
            for (int i = 0; i < KZ_NVA; ++i) {
                L.a0[i] = cb.a0[i];
                L.da1[i] = cb.da1[i];
                L.da2[i] = cb.da2[i];
            }
            ok = cb.ok != 0;
            npieces = cb.npieces;
        } else {
            ok = ok && lane < first_clip;
        }
        L.cx0 = max(L.t.px0, X0);
        L.cx1 = min(L.t.px1, X0 + RW - 1);
        L.cy0 = max(L.t.py0, Y0);
        L.cy1 = min(L.t.py1, Y0 + RW - 1);
        L.live = ok && L.cx0 <= L.cx1 && L.cy0 <= L.cy1;
        L.c = kz_cov_setup(L.t);
        const int32_t cw = L.cx1 - L.cx0 + 1;
        const int32_t cand = L.live ? cw * (L.cy1 - L.cy0 + 1) : 0;
        // a triangle with few candidate pixels here and short edges: one lane
        // walks its candidates (the others: the whole wave, one at a time)
        const bool small = cand <= KZ_SMALL_CAND && L.c.small;
        // every triangle of the batch that can have fragments here has short
        // edges: the resolve evaluates edges in 32 bits
        const bool narrow = kz_ballot(L.live && !L.c.small) == 0;

        // --- phase 1: covered pixels per triangle ------------------------
        // (small triangles keep their sample masks, KZ_SAMPLES bits per
        // candidate pixel, for phase 2: at most 24 x 1 or 16 x 4 bits)
        uint32_t cnt = 0;
        KzCandMask cmask = 0;
        {
            // (the wave's longest walk, a scalar loop bound)
            const int32_t maxc = __builtin_amdgcn_readfirstlane(kz_wave_max(small ? cand : 0, lane));
            KzWalk wk;
#if KZ_SAMPLES == 1
            // Every lane steps maxc times (a lane that walks nothing, or
            // less, steps on past its box, in wrapping 32-bit arithmetic;
            // those steps are dropped below), shifting each pixel's sign bit
            // (set: outside) in at bit 31, so that pixel j of maxc ends at bit
            // 32 - maxc + j: two instructions a pixel, no mask switching.
            kz_walk_init(wk, L.t, L.c, L.cx0, L.cy0, cw);
            int32_t dx = 0;
            uint32_t out = 0;
            for (int32_t j = 0; j < maxc; ++j) {
                const uint32_t v = kz_walk_or(wk);
                const bool wrap = dx == cw - 1;
                kz_walk_step(wk, wrap);
                dx = wrap ? 0 : dx + 1;
                out = (out >> 1) | (v & 0x80000000u);
            }
            if (small && maxc > 0) {
                cmask = ~(out >> (32 - maxc)) & ((1u << cand) - 1u);
                cnt = (uint32_t)__popc(cmask);
            }
#elif KZ_SAMPLES == 4
            // The same at four samples: a pixel's four sign bits (set:
            // sample outside) shift in at the top of a 64-bit word, so that
            // pixel j of maxc (at most 16) ends at bits 4 (16 - maxc + j) ..
            // + 3; the count is of pixels with any sample covered.
            kz_walk_init(wk, L.t, L.c, L.cx0, L.cy0, cw);
            int32_t dx = 0;
            uint64_t out = 0;
            for (int32_t j = 0; j < maxc; ++j) {
                const uint32_t om = kz_walk_out4(wk);
                const bool wrap = dx == cw - 1;
                kz_walk_step(wk, wrap);
                dx = wrap ? 0 : dx + 1;
                out = (out >> 4) | ((uint64_t)om << 60);
            }
            if (small && maxc > 0) {
                const uint64_t in = ~(out >> (64 - 4 * maxc));  // pixel j at bits 4 j .. 4 j + 3
                cmask = cand >= 16 ? in : in & ((1ull << (4 * cand)) - 1ull);
                uint64_t any = cmask | (cmask >> 1);
                any |= any >> 2;
                cnt = (uint32_t)__popcll(any & 0x1111111111111111ull);
            }
#else
            if (small) kz_walk_init(wk, L.t, L.c, L.cx0, L.cy0, cw);
            int32_t dx = 0;
            for (int32_t j = 0; j < maxc; ++j) {
                if (small && j < cand) {
                    const uint32_t m = kz_walk_mask(wk);
                    const bool wrap = dx == cw - 1;
                    kz_walk_step(wk, wrap);
                    dx = wrap ? 0 : dx + 1;
                    if (m) {
                        ++cnt;
                        cmask |= (KzCandMask)m << (j * KZ_SAMPLES);
                    }
                }
            }
#endif
        }
        uint64_t bigm = kz_ballot(L.live && !small);
        while (bigm) {
            const uint32_t b = (uint32_t)__builtin_ctzll(bigm);
            bigm &= bigm - 1;
            KzTri tb;
            tb.x0 = kz_readlane_i(L.t.x0, b);
            tb.y0 = kz_readlane_i(L.t.y0, b);
            tb.d1x = kz_readlane_i(L.t.d1x, b);
            tb.d1y = kz_readlane_i(L.t.d1y, b);
            tb.d2x = kz_readlane_i(L.t.d2x, b);
            tb.d2y = kz_readlane_i(L.t.d2y, b);
            tb.S = (int64_t)(((uint64_t)(uint32_t)kz_readlane_i((int32_t)(uint32_t)(uint64_t)L.t.S, b)) |
                             ((uint64_t)(uint32_t)kz_readlane_i((int32_t)((uint64_t)L.t.S >> 32), b) << 32));
            KzCov cb;
            cb.flip = kz_readlane_i(L.c.flip, b);
            cb.th01 = kz_readlane_i(L.c.th01, b);
            cb.th12 = kz_readlane_i(L.c.th12, b);
            cb.th20 = kz_readlane_i(L.c.th20, b);
            cb.small = kz_readlane_i(L.c.small, b);
            const int32_t bx0 = kz_readlane_i(L.cx0, b), by0 = kz_readlane_i(L.cy0, b);
            const int32_t bw = kz_readlane_i(cw, b), bc = kz_readlane_i(cand, b);
            // candidate i is pixel (i % bw, i / bw) of the box; the row is
            // floor((i + 0.5) / bw) through the reciprocal: that quotient is
            // at least 0.5 / bw from an integer, far above the rounding
            // error for i < 2^21
            const float rbw = __builtin_amdgcn_rcpf((float)bw);
            uint32_t c = 0;
            for (int32_t i0 = 0; i0 < bc; i0 += 64) {
                const int32_t i = i0 + (int32_t)lane;
                bool cov = false;
                if (i < bc) {
                    const int32_t dy = (int32_t)(((float)i + 0.5f) * rbw);
                    cov = kz_cover_mask(tb, cb, bx0 + i - dy * bw, by0 + dy) != 0;
                }
                c += (uint32_t)__popcll(kz_ballot(cov));
            }
            if (lane == b) cnt = c;
        }

        // --- batch cut: keep the longest lane prefix whose fragments fit ---
        const uint32_t incl = kz_wave_incl_scan(cnt, lane);
        // (a batch that is not a clip batch also ends before its first
        // triangle that needs clipping)
        const uint32_t take = min((uint32_t)__popcll(kz_ballot(incl <= KZ_FRAG_CAP)), clip_batch ? 64u : first_clip);
        const bool taken = lane < take;
        const uint32_t total = (uint32_t)kz_readlane_i((int32_t)incl, take - 1);
        const uint32_t excl = incl - cnt;
        // where the next batch starts: past the pieces or the lanes taken
        uint32_t next_k, next_ks = 0;
        if (clip_batch) {
            next_ks = ks + take < npieces ? ks + take : 0u;
            next_k = next_ks ? k : k + 1;
        } else {
            next_k = k + take;
        }
        if (!filter && nx.k != next_k) kz_ahead_ids(nx, ids, next_k, n, lane);

        // --- phase 2: write the fragment queue in primitive order ---------
        if (taken && small) {  // the covered candidates, in row-major order
            // candidate j of the box is pixel (j % cw, j / cw), and
            // floor((j + 0.5) / cw) == j / cw for j < 64, cw <= 64 even with
            // the reciprocal's rounding
            const float rcw = __builtin_amdgcn_rcpf((float)cw);
            uint32_t w = excl;
            KzCandMask bits = cmask;
            while (bits) {
                const uint32_t j = (uint32_t)__builtin_ctzll((uint64_t)bits) / KZ_SAMPLES;
                const uint32_t m = (uint32_t)(bits >> (j * KZ_SAMPLES)) & ((1u << KZ_SAMPLES) - 1u);
                bits &= ~((KzCandMask)((1u << KZ_SAMPLES) - 1u) << (j * KZ_SAMPLES));
                const int32_t dy = (int32_t)(((float)j + 0.5f) * rcw);
                const int32_t cx = L.cx0 + (int32_t)j - dy * cw, cy = L.cy0 + dy;
                s_frag[w++] = kz_pack_frag((uint32_t)((cy - Y0) * KZ_TS + (cx - X0)), lane, m);
            }
        }
        bigm = kz_ballot(taken && L.live && !small);
        while (bigm) {
            const uint32_t b = (uint32_t)__builtin_ctzll(bigm);
            bigm &= bigm - 1;
            KzTri tb;
            tb.x0 = kz_readlane_i(L.t.x0, b);
            tb.y0 = kz_readlane_i(L.t.y0, b);
            tb.d1x = kz_readlane_i(L.t.d1x, b);
            tb.d1y = kz_readlane_i(L.t.d1y, b);
            tb.d2x = kz_readlane_i(L.t.d2x, b);
            tb.d2y = kz_readlane_i(L.t.d2y, b);
            tb.S = (int64_t)(((uint64_t)(uint32_t)kz_readlane_i((int32_t)(uint32_t)(uint64_t)L.t.S, b)) |
                             ((uint64_t)(uint32_t)kz_readlane_i((int32_t)((uint64_t)L.t.S >> 32), b) << 32));
            KzCov cb;
            cb.flip = kz_readlane_i(L.c.flip, b);
            cb.th01 = kz_readlane_i(L.c.th01, b);
            cb.th12 = kz_readlane_i(L.c.th12, b);
            cb.th20 = kz_readlane_i(L.c.th20, b);
            cb.small = kz_readlane_i(L.c.small, b);
            const int32_t bx0 = kz_readlane_i(L.cx0, b), by0 = kz_readlane_i(L.cy0, b);
            const int32_t bw = kz_readlane_i(cw, b), bc = kz_readlane_i(cand, b);
            uint32_t w = (uint32_t)kz_readlane_i((int32_t)excl, b);
            const float rbw = __builtin_amdgcn_rcpf((float)bw);  // (rows as in phase 1)
            for (int32_t i0 = 0; i0 < bc; i0 += 64) {
                const int32_t i = i0 + (int32_t)lane;
                uint32_t m = 0;
                int32_t px = 0, py = 0;
                if (i < bc) {
                    const int32_t dy = (int32_t)(((float)i + 0.5f) * rbw);
                    px = bx0 + i - dy * bw;
                    py = by0 + dy;
                    m = kz_cover_mask(tb, cb, px, py);
                }
                const uint64_t bal = kz_ballot(m != 0);
                if (m) s_frag[w + kz_rank(bal)] = kz_pack_frag((uint32_t)((py - Y0) * KZ_TS + (px - X0)), b, m);
                w += (uint32_t)__popcll(bal);
            }
        }
        __syncthreads();
        // the next batch's draws and vertex indices, during this resolve
        if (!filter) kz_ahead_indices(nx, f);

        // --- resolve the queue 64 fragments at a time --------------------
        if (narrow)
            kz_resolve<true>(f, L, X0, Y0, total, lane, s_color, s_depth, s_claim, s_frag, nfrag, npass);
        else
            kz_resolve<false>(f, L, X0, Y0, total, lane, s_color, s_depth, s_claim, s_frag, nfrag, npass);
        k = next_k;
        ks = next_ks;
        __syncthreads();
    }

    for (uint32_t j = lane; j < (uint32_t)(RW * RW); j += 64) {
        const int32_t x = X0 + (int32_t)(j & (RW - 1)), y = Y0 + (int32_t)(j >> RS);
        const uint32_t i = (uint32_t)((y - Y0) * KZ_TS + (x - X0));
        if (x < (int32_t)f.width && y < (int32_t)f.height) {
            const size_t gp = ((size_t)y * f.width + x) * KZ_SAMPLES;
#pragma unroll
            for (int s = 0; s < KZ_SAMPLES; ++s) {
                ((uint32_t *)f.color)[gp + s] = s_color[i * KZ_SAMPLES + s];
                f.depth[gp + s] = s_depth[i * KZ_SAMPLES + s];
            }
        }
    }
    nfrag = kz_wave_sum_u64(nfrag, lane);
    npass = kz_wave_sum_u64(npass, lane);
    if (lane == 0) {
        unsigned long long *st = kz_stats_stripe(f);
        atomicAdd(&st[KZ_STAT_FRAGMENTS], nfrag);
        atomicAdd(&st[KZ_STAT_DEPTH_PASSED], npass);
    }
    __syncthreads();
    return nfrag;
}

// Occupancy hints (resident waves per SIMD the compiler must allow for;
// 0: its own choice), one per raster kernel
#if KZ_RASTER_WAVES
#define KZ_RASTER_ATTR __attribute__((amdgpu_waves_per_eu(KZ_RASTER_WAVES)))
#else
#define KZ_RASTER_ATTR
#endif
#if KZ_PERSIST_WAVES
#define KZ_PERSIST_ATTR __attribute__((amdgpu_waves_per_eu(KZ_PERSIST_WAVES)))
#else
#define KZ_PERSIST_ATTR
#endif

// Static mode: one workgroup per screen tile, in tile order.

extern "C" __global__ __launch_bounds__(64) KZ_RASTER_ATTR void kz_raster(KzFrame f) {
    __shared__ uint32_t s_color[KZ_TP * KZ_SAMPLES];
    __shared__ float s_depth[KZ_TP * KZ_SAMPLES];
    __shared__ uint32_t s_claim[KZ_TP];
    __shared__ uint32_t s_frag[KZ_FRAG_CAP];
    const uint32_t tile = kz_xcd_remap(blockIdx.x, gridDim.x);
    const int32_t X0 = (int32_t)((tile % f.tiles_x) << KZ_TILE_SHIFT);
    const int32_t Y0 = (int32_t)((tile / f.tiles_x) << KZ_TILE_SHIFT);
    uint32_t off, n;
    kz_tile_list(f, tile, off, n);
    const uint64_t t0 = f.tile_cycles ? clock64() : 0;
    // A tile flagged by a cut (the pairs of the triangles from the first
    // cut chunk on were dropped) draws its list, then the triangles from
    // the cut on by an in-order scan of their pixel boxes; a list past the
    // capacity without a flag (sort binning) is a scan of every triangle.
    // (An empty tile returns at once unless clears are folded.)
    const bool cut = f.tile_flag && f.tile_flag[tile];
    const bool scan = cut || (uint64_t)off + n > (uint64_t)f.list_capacity;
    uint32_t lcnt = scan ? 0u : n;
    const uint32_t *lids = f.tile_lists + off;
    if (cut) {
        lcnt = f.tile_counts[tile];
        lids = f.tile_lists + (lcnt ? f.tile_offsets[tile] : 0u);
    }
    unsigned long long frags =
        kz_raster_rect(f, X0, Y0, KZ_TILE_SHIFT, lids, lcnt, false, f.clear_mask, s_color, s_depth, s_claim, s_frag);
    if (scan)
        frags += kz_raster_rect(f, X0, Y0, KZ_TILE_SHIFT, nullptr, f.total_tris, true, 0u, s_color, s_depth, s_claim,
                                s_frag);
    // tile profile: the cost model's label (fragments) and raster cycles
    if (f.tile_cycles && kz_lane() == 0) {
        const uint64_t dt = n ? clock64() - t0 : 0;
        f.tile_cycles[tile] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
        f.tile_frags[tile] = frags > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)frags;
    }
}

// Balanced mode: persistent workgroups take work items (tiles or
// sub-tiles) from the queue, most expensive first. The queue is dealt into
// KZ_QUEUES sub-queues (item i to sub-queue i % KZ_QUEUES: each stays in
// cost order, and their totals are close), one per XCD: workgroup b (on XCD
// b % 8) takes its sub-queue's item b / 8 in the first round, then the next
// untaken one from the sub-queue's counter. One counter for the whole
// queue, an atomic per item on one address from every workgroup, cost more
// than the balance it bought (a 4M-triangle cluster: 158 us of raster,
// against 103 us with a fixed order); eight counters keep the balance. With
// folded clears, the tiles without triangles (they have no items) follow,
// strided over the workgroups. One call site per filter mode of
// kz_raster_rect serves every case.
extern "C" __global__ __launch_bounds__(64) KZ_PERSIST_ATTR void kz_raster_persistent(KzFrame f) {
    __shared__ uint32_t s_color[KZ_TP * KZ_SAMPLES];
    __shared__ float s_depth[KZ_TP * KZ_SAMPLES];
    __shared__ uint32_t s_claim[KZ_TP];
    __shared__ uint32_t s_frag[KZ_FRAG_CAP];
    const uint32_t lane = kz_lane();
    const uint32_t n_items = f.ctrl[KZ_CTRL_ITEMS];
    const uint32_t ntiles = f.tiles_x * f.tiles_y;
    // sub-queue x of K, this workgroup's place k in its first round, Gx
    // workgroups on it (a grid smaller than KZ_QUEUES uses one queue)
    const uint32_t K = gridDim.x >= KZ_QUEUES ? KZ_QUEUES : 1u;
    const uint32_t x = blockIdx.x % K, Gx = (gridDim.x - x + K - 1u) / K;
    uint32_t q = x + K * (blockIdx.x / K);  // the item in hand
    uint32_t e = blockIdx.x;                // next tile to check for emptiness
    for (;;) {
        int32_t X0, Y0;
        uint32_t RS, n;
        const uint32_t *ids;
        bool filter = false, cut = false;
        const bool item = q < n_items;
        if (item) {
            const uint32_t geo = f.queue[q];
            const uint32_t tile = KZ_ITEM_TILE(geo), level = KZ_ITEM_LEVEL(geo), sub = KZ_ITEM_SUB(geo);
            RS = KZ_TILE_SHIFT - level;
            X0 = (int32_t)(((tile % f.tiles_x) << KZ_TILE_SHIFT) + ((sub & ((1u << level) - 1)) << RS));
            Y0 = (int32_t)(((tile / f.tiles_x) << KZ_TILE_SHIFT) + ((sub >> level) << RS));
            uint32_t toff, tn;
            kz_tile_list(f, tile, toff, tn);
            if (level == 0) {
                cut = f.tile_flag && f.tile_flag[tile];
                if (cut) {  // its list, then (pass 1) the triangles from the cut on
                    n = f.tile_counts[tile];
                    ids = n ? f.tile_lists + f.tile_offsets[tile] : nullptr;
                } else {
                    filter = (uint64_t)toff + tn > (uint64_t)f.list_capacity;
                    ids = filter ? nullptr : f.tile_lists + toff;
                    n = filter ? f.total_tris : tn;
                }
            } else {
                // a split tile's list always fits; its sub-tile lists may
                // not, and then the sub-tile filters the tile's list
                const uint32_t i = f.split_slot[tile] * 16 + sub;
                const uint32_t soff = f.split_offsets[i], sn = f.split_counts[i];
                filter = (uint64_t)soff + sn > (uint64_t)f.split_capacity;
                ids = filter ? f.tile_lists + toff : f.split_lists + soff;
                n = filter ? tn : sn;
            }
        } else {
            if (!f.clear_mask) break;
            uint32_t off = 0, cnt = 1;
            for (; e < ntiles; e += gridDim.x) {
                kz_tile_list(f, e, off, cnt);
                if (cnt == 0) break;
            }
            if (e >= ntiles) break;
            RS = KZ_TILE_SHIFT;
            X0 = (int32_t)((e % f.tiles_x) << KZ_TILE_SHIFT);
            Y0 = (int32_t)((e / f.tiles_x) << KZ_TILE_SHIFT);
            ids = nullptr;
            n = 0;
            e += gridDim.x;
        }
        // (a tile flagged by a cut: its list, then the triangles from the
        // cut on by an in-order scan of their pixel boxes)
        for (uint32_t pass = 0; pass < (cut ? 2u : 1u); ++pass) {
            if (pass == 1) {
                ids = nullptr;
                n = f.total_tris;
                filter = true;
            }
            // (a call per filter mode: each inlined copy is specialised)
            const uint32_t clear = pass == 0 ? f.clear_mask : 0u;
            if (filter)
                kz_raster_rect(f, X0, Y0, RS, ids, n, true, clear, s_color, s_depth, s_claim, s_frag);
            else
                kz_raster_rect(f, X0, Y0, RS, ids, n, false, clear, s_color, s_depth, s_claim, s_frag);
        }
        if (item) {
            uint32_t next = 0;
            if (lane == 0) next = atomicAdd(&f.ctrl[KZ_CTRL_HEADS + x * KZ_CTRL_HEAD_STRIDE], 1u);
            q = x + K * (Gx + (uint32_t)kz_readlane_i((int32_t)next, 0));
        }
    }
}
