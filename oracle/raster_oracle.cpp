// raster_oracle.cpp — CPU restatement of the Kazan draw path.
//
// TEST INFRASTRUCTURE: the parity oracle and the CPU baseline. The product
// (libkazan_driver.so) never links or calls this file. Build: oracle/Makefile
// (g++ -O2 -ffp-contract=off, no fast-math, SSE f32 arithmetic).
//
// The rules restated here are listed in raster_oracle.h with the reference
// lines that fix them (Kazan's advertised limits in
// vulkan-driver/src/api_impl.rs:1501-1612). The implementation is the plain
// one on purpose: one triangle at a time, every pixel of its bounding box,
// every sample, int64 edge functions — nothing in it mirrors the tiled GPU
// kernels.
#include "raster_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int kMaxVar = 8;

struct Vtx {
    float pos[4];
    float var[kMaxVar];
};

struct Tri {
    int32_t x[3], y[3];
    int64_t S;
    float invS;
    float z0, dz1, dz2;
    float iw[3];
    float a0[kMaxVar], da1[kMaxVar], da2[kMaxVar];
    int32_t px0, px1, py0, py1;
};

const int kSamples1[1][2] = {{8, 8}};
const int kSamples4[4][2] = {{6, 2}, {14, 6}, {2, 10}, {10, 14}};

int num_varyings(uint32_t vs) {
    switch (vs) {
    case KZO_VS_POS4_COLOR4: return 4;
    case KZO_VS_POS3_UV2: return 2;
    case KZO_VS_MVP_POS3_COLOR4: return 4;
    case KZO_VS_MVP_POS3_NORMAL3: return 6;
    }
    return 0;
}

uint32_t stride_floats(uint32_t vs) {
    switch (vs) {
    case KZO_VS_POS4_COLOR4: return 8;
    case KZO_VS_POS3_UV2: return 5;
    case KZO_VS_MVP_POS3_COLOR4: return 7;
    case KZO_VS_MVP_POS3_NORMAL3: return 6;
    }
    return 0;
}

// OpMatrixTimesVector as the HIP backend lowers it: r = M[0]*v.x, then
// r += M[c]*v[c] for c = 1..3, one rounding per operation.
void mat_vec(const float *m, const float v[4], float out[4]) {
    for (int i = 0; i < 4; ++i) {
        float r = m[0 * 4 + i] * v[0];
        r = r + m[1 * 4 + i] * v[1];
        r = r + m[2 * 4 + i] * v[2];
        r = r + m[3 * 4 + i] * v[3];
        out[i] = r;
    }
}

void run_vs(const KzoState &st, const KzoDraw &d, uint32_t vi, Vtx &o) {
    const float *v = d.vertices + (size_t)vi * stride_floats(st.vs_kind);
    switch (st.vs_kind) {
    case KZO_VS_POS4_COLOR4:
        for (int i = 0; i < 4; ++i) o.pos[i] = v[i], o.var[i] = v[4 + i];
        break;
    case KZO_VS_POS3_UV2:
        o.pos[0] = v[0]; o.pos[1] = v[1]; o.pos[2] = v[2]; o.pos[3] = 1.0f;
        o.var[0] = v[3]; o.var[1] = v[4];
        break;
    case KZO_VS_MVP_POS3_COLOR4: {
        float p[4] = {v[0], v[1], v[2], 1.0f};
        mat_vec(d.mvp, p, o.pos);
        for (int i = 0; i < 4; ++i) o.var[i] = v[3 + i];
        break;
    }
    case KZO_VS_MVP_POS3_NORMAL3: {
        float p[4] = {v[0], v[1], v[2], 1.0f};
        mat_vec(d.mvp, p, o.pos);
        for (int i = 0; i < 6; ++i) o.var[i] = v[i];
        break;
    }
    }
}

inline int64_t floor_div16(int64_t v) { return v >> 4; }

// Returns 0 = drop (culled/degenerate/clip-rejected), 1 = set up,
// 2 = needs clipping (R-CLIP: the caller clips it and sets up the pieces).
// `piece`: v is a piece of a clipped triangle; the clip checks are skipped
// and a vertex with w <= 0 or outside the guard band drops the piece.
int setup_triangle(const KzoState &st, const Vtx *v, Tri &t, bool piece = false) {
    if (!piece) {
        // Trivial reject against the clip volume -w<=x,y<=w, 0<=z<=w.
        for (int plane = 0; plane < 6; ++plane) {
            int out = 0;
            for (int k = 0; k < 3; ++k) {
                const float *p = v[k].pos;
                bool o = false;
                switch (plane) {
                case 0: o = p[0] < -p[3]; break;
                case 1: o = p[0] > p[3]; break;
                case 2: o = p[1] < -p[3]; break;
                case 3: o = p[1] > p[3]; break;
                case 4: o = p[2] < 0.0f; break;
                case 5: o = p[2] > p[3]; break;
                }
                out += o;
            }
            if (out == 3) return 0;
        }
        for (int k = 0; k < 3; ++k) {
            const float *p = v[k].pos;
            if (!(p[3] > 0.0f) || !(p[2] >= 0.0f) || !(p[2] <= p[3])) return 2;
        }
    } else {
        for (int k = 0; k < 3; ++k)
            if (!(v[k].pos[3] > 0.0f)) return 0;
    }
    const float hw = st.vp_w * 0.5f, hh = st.vp_h * 0.5f;
    const float ox = st.vp_x + hw, oy = st.vp_y + hh;
    const float dz = st.vp_max_depth - st.vp_min_depth;
    float zf[3];
    for (int k = 0; k < 3; ++k) {
        const float *p = v[k].pos;
        float xd = p[0] / p[3], yd = p[1] / p[3], zd = p[2] / p[3];
        float xf = xd * hw + ox;
        float yf = yd * hh + oy;
        zf[k] = zd * dz + st.vp_min_depth;
        if (!(xf >= -8192.0f && xf <= 8191.0f && yf >= -8192.0f && yf <= 8191.0f)) return piece ? 0 : 2;
        t.x[k] = kzo_snap(xf);
        t.y[k] = kzo_snap(yf);
        t.iw[k] = 1.0f / p[3];
    }
    const int64_t S = (int64_t)(t.x[1] - t.x[0]) * (t.y[2] - t.y[0]) -
                      (int64_t)(t.x[2] - t.x[0]) * (t.y[1] - t.y[0]);
    if (S == 0) return 0;
    // Vulkan's signed area a = -S/2 (y down); CCW front iff a > 0.
    const bool front = (st.front_face == 0) ? (S < 0) : (S > 0);
    if ((st.cull_mode & 1u) && front) return 0;
    if ((st.cull_mode & 2u) && !front) return 0;
    t.S = S;
    t.invS = 1.0f / kzo_i64_to_f32(S);
    t.z0 = zf[0];
    t.dz1 = zf[1] - zf[0];
    t.dz2 = zf[2] - zf[0];
    const int nv = num_varyings(st.vs_kind);
    for (int i = 0; i < nv; ++i) {
        t.a0[i] = v[0].var[i];
        t.da1[i] = v[1].var[i] - v[0].var[i];
        t.da2[i] = v[2].var[i] - v[0].var[i];
    }
    int32_t xmin = std::min({t.x[0], t.x[1], t.x[2]}), xmax = std::max({t.x[0], t.x[1], t.x[2]});
    int32_t ymin = std::min({t.y[0], t.y[1], t.y[2]}), ymax = std::max({t.y[0], t.y[1], t.y[2]});
    int64_t px0 = floor_div16(xmin), px1 = floor_div16(xmax);
    int64_t py0 = floor_div16(ymin), py1 = floor_div16(ymax);
    int64_t cx0 = std::max<int64_t>(0, st.sc_x), cy0 = std::max<int64_t>(0, st.sc_y);
    int64_t cx1 = std::min<int64_t>((int64_t)st.width - 1, (int64_t)st.sc_x + st.sc_w - 1);
    int64_t cy1 = std::min<int64_t>((int64_t)st.height - 1, (int64_t)st.sc_y + st.sc_h - 1);
    t.px0 = (int32_t)std::max(px0, cx0);
    t.px1 = (int32_t)std::min(px1, cx1);
    t.py0 = (int32_t)std::max(py0, cy0);
    t.py1 = (int32_t)std::min(py1, cy1);
    return 1;
}

// R-CLIP: Sutherland-Hodgman clipping of a triangle that needs it (see
// setup_triangle) against z >= 0, z <= w and the guard band's four planes
// x = gx * w, y = gy * w (screen x, y = -8000 and 8000 px, inside
// viewportBoundsRange [-8192, 8191], api_impl.rs:1568). A polygon vertex is
// a vertex record: clip position and varyings. Per plane the vertices are
// visited in order: an inside vertex (d >= 0) is kept, and an edge whose
// ends are on different sides adds its intersection, always interpolated
// from the inside end, every component of the record alike:
// t = d_in / (d_in - d_out), x = x_in + t * (x_out - x_in) (so the shared
// edge of two triangles gets the same point from both). At most kClipMax
// vertices are kept; a plane that leaves fewer than 3 drops the triangle.
// The polygon is drawn as the fan (0, j + 1, j + 2), j = 0 .. n - 3, each
// piece set up like a triangle, in place of the original in primitive
// order.
constexpr int kClipMax = 12;

struct ClipBand {
    float xl, xh, yl, yh;
};

ClipBand clip_band(const KzoState &st) {
    const float hw = st.vp_w * 0.5f, hh = st.vp_h * 0.5f;
    const float ox = st.vp_x + hw, oy = st.vp_y + hh;
    const float ax = (-8000.0f - ox) / hw, bx = (8000.0f - ox) / hw;
    const float ay = (-8000.0f - oy) / hh, by = (8000.0f - oy) / hh;
    return {std::fmin(ax, bx), std::fmax(ax, bx), std::fmin(ay, by), std::fmax(ay, by)};
}

inline float clip_distance(int plane, const ClipBand &g, const float P[4]) {
    switch (plane) {
    case 0: return P[2];
    case 1: return P[3] - P[2];
    case 2: return P[0] - g.xl * P[3];
    case 3: return g.xh * P[3] - P[0];
    case 4: return P[1] - g.yl * P[3];
    default: return g.yh * P[3] - P[1];
    }
}

inline float clip_lerp(float in, float out, float t) { return in + t * (out - in); }

// The clipped polygon's vertex records; returns its vertex count (0: nothing left).
int clip_polygon(const KzoState &st, const Vtx *v, Vtx poly[kClipMax]) {
    const ClipBand g = clip_band(st);
    Vtx a[kClipMax], b[kClipMax];
    for (int k = 0; k < 3; ++k) a[k] = v[k];
    int n = 3;
    for (int plane = 0; plane < 6; ++plane) {
        int m = 0;
        for (int i = 0; i < n; ++i) {
            const int j = i + 1 == n ? 0 : i + 1;
            const float di = clip_distance(plane, g, a[i].pos), dj = clip_distance(plane, g, a[j].pos);
            const bool ini = di >= 0.0f, inj = dj >= 0.0f;
            if (ini && m < kClipMax) b[m++] = a[i];
            if (ini != inj && m < kClipMax) {
                const Vtx &vin = ini ? a[i] : a[j], &vout = ini ? a[j] : a[i];
                const float din = ini ? di : dj, dout = ini ? dj : di;
                const float tt = din / (din - dout);
                for (int c = 0; c < 4; ++c) b[m].pos[c] = clip_lerp(vin.pos[c], vout.pos[c], tt);
                for (int c = 0; c < kMaxVar; ++c) b[m].var[c] = clip_lerp(vin.var[c], vout.var[c], tt);
                ++m;
            }
        }
        if (m < 3) return 0;
        for (int i = 0; i < m; ++i) a[i] = b[i];
        n = m;
    }
    for (int i = 0; i < n; ++i) poly[i] = a[i];
    return n;
}

// The fan pieces of a triangle that needs clipping that set up (R-CLIP).
int clip_pieces(const KzoState &st, const Vtx *v, Tri *out) {
    Vtx poly[kClipMax];
    const int n = clip_polygon(st, v, poly);
    int k = 0;
    for (int j = 0; j + 2 < n; ++j) {
        const Vtx p[3] = {poly[0], poly[j + 1], poly[j + 2]};
        if (setup_triangle(st, p, out[k], true) == 1) ++k;
    }
    return k;
}

struct Edges {
    int64_t e[3];  // unflipped E12, E20, E01 (weights of v0, v1, v2)
};

inline void eval_edges(const Tri &t, int64_t px, int64_t py, Edges &o) {
    // E_ab(p) = (xb - xa)*(py - ya) - (yb - ya)*(px - xa)
    o.e[0] = (int64_t)(t.x[2] - t.x[1]) * (py - t.y[1]) - (int64_t)(t.y[2] - t.y[1]) * (px - t.x[1]);
    o.e[1] = (int64_t)(t.x[0] - t.x[2]) * (py - t.y[2]) - (int64_t)(t.y[0] - t.y[2]) * (px - t.x[2]);
    o.e[2] = (int64_t)(t.x[1] - t.x[0]) * (py - t.y[0]) - (int64_t)(t.y[1] - t.y[0]) * (px - t.x[0]);
}

inline bool is_top_left(int64_t A, int64_t B) { return A > 0 || (A == 0 && B > 0); }

inline bool covered(const Tri &t, const Edges &E) {
    const int64_t sgn = t.S > 0 ? 1 : -1;
    // Coefficients of E12, E20, E01: A = -(yb - ya), B = (xb - xa).
    const int64_t A[3] = {-(int64_t)(t.y[2] - t.y[1]), -(int64_t)(t.y[0] - t.y[2]), -(int64_t)(t.y[1] - t.y[0])};
    const int64_t B[3] = {(int64_t)(t.x[2] - t.x[1]), (int64_t)(t.x[0] - t.x[2]), (int64_t)(t.x[1] - t.x[0])};
    for (int i = 0; i < 3; ++i) {
        const int64_t e = sgn * E.e[i];
        if (e > 0) continue;
        if (e == 0 && is_top_left(sgn * A[i], sgn * B[i])) continue;
        return false;
    }
    return true;
}

inline float clamp01(float v) { return std::fmin(std::fmax(v, 0.0f), 1.0f); }

bool depth_compare(uint32_t op, float z, float d) {
    switch (op) {
    case 0: return false;
    case 1: return z < d;
    case 2: return z == d;
    case 3: return z <= d;
    case 4: return z > d;
    case 5: return z != d;
    case 6: return z >= d;
    default: return true;
    }
}

float blend_factor(uint32_t f, const float s[4], const float d[4], int c) {
    switch (f) {
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
    case 14: return c == 3 ? 1.0f : std::fmin(s[3], 1.0f - d[3]);
    default: return 0.0f;  // constant factors are not used by the draw path
    }
}

float blend_op(uint32_t op, float s, float sf, float d, float df) {
    switch (op) {
    case 0: return s * sf + d * df;
    case 1: return s * sf - d * df;
    case 2: return d * df - s * sf;
    case 3: return std::fmin(s, d);
    default: return std::fmax(s, d);
    }
}

inline float dot3(const float *a, const float *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

uint32_t wrap_coord(int64_t i, uint32_t n, uint32_t mode) {
    if (mode == 0) {  // REPEAT
        int64_t m = i % (int64_t)n;
        return (uint32_t)(m < 0 ? m + n : m);
    }
    if (i < 0) return 0;  // CLAMP_TO_EDGE
    if (i >= (int64_t)n) return n - 1;
    return (uint32_t)i;
}

void fs_phong(const KzoState &st, const float *var, float out[4]) {
    const float *vp = var, *vn = var + 3;
    const float rn = 1.0f / std::sqrt(dot3(vn, vn));
    const float n[3] = {vn[0] * rn, vn[1] * rn, vn[2] * rn};
    const uint32_t nl = st.num_lights;
    const float *amb = st.ubo + 8 * nl;
    float c[3] = {amb[0], amb[1], amb[2]};
    for (uint32_t i = 0; i < nl; ++i) {
        const float *lp = st.ubo + 4 * i, *lc = st.ubo + 4 * (nl + i);
        const float L[3] = {lp[0] - vp[0], lp[1] - vp[1], lp[2] - vp[2]};
        const float d2 = dot3(L, L);
        const float rl = 1.0f / std::sqrt(d2);
        const float l[3] = {L[0] * rl, L[1] * rl, L[2] * rl};
        const float ndl = std::fmax(dot3(n, l), 0.0f);
        const float hv[3] = {l[0], l[1], l[2] + 1.0f};
        const float rh = 1.0f / std::sqrt(dot3(hv, hv));
        const float h[3] = {hv[0] * rh, hv[1] * rh, hv[2] * rh};
        float s = std::fmax(dot3(n, h), 0.0f);
        s = s * s; s = s * s; s = s * s; s = s * s;
        const float att = 1.0f / (1.0f + d2);
        const float k = (ndl + s) * att;
        c[0] = c[0] + lc[0] * k;
        c[1] = c[1] + lc[1] * k;
        c[2] = c[2] + lc[2] * k;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = 1.0f;
}

// Returns false when the fragment shader discards the fragment.
bool run_fs(const KzoState &st, const float *var, float out[4]) {
    switch (st.fs_kind) {
    case KZO_FS_COLOR:
        for (int i = 0; i < 4; ++i) out[i] = var[i];
        break;
    case KZO_FS_ALPHA_TEST:
        for (int i = 0; i < 4; ++i) out[i] = var[i];
        return !(var[3] < 0.5f);
    case KZO_FS_TEXTURE:
        if (st.tex_filter == 0) {
            // NEAREST: texel floor(u * W), floor(v * H), same addressing
            auto coord = [](float c) {
                c = std::fmin(std::fmax(c, -1.0e8f), 1.0e8f);
                return c != c ? 0.0f : c;
            };
            const int64_t x = (int64_t)std::floor(coord(var[0] * (float)st.tex_w));
            const int64_t y = (int64_t)std::floor(coord(var[1] * (float)st.tex_h));
            const uint8_t *t = st.tex + ((size_t)wrap_coord(y, st.tex_h, st.tex_addr_v) * st.tex_w +
                                         wrap_coord(x, st.tex_w, st.tex_addr_u)) * 4;
            for (int c = 0; c < 4; ++c) out[c] = (float)t[c] * (1.0f / 255.0f);
        } else {
            kzo_bilinear(st.tex, st.tex_w, st.tex_h, st.tex_addr_u, st.tex_addr_v, var[0], var[1], out);
        }
        break;
    case KZO_FS_PHONG:
        fs_phong(st, var, out);
        break;
    default:
        out[0] = out[1] = out[2] = out[3] = 0.0f;
    }
    return true;
}

bool fs_can_discard(uint32_t fs) { return fs == KZO_FS_ALPHA_TEST; }

struct BandStats {
    uint64_t fragments = 0, depth_passed = 0;
};

void raster_band(const KzoState &st, const std::vector<Tri> &tris, int row0, int row1,
                 uint8_t *color, float *depth, BandStats &bs) {
    const uint32_t ns = st.samples;
    const int(*sp)[2] = ns == 4 ? kSamples4 : kSamples1;
    const int nv = num_varyings(st.vs_kind);
    const float zlo = std::fmin(st.vp_min_depth, st.vp_max_depth);
    const float zhi = std::fmax(st.vp_min_depth, st.vp_max_depth);
    for (const Tri &t : tris) {
        const int y0 = std::max(t.py0, row0), y1 = std::min(t.py1, row1 - 1);
        if (y0 > y1 || t.px0 > t.px1) continue;
        for (int py = y0; py <= y1; ++py) {
            for (int px = t.px0; px <= t.px1; ++px) {
                uint32_t mask = 0;
                float zs[4];
                for (uint32_t s = 0; s < ns; ++s) {
                    Edges E;
                    eval_edges(t, (int64_t)px * 16 + sp[s][0], (int64_t)py * 16 + sp[s][1], E);
                    if (!covered(t, E)) continue;
                    mask |= 1u << s;
                    const float l1 = kzo_i64_to_f32(E.e[1]) * t.invS;
                    const float l2 = kzo_i64_to_f32(E.e[2]) * t.invS;
                    float z = (t.z0 + l1 * t.dz1) + l2 * t.dz2;
                    zs[s] = std::fmin(std::fmax(z, zlo), zhi);
                }
                if (!mask) continue;
                bs.fragments += __builtin_popcount(mask);
                const size_t pix = (size_t)py * st.width + px;
                // varyings and the fragment shader at the pixel centre: before
                // the depth test when the shader can discard (R-KILL), else
                // only for fragments that pass it (the same result)
                float src[4];
                auto shade = [&]() {
                    Edges E;
                    eval_edges(t, (int64_t)px * 16 + 8, (int64_t)py * 16 + 8, E);
                    const float l0 = kzo_i64_to_f32(E.e[0]) * t.invS;
                    const float l1 = kzo_i64_to_f32(E.e[1]) * t.invS;
                    const float l2 = kzo_i64_to_f32(E.e[2]) * t.invS;
                    const float q = ((l0 * t.iw[0]) + (l1 * t.iw[1])) + (l2 * t.iw[2]);
                    const float rq = 1.0f / q;
                    const float b1 = (l1 * t.iw[1]) * rq;
                    const float b2 = (l2 * t.iw[2]) * rq;
                    float var[kMaxVar];
                    for (int i = 0; i < nv; ++i) var[i] = (t.a0[i] + b1 * t.da1[i]) + b2 * t.da2[i];
                    const bool keep = run_fs(st, var, src);
                    for (int c = 0; c < 4; ++c) src[c] = clamp01(src[c]);
                    return keep;
                };
                const bool late = fs_can_discard(st.fs_kind);
                if (late && !shade()) continue;
                uint32_t pass = 0;
                for (uint32_t s = 0; s < ns; ++s) {
                    if (!(mask >> s & 1u)) continue;
                    float &d = depth[pix * ns + s];
                    if (st.depth_test && !depth_compare(st.depth_compare, zs[s], d)) continue;
                    pass |= 1u << s;
                    if (st.depth_test && st.depth_write) d = zs[s];
                }
                if (!pass) continue;
                bs.depth_passed += __builtin_popcount(pass);
                if (!late) shade();
                for (uint32_t s = 0; s < ns; ++s) {
                    if (!(pass >> s & 1u)) continue;
                    uint8_t *cp = color + (pix * ns + s) * 4;
                    float res[4];
                    if (st.blend_enable) {
                        float dst[4];
                        for (int c = 0; c < 4; ++c) dst[c] = (float)cp[c] * (1.0f / 255.0f);
                        for (int c = 0; c < 3; ++c)
                            res[c] = blend_op(st.color_op, src[c], blend_factor(st.src_color, src, dst, c), dst[c],
                                              blend_factor(st.dst_color, src, dst, c));
                        res[3] = blend_op(st.alpha_op, src[3], blend_factor(st.src_alpha, src, dst, 3), dst[3],
                                          blend_factor(st.dst_alpha, src, dst, 3));
                    } else {
                        for (int c = 0; c < 4; ++c) res[c] = src[c];
                    }
                    for (int c = 0; c < 4; ++c)
                        if (st.write_mask >> c & 1u) cp[c] = (uint8_t)kzo_encode_unorm8(res[c]);
                }
            }
        }
    }
}

}  // namespace

extern "C" {

int32_t kzo_snap(float xf) { return (int32_t)std::rint(xf * 16.0f); }

float kzo_i64_to_f32(int64_t v) { return (float)v; }

uint32_t kzo_encode_unorm8(float v) { return (uint32_t)(clamp01(v) * 255.0f + 0.5f); }

void kzo_bilinear(const uint8_t *tex, uint32_t w, uint32_t h, uint32_t addr_u, uint32_t addr_v, float u, float v,
                  float out[4]) {
    // f = u * (W*16) - 8 in 1/16 texel units relative to texel centres.
    float fu = u * (float)(w * 16) - 8.0f;
    float fv = v * (float)(h * 16) - 8.0f;
    // Bound the range so the integer conversion is defined; any value this
    // far outside the texture addresses the same texels.
    const float lim = 1.0e8f;
    fu = std::fmin(std::fmax(fu, -lim), lim);
    fv = std::fmin(std::fmax(fv, -lim), lim);
    if (fu != fu) fu = 0.0f;
    if (fv != fv) fv = 0.0f;
    const int64_t iu = (int64_t)std::floor(fu), iv = (int64_t)std::floor(fv);
    const int64_t x0 = iu >> 4, y0 = iv >> 4;
    const uint32_t ax = (uint32_t)(iu & 15), ay = (uint32_t)(iv & 15);
    const uint32_t xa = wrap_coord(x0, w, addr_u), xb = wrap_coord(x0 + 1, w, addr_u);
    const uint32_t ya = wrap_coord(y0, h, addr_v), yb = wrap_coord(y0 + 1, h, addr_v);
    const uint8_t *t00 = tex + ((size_t)ya * w + xa) * 4, *t10 = tex + ((size_t)ya * w + xb) * 4;
    const uint8_t *t01 = tex + ((size_t)yb * w + xa) * 4, *t11 = tex + ((size_t)yb * w + xb) * 4;
    for (int c = 0; c < 4; ++c) {
        const uint32_t top = t00[c] * (16 - ax) + t10[c] * ax;
        const uint32_t bot = t01[c] * (16 - ax) + t11[c] * ax;
        const uint32_t val = top * (16 - ay) + bot * ay;
        out[c] = (float)val * (1.0f / 65280.0f);
    }
}

void kzo_vertex(const KzoState *st, const KzoDraw *d, uint32_t vi, float pos[4], float var[8]) {
    Vtx v{};
    run_vs(*st, *d, vi, v);
    for (int i = 0; i < 4; ++i) pos[i] = v.pos[i];
    for (int i = 0; i < kMaxVar; ++i) var[i] = v.var[i];
}

int kzo_fragment(const KzoState *st, const float *var, float out[4]) { return run_fs(*st, var, out) ? 1 : 0; }

int kzo_clip_polygon(const KzoState *st, const float *pos, float *poly) {
    Vtx v[3]{};
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 4; ++c) v[k].pos[c] = pos[4 * k + c];
    Vtx p[kClipMax];
    const int n = clip_polygon(*st, v, p);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 4; ++c) poly[4 * i + c] = p[i].pos[c];
    return n;
}

int kzo_render(const KzoState *stp, const KzoDraw *draws, uint32_t ndraws, uint8_t *color, float *depth,
               uint8_t *resolved, KzoStats *stats, int threads) {
    const KzoState &st = *stp;
    if (st.samples != 1 && st.samples != 4) return -1;
    if (num_varyings(st.vs_kind) == 0 && st.vs_kind != KZO_VS_POS4_COLOR4) return -1;
    const size_t npix = (size_t)st.width * st.height * st.samples;
    if (st.clear_color_enable) {
        uint8_t c[4];
        for (int i = 0; i < 4; ++i) c[i] = (uint8_t)kzo_encode_unorm8(st.clear_color[i]);
        for (size_t i = 0; i < npix; ++i) std::memcpy(color + i * 4, c, 4);
    }
    if (st.clear_depth_enable)
        for (size_t i = 0; i < npix; ++i) depth[i] = st.clear_depth;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());

    // Primitive assembly + vertex shader + setup, in primitive order.
    size_t total = 0;
    for (uint32_t d = 0; d < ndraws; ++d) total += draws[d].count / 3;
    std::vector<Tri> all(total);
    std::vector<uint8_t> status(total);
    std::vector<std::pair<uint32_t, uint32_t>> where(total);
    {
        size_t k = 0;
        for (uint32_t d = 0; d < ndraws; ++d)
            for (uint32_t t = 0; t < draws[d].count / 3; ++t) where[k++] = {d, t};
    }
    // the pieces of clipped triangles, per setup thread, in primitive order
    std::vector<std::vector<std::pair<size_t, std::vector<Tri>>>> clipped(threads);
    auto setup_range = [&](size_t a, size_t b, int tid) {
        for (size_t k = a; k < b; ++k) {
            const KzoDraw &d = draws[where[k].first];
            const uint32_t t = where[k].second;
            Vtx v[3];
            for (int j = 0; j < 3; ++j) {
                uint32_t vi;
                if (d.indices)
                    vi = (uint32_t)((int64_t)d.indices[d.first + 3 * t + j] + d.vertex_offset);
                else
                    vi = d.first + 3 * t + j;
                run_vs(st, d, vi, v[j]);
            }
            status[k] = (uint8_t)setup_triangle(st, v, all[k]);
            if (status[k] == 2) {
                Tri pieces[kClipMax];
                const int np = clip_pieces(st, v, pieces);
                clipped[tid].emplace_back(k, std::vector<Tri>(pieces, pieces + np));
            }
        }
    };
    {
        std::vector<std::thread> pool;
        const size_t chunk = (total + threads - 1) / threads;
        for (int i = 0; i < threads; ++i) {
            size_t a = i * chunk, b = std::min(total, a + chunk);
            if (a < b) pool.emplace_back(setup_range, a, b, i);
        }
        for (auto &th : pool) th.join();
    }
    std::vector<Tri> tris;
    tris.reserve(total);
    uint64_t n_clipped = 0, n_setup = 0;
    {
        int ci = 0;
        size_t cj = 0;
        for (size_t k = 0; k < total; ++k) {
            if (status[k] == 1) {
                tris.push_back(all[k]);
                ++n_setup;
            } else if (status[k] == 2) {
                while (cj >= clipped[ci].size()) ++ci, cj = 0;
                const auto &pc = clipped[ci][cj++];
                ++n_clipped;
                n_setup += pc.second.empty() ? 0 : 1;
                for (const Tri &q : pc.second) tris.push_back(q);
            }
        }
    }
    all.clear();
    all.shrink_to_fit();

    // Raster in horizontal bands; each band walks all triangles in order.
    const int H = (int)st.height;
    const int nb = std::max(1, std::min(threads, H));
    std::vector<BandStats> bstats(nb);
    {
        std::vector<std::thread> pool;
        for (int b = 0; b < nb; ++b) {
            int r0 = (int)((int64_t)H * b / nb), r1 = (int)((int64_t)H * (b + 1) / nb);
            pool.emplace_back([&, b, r0, r1] { raster_band(st, tris, r0, r1, color, depth, bstats[b]); });
        }
        for (auto &th : pool) th.join();
    }
    if (resolved && st.samples == 4) {
        const size_t n1 = (size_t)st.width * st.height;
        for (size_t p = 0; p < n1; ++p)
            for (int c = 0; c < 4; ++c) {
                uint32_t s = 2;
                for (int k = 0; k < 4; ++k) s += color[(p * 4 + k) * 4 + c];
                resolved[p * 4 + c] = (uint8_t)(s >> 2);
            }
    } else if (resolved) {
        std::memcpy(resolved, color, (size_t)st.width * st.height * 4);
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->triangles_in = total;
        stats->triangles_setup = n_setup;
        stats->triangles_clipped = n_clipped;
        for (auto &b : bstats) stats->fragments += b.fragments, stats->depth_passed += b.depth_passed;
    }
    return 0;
}

}  // extern "C"
