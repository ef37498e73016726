"""More pins of the CPU oracle (oracle/raster_oracle.cpp), one per rule that
the first set (test_oracle.py) left open: perspective-correct varyings at
w != 1, the MVP vertex shader, the Phong fragment shader, NEAREST filtering
and REPEAT wrapping, the 4x resolve rounding, cull mode x front face, and
every depth compare op.

Each expected value comes from somewhere other than the oracle's own
formula: exact rationals for barycentric weights, float64 products, closed
forms of the lighting model at special geometry, the texel a pixel centre
lands in, the Vulkan definitions of facing and of each VkCompareOp. Each is
chosen so that a plausible mistake (a dropped 1/w, a transposed matrix, a
missing normalisation, round-to-even instead of round-half-up, a swapped
cull bit, a reversed comparison) fails it.
"""
from fractions import Fraction

import numpy as np
import pytest

from kazan_amd.scenes import generators as g
from kazan_amd.scenes import rng
from scene_kit import px_to_ndc, rect_tris, scene_from_pixels


def _area2(a, b, q):
    return (b[0] - a[0]) * (q[1] - a[1]) - (b[1] - a[1]) * (q[0] - a[0])


def _snapped(tri):
    return [(Fraction(int(round(x * 16)), 16), Fraction(int(round(y * 16)), 16)) for x, y in tri]


# --- perspective-correct varyings -------------------------------------------

@pytest.mark.parametrize("ws", [(1.0, 4.0, 0.25), (0.5, 8.0, 2.0), (16.0, 1.0, 0.125)])
def test_perspective_correct_varyings_at_w_not_one(oracle, ws):
    """Vertex colours red/green/blue with clip w = ws: at a pixel centre
    channel i is (l_i / w_i) / sum_j (l_j / w_j), l the exact screen-space
    barycentrics. Powers of two keep x/w exact, so the snapped vertices are
    known; a dropped 1/w (plain l_i) is off by tens of LSB here."""
    W = H = 64
    tri = [(4.25, 3.5), (60.0, 14.75), (17.5, 61.0)]
    v = np.zeros((3, 8), np.float32)
    for k, ((x, y), w) in enumerate(zip(tri, ws)):
        w = np.float32(w)
        v[k, 0] = px_to_ndc(x, W) * w
        v[k, 1] = px_to_ndc(y, H) * w
        v[k, 2] = np.float32(0.5) * w
        v[k, 3] = w
        v[k, 4 + k] = 1.0
        v[k, 7] = 1.0
    s = g.Scene("persp", W, H, g.VS_POS4_COLOR4, g.FS_COLOR, v)
    s.depth_test = 0
    c, _, _, st = oracle.render(s)
    P = _snapped(tri)
    S = _area2(P[0], P[1], P[2])
    wf = [Fraction(w) for w in ws]
    checked = 0
    worst_linear = 0.0
    for py in range(0, H, 2):
        for px in range(0, W, 2):
            q = (Fraction(2 * px + 1, 2), Fraction(2 * py + 1, 2))
            lam = [_area2(P[1], P[2], q) / S, _area2(P[2], P[0], q) / S, _area2(P[0], P[1], q) / S]
            if min(lam) <= 0:
                continue
            den = sum(l / w for l, w in zip(lam, wf))
            want = np.array([float((l / w) / den) * 255 for l, w in zip(lam, wf)])
            got = c[py, px, 0, :3].astype(float)
            assert np.abs(got - want).max() <= 1.0, (px, py, got, want)
            worst_linear = max(worst_linear, np.abs(np.array([float(l) * 255 for l in lam]) - want).max())
            checked += 1
    assert checked > 100
    # the check can tell perspective-correct from screen-linear interpolation
    assert worst_linear > 20


# --- the MVP vertex shader --------------------------------------------------

def _mvp_scene(tris_px, mvp, W=64, H=64, colors=None):
    tris_px = np.asarray(tris_px, np.float64).reshape(-1, 3, 2)
    n = len(tris_px)
    v = np.zeros((n * 3, 7), np.float32)
    v[:, 0] = px_to_ndc(tris_px[:, :, 0].ravel(), W)
    v[:, 1] = px_to_ndc(tris_px[:, :, 1].ravel(), H)
    v[:, 2] = 0.5
    v[:, 3:7] = 1.0 if colors is None else np.asarray(colors, np.float32).reshape(-1, 4)
    s = g.Scene("mvp", W, H, g.VS_MVP_POS3_COLOR4, g.FS_COLOR, v,
                draws=[dict(count=3 * n, first=0, vertex_offset=0, mvp=np.asarray(mvp, np.float32))])
    return s


def test_identity_mvp_matches_pass_through(oracle):
    r = np.random.default_rng(21)
    tris = np.round(r.uniform(-4, 68, (40, 3, 2)) * 16) / 16
    cols = r.uniform(0, 1, (40, 3, 4)).astype(np.float32)
    zs = r.uniform(0, 1, (40, 3)).astype(np.float32)
    a = scene_from_pixels(tris, colors=cols, z=zs)
    b = _mvp_scene(tris, np.eye(4, dtype=np.float32).ravel(), colors=cols)
    b.vertices[:, 2] = zs.ravel()
    ca, da, _, sa = oracle.render(a)
    cb, db, _, sb = oracle.render(b)
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_array_equal(da.view(np.uint32), db.view(np.uint32))
    assert sa["fragments"] == sb["fragments"] > 0


def test_translation_mvp_shifts_coverage(oracle):
    r = np.random.default_rng(22)
    tris = np.round(r.uniform(8, 40, (20, 3, 2)) * 16) / 16
    eye = np.eye(4, dtype=np.float32)
    t = eye.copy()
    # column-major: the translation is the fourth column, elements 12..14
    t[3, 0] = np.float32(2 * 5 / 64)
    t[3, 1] = np.float32(2 * 3 / 64)
    s0, s1 = _mvp_scene(tris, eye.ravel()), _mvp_scene(tris, t.ravel())
    for s in (s0, s1):
        s.blend = (1, g.BLEND_ONE, g.BLEND_ONE, 0, g.BLEND_ONE, g.BLEND_ZERO, 0)
        s.depth_test = 0
        s.vertices[:, 3:7] = [1.0 / 255, 0, 0, 1]
    c0 = oracle.render(s0)[0][:, :, 0, 0].astype(int)
    c1 = oracle.render(s1)[0][:, :, 0, 0].astype(int)
    assert c0.sum() > 500
    np.testing.assert_array_equal(c1[3:, 5:], c0[:-3, :-5])


def test_mvp_vertex_against_float64_product(oracle):
    """A random non-symmetric perspective MVP (column-major, as Vulkan push
    constants hold it): the shader's clip position agrees with the float64
    product M @ (x, y, z, 1) to float32 rounding; a transposed matrix or a
    row-major read is off by O(1)."""
    r = np.random.default_rng(23)
    M = r.uniform(-2, 2, (4, 4))
    M[3] = [0.1, -0.2, -1.0, 0.3]
    pts = r.uniform(-1, 1, (16, 3)).astype(np.float32)
    v = np.zeros((16, 7), np.float32)
    v[:, 0:3] = pts
    s = g.Scene("mvp", 64, 64, g.VS_MVP_POS3_COLOR4, g.FS_COLOR, v,
                draws=[dict(count=15, first=0, vertex_offset=0, mvp=M.T.astype(np.float32).ravel())])
    Mf = M.astype(np.float32).astype(np.float64)
    for i in range(16):
        pos, var = oracle.vertex(s, 0, i)
        want = Mf @ np.append(pts[i].astype(np.float64), 1.0)
        scale = np.abs(Mf) @ np.append(np.abs(pts[i].astype(np.float64)), 1.0)
        assert (np.abs(pos - want) <= 4 * np.finfo(np.float32).eps * scale).all(), (i, pos, want)
        np.testing.assert_array_equal(var[:4], v[i, 3:7])


# --- the Phong fragment shader ----------------------------------------------

def _phong_scene(lights, ambient):
    """lights: [(pos3, colour3)]; the uniform block {pos[n], colour[n], ambient}."""
    n = len(lights)
    ubo = np.zeros((2 * n + 1, 4), np.float32)
    for i, (p, c) in enumerate(lights):
        ubo[i, :3] = p
        ubo[n + i, :3] = c
    ubo[2 * n, :3] = ambient
    s = g.Scene("phong", 8, 8, g.VS_MVP_POS3_NORMAL3, g.FS_PHONG, np.zeros((3, 6), np.float32))
    s.ubo = ubo.ravel()
    s.num_lights = n
    return s


AMB = (0.05, 0.1, 0.15)
LC = (0.5, 0.25, 0.125)


def test_phong_no_lights_is_ambient(oracle):
    out = oracle.fragment(_phong_scene([], AMB), [0.3, -0.2, 0.1, 0.0, 0.0, 2.0])
    np.testing.assert_array_equal(out, np.array(AMB + (1.0,), np.float32))


@pytest.mark.parametrize("d", [0.5, 1.0, 3.0])
def test_phong_light_along_the_normal(oracle, d):
    """Light at distance d straight along the normal: n.l = 1, the half
    vector is the normal, so diffuse 1 and specular 1, attenuation
    1 / (1 + d^2): c = ambient + colour * 2 / (1 + d^2)."""
    p = np.array([0.25, -0.5, 0.75])
    out = oracle.fragment(_phong_scene([(p + [0, 0, d], LC)], AMB), list(p) + [0.0, 0.0, 3.0])
    want = np.array(AMB) + np.array(LC) * 2.0 / (1.0 + d * d)
    np.testing.assert_allclose(out[:3], want, rtol=2e-6, atol=1e-7)
    assert out[3] == 1.0


@pytest.mark.parametrize("d", [1.0, 2.0])
def test_phong_light_in_the_tangent_plane(oracle, d):
    """Light along +x with normal +z: n.l = 0 (no diffuse); the half vector
    (1, 0, 1)/sqrt(2) gives n.h = 1/sqrt(2), specular (1/sqrt(2))^16 =
    1/256: c = ambient + colour * (1/256) / (1 + d^2)."""
    out = oracle.fragment(_phong_scene([((d, 0.0, 0.0), LC)], AMB), [0, 0, 0, 0, 0, 0.5])
    want = np.array(AMB) + np.array(LC) * (1.0 / 256) / (1.0 + d * d)
    np.testing.assert_allclose(out[:3], want, rtol=1e-5, atol=1e-8)


def test_phong_light_behind_the_surface(oracle):
    """Light below the surface at (1, 0, -1) * d, normal +z: no diffuse;
    the half vector of l = (1, 0, -1)/sqrt(2) and v = +z is
    (1/sqrt(2), 0, 1 - 1/sqrt(2)) normalised, whose z is
    (1 - 1/sqrt(2)) / sqrt(2 - sqrt(2)); specular is that to the 16th."""
    d = 2.0
    out = oracle.fragment(_phong_scene([((d, 0.0, -d), LC)], AMB), [0, 0, 0, 0, 0, 1.0])
    r2 = np.sqrt(2.0)
    nh = (1 - 1 / r2) / np.sqrt(2 - r2)
    want = np.array(AMB) + np.array(LC) * (nh ** 16) / (1.0 + 2 * d * d)
    np.testing.assert_allclose(out[:3], want, rtol=1e-5, atol=1e-9)


def test_phong_lights_add(oracle):
    pos = [0.1, 0.2, 0.3]
    nrm = [0.3, -0.4, 1.2]
    la = ((1.0, -0.5, 2.0), (0.3, 0.2, 0.1))
    lb = ((-1.5, 0.5, 1.0), (0.1, 0.4, 0.2))
    both = oracle.fragment(_phong_scene([la, lb], AMB), pos + nrm)
    a = oracle.fragment(_phong_scene([la], AMB), pos + nrm)
    b = oracle.fragment(_phong_scene([lb], AMB), pos + nrm)
    np.testing.assert_allclose(both[:3], a[:3] + b[:3] - np.array(AMB), rtol=1e-6, atol=1e-7)
    # the interpolated normal is normalised: its length does not matter
    a3 = oracle.fragment(_phong_scene([la], AMB), pos + [3 * x for x in nrm])
    np.testing.assert_allclose(a3, a, rtol=1e-6, atol=1e-7)


# --- NEAREST filtering and REPEAT wrapping ----------------------------------

def _uv_quad_scene(W, H, tex, u0, u1, v0, v1, filt, addr):
    # one full-target quad, uv linear in screen space from (u0, v0) at the
    # top-left corner to (u1, v1) at the bottom-right one
    corners = [(-1, -1, u0, v0), (1, -1, u1, v0), (1, 1, u1, v1), (-1, 1, u0, v1)]
    v = np.array([[x, y, 0.5, u, vv] for x, y, u, vv in corners], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    s = g.Scene("uvquad", W, H, g.VS_POS3_UV2, g.FS_TEXTURE, v, idx)
    s.texture = tex
    s.tex_filter = filt
    s.tex_addr = (addr, addr)
    return s


@pytest.mark.parametrize("u0,u1,addr", [(0.0, 1.0, 2), (-1.0, 2.0, 0), (1.0, 3.0, 0), (-2.5, 0.5, 0),
                                         (-0.5, 1.5, 2)])
def test_nearest_texel_selection(oracle, u0, u1, addr):
    """NEAREST: pixel (px, py) samples texel floor(u * 16) of a 16^2 texture,
    wrapped by REPEAT (mod 16) or clamped (CLAMP_TO_EDGE), with u exact in
    rationals; its colour comes back unchanged (u8 -> /255 -> RGBA8)."""
    W = H = 64
    tex = rng.uniform_u8(31, 16 * 16 * 4).reshape(16, 16, 4)
    s = _uv_quad_scene(W, H, tex, u0, u1, u0, u1, 0, addr)
    c, _, _, _ = oracle.render(s)

    def texel(p):
        u = Fraction(u0) + (Fraction(u1) - Fraction(u0)) * Fraction(2 * p + 1, 2 * W)
        t = u * 16
        assert abs(t - round(t)) >= Fraction(1, 10)  # far from a texel boundary
        i = int(np.floor(float(t)))
        return i % 16 if addr == 0 else min(max(i, 0), 15)

    tx = [texel(p) for p in range(W)]
    want = tex[np.array(tx)[:, None], np.array(tx)[None, :]]  # [py, px]
    np.testing.assert_array_equal(c[:, :, 0, :], want)


def test_bilinear_repeat_wraps(oracle):
    import ctypes
    lib = oracle.lib()
    tex = rng.uniform_u8(33, 8 * 4 * 4).reshape(4, 8, 4)
    out = (ctypes.c_float * 4)()
    for x in range(8):
        want = (tex[2, x].astype(np.float32) * np.float32(256)) * np.float32(1.0 / 65280.0)
        for shift in (1.0, -1.0, -3.0, 2.0):
            lib.kzo_bilinear(tex.ctypes.data, 8, 4, 0, 0, (x + 0.5) / 8 + shift, 2.5 / 4, out)
            np.testing.assert_array_equal(np.array(out, np.float32), want)
    # u = 0 lies halfway between the last and the first texel under REPEAT,
    # and on the first texel's edge under CLAMP_TO_EDGE
    lib.kzo_bilinear(tex.ctypes.data, 8, 4, 0, 0, 0.0, 2.5 / 4, out)
    want = (tex[2, 7].astype(np.int64) * 8 + tex[2, 0].astype(np.int64) * 8) * 16 / 65280
    np.testing.assert_allclose(np.array(out), want, rtol=0, atol=1e-7)
    lib.kzo_bilinear(tex.ctypes.data, 8, 4, 2, 2, 0.0, 2.5 / 4, out)
    want = (tex[2, 0].astype(np.float32) * np.float32(256)) * np.float32(1.0 / 65280.0)
    np.testing.assert_array_equal(np.array(out, np.float32), want)


# --- the 4x resolve ---------------------------------------------------------

def test_resolve4_rounds_half_up(oracle):
    """A rectangle over x in [10.5, 20.5] px covers 2 of the 4 standard
    samples in columns 10 and 20, all 4 between, none outside. The resolve
    is (s0 + s1 + s2 + s3 + 2) >> 2 per channel: with rectangle value r
    and clear value k the half-covered pixels read (2r + 2k + 2) >> 2.
    The channel values make truncation, round-half-even and
    round-half-up disagree."""
    r8 = np.array([1, 5, 200, 3])
    k8 = np.array([0, 0, 51, 0])
    s = scene_from_pixels(rect_tris(10.5, 0, 20.5, 64), colors=[[list(r8 / 255.0)] * 3] * 2, samples=4)
    s.clear_color = tuple(k8 / 255.0)
    c, _, res, _ = oracle.render(s)
    assert (c[5, 10, :, :] == np.where(np.array([0, 1, 0, 1])[:, None] == 1, r8, k8)).all()
    half = (2 * r8 + 2 * k8 + 2) >> 2
    assert list(half) == [1, 3, 126, 2]
    for y in (0, 31, 63):
        assert (res[y, 10] == half).all() and (res[y, 20] == half).all()
        assert (res[y, 15] == r8).all() and (res[y, 3] == k8).all() and (res[y, 25] == k8).all()


# --- cull mode x front face -------------------------------------------------

@pytest.mark.parametrize("front_face", [0, 1])
@pytest.mark.parametrize("cull", [0, 1, 2, 3])
def test_cull_modes_by_signed_area(oracle, cull, front_face):
    """Vulkan's signed area a = -1/2 sum(x_i y_(i+1) - x_(i+1) y_i) in
    framebuffer coordinates; with VK_FRONT_FACE_COUNTER_CLOCKWISE (0) a
    triangle is front-facing iff a > 0, with CLOCKWISE (1) iff a < 0. Cull
    FRONT (1) drops front-facing triangles, BACK (2) back-facing ones,
    FRONT_AND_BACK (3) both. The culled render must equal the render of
    exactly the survivors with culling off."""
    r = np.random.default_rng(40 + 4 * cull + front_face)
    tris = np.round(r.uniform(-2, 66, (60, 3, 2)) * 16) / 16

    def facing_front(t):
        P = _snapped(t)
        a = -sum(P[i][0] * P[(i + 1) % 3][1] - P[(i + 1) % 3][0] * P[i][1] for i in range(3)) / 2
        if a == 0:
            return None
        return (a > 0) if front_face == 0 else (a < 0)

    keep = []
    for t in tris:
        f = facing_front(t)
        if f is None:
            continue
        if (cull & 1 and f) or (cull & 2 and not f):
            continue
        keep.append(t)
    s = scene_from_pixels(tris, count_image=True)
    s.cull_mode, s.front_face = cull, front_face
    c, _, _, st = oracle.render(s)
    if keep:
        ref = scene_from_pixels(np.array(keep), count_image=True)
        rc, _, _, rst = oracle.render(ref)
        np.testing.assert_array_equal(c, rc)
        assert st["fragments"] == rst["fragments"] > 0
    else:
        assert st["fragments"] == 0 and cull == 3
    assert st["triangles_setup"] == len(keep)


def test_mirroring_swaps_front_and_back(oracle):
    r = np.random.default_rng(44)
    tris = np.round(r.uniform(0, 64, (60, 3, 2)) * 16) / 16
    mirrored = tris.copy()
    mirrored[:, :, 1] = 64 - mirrored[:, :, 1]
    a = scene_from_pixels(tris, count_image=True)
    a.cull_mode = 2
    b = scene_from_pixels(mirrored, count_image=True)
    b.cull_mode = 1
    ca, _, _, sa = oracle.render(a)
    cb, _, _, sb = oracle.render(b)
    # the culled sets are mirror images: the same triangles survive, and the
    # images agree row for row (row y of b samples at 64 - (y + 0.5), the
    # centre of row 63 - y of a) except at samples exactly on an edge, where
    # the top-left rule is not mirror symmetric
    assert 0 < sa["triangles_setup"] == sb["triangles_setup"] < len(tris)
    diff = ca[:, :, 0, 0].astype(int) - cb[::-1, :, 0, 0].astype(int)
    assert np.abs(diff).sum() <= 0.02 * ca[:, :, 0, 0].astype(int).sum()


# --- depth compare ops ------------------------------------------------------

COMPARE = {
    0: lambda z, d: False,   # NEVER
    1: lambda z, d: z < d,   # LESS
    2: lambda z, d: z == d,  # EQUAL
    3: lambda z, d: z <= d,  # LESS_OR_EQUAL
    4: lambda z, d: z > d,   # GREATER
    5: lambda z, d: z != d,  # NOT_EQUAL
    6: lambda z, d: z >= d,  # GREATER_OR_EQUAL
    7: lambda z, d: True,    # ALWAYS
}


@pytest.mark.parametrize("op", sorted(COMPARE))
def test_depth_compare_ops(oracle, op):
    """Two full-target rectangles at constant depths za then zb over a depth
    buffer cleared to 0.5, depth writes on: the first passes iff
    op(za, 0.5) and then stores za; the second passes iff op(zb, stored).
    The final colour and depth follow from the VkCompareOp definitions."""
    red, green = [1.0, 0.0, 0.0, 1.0], [0.0, 1.0, 0.0, 1.0]
    for za in (0.25, 0.5, 0.75):
        for zb in (0.25, 0.5, 0.75):
            tris = rect_tris(0, 0, 16, 16) + rect_tris(0, 0, 16, 16)
            s = scene_from_pixels(tris, width=16, height=16, colors=[[red] * 3] * 2 + [[green] * 3] * 2,
                                  z=[[za] * 3] * 2 + [[zb] * 3] * 2)
            s.depth_op = op
            s.clear_depth = 0.5
            s.clear_color = (0.0, 0.0, 1.0, 1.0)
            c, d, _, st = oracle.render(s)
            stored, colour = 0.5, [0, 0, 255, 255]
            if COMPARE[op](za, stored):
                stored, colour = za, [255, 0, 0, 255]
            if COMPARE[op](zb, stored):
                stored, colour = zb, [0, 255, 0, 255]
            assert (d[:, :, 0] == np.float32(stored)).all(), (op, za, zb)
            assert (c[:, :, 0] == colour).all(), (op, za, zb)
