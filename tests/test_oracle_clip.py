"""Pins of the oracle's clipping rule R-CLIP (oracle/raster_oracle.cpp,
clip_polygon and clip_pieces): triangles that cross the near or far plane,
reach behind the eye, or leave the guard band.

The expected values come from outside the clipper:
- a closed form: a quad whose clip-space z is 0.5 * y * w is cut by z = 0
  exactly along the screen's middle row boundary, so the covered rows and
  the depth (0.5 * y_ndc) are known;
- clipping must not change attribute interpolation: the same quad moved to
  z >= 0 (no clipping) gives the same colours where both cover;
- brute force: for each pixel centre, the point of the (unclipped)
  triangle's plane that projects there, solved in float64 from the clip
  positions (homogeneous barycentrics); it is visible iff inside the
  triangle with 0 <= z <= w, its depth is z / w and its colour the
  barycentric mix of the vertex colours;
- the guard band: a quad with vertices a thousand viewports away covers
  every pixel exactly once with the colours of its linear function;
- watertightness: a mesh cut by the near plane covers every visible pixel
  exactly once.
A wrong plane, a flipped distance sign, interpolation from the wrong end,
a dropped 1/w in the pieces or a fan in the wrong order fails one of them.
"""
import numpy as np
import pytest

from kazan_amd.scenes import generators as g

W = H = 64
COMPARE_ALWAYS = 7


def _scene(v, w=W, h=H, name="clip"):
    s = g.Scene(name, w, h, g.VS_POS4_COLOR4, g.FS_COLOR, np.asarray(v, np.float32))
    s.depth_op = COMPARE_ALWAYS
    return s


def _quad(z_of_y, ws=(1.0, 2.0, 4.0, 1.0), colors=None):
    """Two triangles over NDC [-1, 1]^2; corner k = (X, Y) with clip w ws[k],
    clip z = z_of_y(Y) * w. Corners: (-1,-1), (1,-1), (1,1), (-1,1)."""
    corners = [(-1.0, -1.0), (1.0, -1.0), (1.0, 1.0), (-1.0, 1.0)]
    if colors is None:
        colors = np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1], [1, 1, 0, 1]], np.float32)
    rows = []
    for tri in ((0, 1, 2), (0, 2, 3)):
        for k in tri:
            X, Y = corners[k]
            w = np.float32(ws[k])
            rows.append([X * w, Y * w, np.float32(z_of_y(Y)) * w, w, *colors[k]])
    return np.array(rows, np.float32)


def test_near_plane_cut_closed_form(oracle):
    """z = 0.5 * Y * w: z < 0 on the upper half (Y < 0, y down), so exactly
    rows 32..63 remain, each pixel once, with depth 0.5 * Y at its centre."""
    s = _scene(_quad(lambda Y: 0.5 * Y))
    c, d, _, st = oracle.render(s)
    cov = d[:, :, 0] != np.float32(1.0)
    assert st["triangles_clipped"] == 2 and st["triangles_setup"] == 2
    assert not cov[:32].any() and cov[32:].all()
    assert st["fragments"] == 32 * W
    Y = (np.arange(32, 64) + 0.5) / 32.0 - 1.0
    want = np.repeat((0.5 * Y)[:, None], W, axis=1)
    assert np.abs(d[32:, :, 0] - want).max() < 1e-3


def test_far_plane_cut_closed_form(oracle):
    """z = (0.75 + 0.5 Y) * w: z > w below Y = 0.5 (screen row 48), so rows
    0..47 remain."""
    s = _scene(_quad(lambda Y: 0.75 + 0.5 * Y))
    _, d, _, st = oracle.render(s)
    cov = d[:, :, 0] != np.float32(1.0)
    assert cov[:48].all() and not cov[48:].any()
    assert st["fragments"] == 48 * W and st["triangles_clipped"] == 2


def test_clipping_keeps_attribute_interpolation(oracle):
    """The same quad with z >= 0 everywhere (no clipping) colours the lower
    half as the clipped one does: the pieces interpolate the original
    triangle's varyings perspective-correctly (w = 1, 2, 4, 1 per corner)."""
    clipped = _scene(_quad(lambda Y: 0.5 * Y))
    whole = _scene(_quad(lambda Y: 0.25 * Y + 0.5))
    c1, _, _, st1 = oracle.render(clipped)
    c2, _, _, st2 = oracle.render(whole)
    assert st1["triangles_clipped"] == 2 and st2["triangles_clipped"] == 0
    diff = np.abs(c1[32:].astype(int) - c2[32:].astype(int))
    assert diff.max() <= 1
    # and the colours really vary (perspective), so the comparison is not vacuous
    assert len(np.unique(c2[32:, :, 0, 0])) > 50


def _classify(P, X, Y):
    """(visible, lambda, clip point) of the unclipped triangle at NDC (X, Y),
    in float64; visible is None when the point is degenerate."""
    M = np.array([P[:, 0] - X * P[:, 3], P[:, 1] - Y * P[:, 3], np.ones(3)])
    if abs(np.linalg.det(M)) < 1e-12:
        return None, None, None
    lam = np.linalg.solve(M, [0.0, 0.0, 1.0])
    q = lam @ P
    return bool((lam > 0).all() and q[3] > 0 and 0 <= q[2] <= q[3]), lam, q


def _boundary_lines(P):
    """The screen lines (homogeneous, NDC) that bound the visible part of
    the triangle: its three edges and the traces of the planes z = 0 and
    z = w on it. Each is {lambda . c = 0, sum(lambda) = 1} mapped through
    lambda -> (x, y, w)(lambda), so vertices behind the eye need no care."""
    lines = []
    for c in (np.eye(3)[0], np.eye(3)[1], np.eye(3)[2], P[:, 2], P[:, 3] - P[:, 2]):
        A = np.array([c, np.ones(3)])
        d = np.cross(A[0], A[1])
        if np.linalg.norm(d) < 1e-12:
            continue
        lp = np.linalg.lstsq(A, [0.0, 1.0], rcond=None)[0]
        h1 = lp @ P[:, [0, 1, 3]]
        h2 = (lp + d) @ P[:, [0, 1, 3]]
        L = np.cross(h1, h2)
        if np.hypot(L[0], L[1]) > 1e-12:
            lines.append(L / np.hypot(L[0], L[1]))
    return lines


def _brute_force(P, colors, w, h, delta_px=0.1):
    """Per pixel centre: (visible, depth z/w, colour) from the float64
    homogeneous barycentrics of the unclipped triangle, plus an 'ambiguous'
    mask for centres within delta_px of a boundary line (snapping to
    1/16 px moves an edge by at most 1/32 px)."""
    P = np.asarray(P, np.float64)
    lines = _boundary_lines(P)
    vis = np.zeros((h, w), bool)
    amb = np.zeros((h, w), bool)
    dep = np.zeros((h, w))
    col = np.zeros((h, w, 4))
    grad = np.zeros((h, w))  # colour change per pixel (the largest channel's, either axis)
    half = 1.0 / w           # half a pixel in NDC (w == h)
    for py in range(h):
        Y = (py + 0.5) * 2.0 / h - 1.0
        for px in range(w):
            X = (px + 0.5) * 2.0 / w - 1.0
            v, lam, q = _classify(P, X, Y)
            # distance to each line in pixels (x and y scale alike here: w == h)
            near = any(abs(L @ [X, Y, 1.0]) * w / 2.0 < delta_px for L in lines)
            if v is None or near:
                amb[py, px] = True
                continue
            if v:
                vis[py, px] = True
                dep[py, px] = q[2] / q[3]
                col[py, px] = lam @ colors
                for ox, oy in ((half, 0.0), (0.0, half)):
                    _, l1, _ = _classify(P, X + ox, Y + oy)
                    _, l2, _ = _classify(P, X - ox, Y - oy)
                    if l1 is not None and l2 is not None:
                        grad[py, px] = max(grad[py, px], float(np.abs((l1 - l2) @ colors).max()))
    return vis, amb, dep, col, grad


def _camera_triangles(n, seed):
    """Triangles in camera space around and behind a 60-degree camera at
    the origin (near 0.5, far 20): many cross the near plane, some have a
    vertex behind the eye (w < 0) or beyond the far plane."""
    r = np.random.default_rng(seed)
    proj = g._perspective(60.0, 1.0, 0.5, 20.0)
    out = []
    while len(out) < n:
        c = np.array([r.uniform(-1.5, 1.5), r.uniform(-1.5, 1.5), r.uniform(-4.0, 0.5)])
        tri = c + r.uniform(-3.0, 3.0, (3, 3))
        clip = (proj @ np.c_[tri, np.ones(3)].T).T.astype(np.float32)
        out.append(clip)
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_clipped_triangles_against_brute_force(oracle, seed):
    r = np.random.default_rng(100 + seed)
    n_checked = n_cut = 0
    worst_col = worst_dep = 0.0
    for P in _camera_triangles(80, seed):
        colors = r.uniform(0, 1, (3, 4)).astype(np.float32)
        colors[:, 3] = 1.0
        v = np.c_[P, colors]
        s = _scene(v, 32, 32)
        c, d, _, st = oracle.render(s)
        vis, amb, dep, col, grad = _brute_force(P, colors.astype(np.float64), 32, 32)
        cov = d[:, :, 0] != np.float32(1.0)
        sure = ~amb
        assert (cov[sure] == vis[sure]).all(), (seed, P)
        n_checked += int(sure.sum())
        n_cut += st["triangles_clipped"]
        m = vis & sure
        if m.any():
            worst_dep = max(worst_dep, float(np.abs(d[m, 0] - dep[m]).max()))
            # vertices snap to 1/16 px, so a colour may move by up to
            # 1/16 px of its gradient beyond the 1 LSB of rounding
            err = np.abs(c[m, 0].astype(np.float64) - col[m] * 255.0).max(axis=1)
            worst_col = max(worst_col, float((err - 255.0 * grad[m] / 16.0).max()))
    print("clipped %d of 80, %d pixels checked, worst depth %.2e, worst colour excess %.2f LSB" %
          (n_cut, n_checked, worst_dep, worst_col))
    assert n_cut >= 20, n_cut  # the set really exercises the clipper
    assert n_checked > 40000
    assert worst_dep < 2e-2, worst_dep
    assert worst_col <= 1.0, worst_col


def test_guard_band_quad_covers_every_pixel_once(oracle):
    """A quad with corners at NDC +-1000 (64,000 px outside the viewport)
    and colours R = 0.5 + 0.25 X, G = 0.5 + 0.25 Y at its corners: every
    pixel is covered once and shows that linear function."""
    corners = [(-1000.0, -1000.0), (1000.0, -1000.0), (1000.0, 1000.0), (-1000.0, 1000.0)]
    rows = []
    for tri in ((0, 1, 2), (0, 2, 3)):
        for k in tri:
            X, Y = corners[k]
            rows.append([X, Y, 0.5, 1.0, 0.5 + 0.25 * X, 0.5 + 0.25 * Y, 0.0, 1.0])
    s = _scene(np.array(rows, np.float32))
    c, d, _, st = oracle.render(s)
    assert st["triangles_clipped"] == 2 and st["fragments"] == W * H
    assert (d[:, :, 0] == np.float32(0.5)).all()
    X = (np.arange(W) + 0.5) / 32.0 - 1.0
    Y = (np.arange(H) + 0.5) / 32.0 - 1.0
    want_r = np.clip(0.5 + 0.25 * X, 0, 1) * 255.0
    want_g = np.clip(0.5 + 0.25 * Y, 0, 1) * 255.0
    assert np.abs(c[:, :, 0, 0] - want_r[None, :]).max() <= 1.0
    assert np.abs(c[:, :, 0, 1] - want_g[:, None]).max() <= 1.0


def test_mesh_cut_by_near_plane_is_watertight(oracle):
    """A 12 x 12 quad grid on a floor plane running from behind the camera
    to far in front: with additive counting, every covered pixel is hit
    exactly once, and every pixel whose centre sees the floor is covered."""
    proj = g._perspective(70.0, 1.0, 0.5, 50.0)
    n = 12
    xs = np.linspace(-6, 6, n + 1)
    zs = np.linspace(3, -30, n + 1)  # camera looks down -z
    verts = {}
    for i, x in enumerate(xs):
        for j, z in enumerate(zs):
            verts[i, j] = proj @ np.array([x, -1.0, z, 1.0])
    rows = []
    for i in range(n):
        for j in range(n):
            for tri in (((i, j), (i + 1, j), (i + 1, j + 1)), ((i, j), (i + 1, j + 1), (i, j + 1))):
                for k in tri:
                    rows.append([*verts[k].astype(np.float32), 1.0 / 255.0, 0.0, 0.0, 1.0])
    s = _scene(np.array(rows, np.float32))
    s.depth_test = 0
    s.depth_write = 0
    s.blend = (1, g.BLEND_ONE, g.BLEND_ONE, 0, g.BLEND_ONE, g.BLEND_ZERO, 0)
    c, _, _, st = oracle.render(s)
    assert st["triangles_clipped"] > 0
    count = c[:, :, 0, 0]
    assert count.max() == 1, "pixels covered twice: %d" % int((count > 1).sum())
    # the floor y = -1 is visible below the horizon (lower half of the screen)
    # for z between the near plane and z = -30 at x within the grid
    inv = np.linalg.inv(proj)
    miss = 0
    for py in range(H):
        for px in range(W):
            X, Y = (px + 0.5) / 32.0 - 1.0, (py + 0.5) / 32.0 - 1.0
            # the view ray through the pixel centre, hit with the floor plane
            a = inv @ np.array([X, Y, 0.0, 1.0])
            b = inv @ np.array([X, Y, 1.0, 1.0])
            a, b = a[:3] / a[3], b[:3] / b[3]
            dvec = b - a
            if abs(dvec[1]) < 1e-12:
                continue
            t = (-1.0 - a[1]) / dvec[1]
            if not 0.0 < t < 1.0:
                continue
            hit = a + t * dvec
            inside = -5.9 < hit[0] < 5.9 and -29.5 < hit[2] < 2.9
            if inside and count[py, px] != 1:
                miss += 1
    assert miss == 0
    assert st["fragments"] == int((count == 1).sum())


def test_triangle_entirely_behind_the_eye_is_dropped(oracle):
    v = np.array([[0, 0, 0.5, -1, 1, 1, 1, 1], [1, 0, 0.5, -1, 1, 1, 1, 1], [0, 1, 0.5, -2, 1, 1, 1, 1]],
                 np.float32)
    _, d, _, st = oracle.render(_scene(v))
    assert st["fragments"] == 0 and (d == 1.0).all()


def test_clip_polygon_vertex_order(oracle):
    """The polygon of a triangle cut by z = 0: kept vertices and crossing
    points in edge order, each crossing interpolated from its inside end
    (exact here: t = 0.5 and 0.25), given as weights over the triangle's
    vertices."""
    s = _scene(np.zeros((3, 8), np.float32))

    def expect(weights, P):
        return (np.array(weights, np.float64) @ P.astype(np.float64)).astype(np.float32)

    # v0 inside (z = 0.5), v1 and v2 outside (z = -0.5, -1.5)
    P = np.array([[0, 0, 0.5, 1], [1, 0, -0.5, 1], [0, 1, -1.5, 1]], np.float32)
    np.testing.assert_array_equal(oracle.clip_polygon(s, P), expect([[1, 0, 0], [0.5, 0.5, 0], [0.75, 0, 0.25]], P))
    # v0 outside: the polygon starts with the crossing of edge (v0, v1)
    P2 = np.array([[0, 0, -0.5, 1], [1, 0, 0.5, 1], [0, 1, 0.5, 1]], np.float32)
    np.testing.assert_array_equal(oracle.clip_polygon(s, P2),
                                  expect([[0.5, 0.5, 0], [0, 1, 0], [0, 0, 1], [0.5, 0, 0.5]], P2))
    # cut by z = 0 and z = w: six planes in order, the far plane second
    P3 = np.array([[0, 0, -0.5, 1], [1, 0, 0.5, 1], [0, 1, 1.5, 1]], np.float32)
    np.testing.assert_array_equal(
        oracle.clip_polygon(s, P3),
        expect([[0.5, 0.5, 0], [0, 1, 0], [0, 0.5, 0.5], [0.25, 0, 0.75], [0.75, 0, 0.25]], P3))

