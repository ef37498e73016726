"""Small hand-built scenes for known-answer and invariant tests.

Positions are given in pixels (multiples of 1/16) on power-of-two targets, so
pixel -> NDC -> viewport -> pixel round-trips exactly and the expected
coverage can be stated in closed form.
"""
import numpy as np

from kazan_amd.scenes import generators as g

COUNT = 1.0 / 255.0  # colour increment of a "count image" (blend ONE + ONE)


def px_to_ndc(p, size):
    return np.float32(p) * np.float32(2.0 / size) - np.float32(1.0)


def scene_from_pixels(tris, width=64, height=64, colors=None, z=None, count_image=False, samples=1):
    """tris: (n, 3, 2) pixel coordinates; colors: (n, 3, 4) or None; z: (n, 3) or None."""
    tris = np.asarray(tris, np.float64).reshape(-1, 3, 2)
    n = len(tris)
    v = np.zeros((n * 3, 8), np.float32)
    v[:, 0] = px_to_ndc(tris[:, :, 0].ravel(), width)
    v[:, 1] = px_to_ndc(tris[:, :, 1].ravel(), height)
    v[:, 2] = 0.5 if z is None else np.asarray(z, np.float32).ravel()
    v[:, 3] = 1.0
    if count_image:
        v[:, 4] = COUNT
        v[:, 7] = 1.0
    elif colors is not None:
        v[:, 4:8] = np.asarray(colors, np.float32).reshape(-1, 4)
    else:
        v[:, 4:8] = 1.0
    s = g.Scene("kit", width, height, g.VS_POS4_COLOR4, g.FS_COLOR, v, samples=samples)
    if count_image:
        s.depth_test = 0
        s.depth_write = 0
        s.blend = (1, g.BLEND_ONE, g.BLEND_ONE, 0, g.BLEND_ONE, g.BLEND_ZERO, 0)
        s.clear_color = (0.0, 0.0, 0.0, 0.0)
    return s


def rect_tris(x0, y0, x1, y1, diagonal=0):
    """Two triangles covering [x0, x1] x [y0, y1]."""
    a, b, c, d = (x0, y0), (x1, y0), (x1, y1), (x0, y1)
    if diagonal == 0:
        return [[a, b, c], [a, c, d]]
    return [[a, b, d], [b, c, d]]


def closed_form_count(lo, hi):
    """Samples at 16*p + 8 (1/16 px) with lo <= sample < hi (top/left
    inclusive, bottom/right exclusive), lo/hi in pixels."""
    lo16, hi16 = int(round(lo * 16)), int(round(hi * 16))
    first = -((-(lo16 - 8)) // 16)  # ceil((lo16 - 8) / 16)
    last = -((-(hi16 - 8)) // 16) - 1
    return max(0, last - first + 1), first


def random_partition(width, height, n_interior, rng):
    """A triangulated grid over the whole target with interior vertices
    jittered by at most 0.15 cell on the 1/16 px lattice (too little for any
    quad to fold): every sample lies in exactly one triangle."""
    gx, gy = n_interior + 2, n_interior + 2
    xs = np.linspace(0, width, gx)
    ys = np.linspace(0, height, gy)
    P = np.zeros((gy, gx, 2))
    for j in range(gy):
        for i in range(gx):
            x, y = xs[i], ys[j]
            if 0 < i < gx - 1:
                x += rng.uniform(-0.15, 0.15) * (width / (gx - 1))
            if 0 < j < gy - 1:
                y += rng.uniform(-0.15, 0.15) * (height / (gy - 1))
            P[j, i] = (np.round(x * 16) / 16, np.round(y * 16) / 16)
    tris = []
    for j in range(gy - 1):
        for i in range(gx - 1):
            a, b, c, d = P[j, i], P[j, i + 1], P[j + 1, i + 1], P[j + 1, i]
            if rng.random() < 0.5:
                tris += [[a, b, c], [a, c, d]]
            else:
                tris += [[a, b, d], [b, c, d]]
    return np.array(tris)
