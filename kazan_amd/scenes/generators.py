"""Seeded synthetic scenes for the draw-path workloads.

Every scene is a plain description (geometry arrays + pipeline state) that
both the Vulkan renderer (scenes/renderer.py) and the CPU oracle consume, so
the two see byte-identical inputs. All randomness comes from SplitMix64
streams (scenes/rng.py); `seed` selects the stream.

Workloads
  triangles_256      1,024 random triangles, non-indexed, 256x256, colour FS,
                     depth LESS, no blend                  (the CPU-path case)
  grid_1080p         512x1024 jittered quad grid = 1,048,576 indexed
                     triangles, 1920x1080, bilinear 1024^2 texture, alpha
                     blend SRC_ALPHA/ONE_MINUS_SRC_ALPHA, depth LESS
  (and small variants of both for parity tests)
"""
import numpy as np

from . import rng

VS_POS4_COLOR4 = 0
VS_POS3_UV2 = 1
VS_MVP_POS3_COLOR4 = 2
VS_MVP_POS3_NORMAL3 = 3
FS_COLOR = 0
FS_TEXTURE = 1
FS_PHONG = 2

COMPARE_LESS = 1
BLEND_ONE, BLEND_ZERO = 1, 0
BLEND_SRC_ALPHA, BLEND_ONE_MINUS_SRC_ALPHA = 6, 7


class Scene:
    """A draw-path workload: one render pass, one pipeline, N draws."""

    def __init__(self, name, width, height, vs, fs, vertices, indices=None, draws=None, samples=1):
        self.name = name
        self.width = width
        self.height = height
        self.samples = samples
        self.vs = vs
        self.fs = fs
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float32)
        self.indices = None if indices is None else np.ascontiguousarray(indices, dtype=np.uint32)
        n = len(self.indices) if self.indices is not None else len(self.vertices)
        self.draws = draws if draws is not None else [dict(count=n - n % 3, first=0, vertex_offset=0, mvp=None)]
        self.clear_color = (0.0, 0.0, 0.0, 0.0)
        self.clear_depth = 1.0
        self.depth_test, self.depth_write, self.depth_op = 1, 1, COMPARE_LESS
        self.blend = (0, BLEND_ONE, BLEND_ZERO, 0, BLEND_ONE, BLEND_ZERO, 0)
        self.write_mask = 15
        self.cull_mode, self.front_face = 0, 0
        self.viewport = (0.0, 0.0, float(width), float(height), 0.0, 1.0)
        self.scissor = (0, 0, width, height)
        self.texture = None  # (H, W, 4) uint8
        self.tex_filter = 1  # LINEAR
        self.tex_addr = (2, 2)  # CLAMP_TO_EDGE
        self.ubo = None
        self.num_lights = 0

    @property
    def stride(self):
        return self.vertices.shape[1]

    @property
    def num_triangles(self):
        return sum(d["count"] // 3 for d in self.draws)


def random_triangles(n_tris=1024, width=256, height=256, seed=rng.DEFAULT_SEED):
    """n_tris random triangles, vertex xy ~ U[-1,1]^2 (NDC), z ~ U[0,1], w = 1,
    per-vertex RGBA ~ U[0,1]; pass-through VS, interpolated-colour FS, depth
    LESS, no blend, cleared to (0,0,0,0) / 1.0. Non-indexed."""
    nv = 3 * n_tris
    u = rng.uniform_f32(seed, nv * 7).reshape(nv, 7)
    v = np.empty((nv, 8), np.float32)
    v[:, 0] = u[:, 0] * np.float32(2.0) - np.float32(1.0)
    v[:, 1] = u[:, 1] * np.float32(2.0) - np.float32(1.0)
    v[:, 2] = u[:, 2]
    v[:, 3] = 1.0
    v[:, 4:8] = u[:, 3:7]
    return Scene("triangles_%d" % width, width, height, VS_POS4_COLOR4, FS_COLOR, v)


def noise_texture(size=1024, seed=rng.DEFAULT_SEED + 1):
    return rng.uniform_u8(seed, size * size * 4).reshape(size, size, 4)


def jittered_grid(cols=512, rows=1024, width=1920, height=1080, tex_size=1024, seed=rng.DEFAULT_SEED):
    """A cols x rows quad grid over NDC [-1,1]^2, vertices jittered by
    U[-0.3, 0.3] of a cell in x and y (no folds), z ~ U[0,1], uv = grid
    position (unjittered). Two triangles per quad, indexed u32. Bilinear
    texture (seeded RGBA8 noise), alpha blend SRC_ALPHA/ONE_MINUS_SRC_ALPHA
    for colour and ONE/ONE_MINUS_SRC_ALPHA for alpha, depth LESS."""
    nx, ny = cols + 1, rows + 1
    nv = nx * ny
    u = rng.uniform_f32(seed, nv * 3).reshape(nv, 3)
    j = np.arange(ny, dtype=np.float32).repeat(nx)
    i = np.tile(np.arange(nx, dtype=np.float32), ny)
    cw, ch = np.float32(2.0 / cols), np.float32(2.0 / rows)
    v = np.empty((nv, 5), np.float32)
    v[:, 0] = np.float32(-1.0) + i * cw + (u[:, 0] * np.float32(0.6) - np.float32(0.3)) * cw
    v[:, 1] = np.float32(-1.0) + j * ch + (u[:, 1] * np.float32(0.6) - np.float32(0.3)) * ch
    v[:, 2] = u[:, 2]
    v[:, 3] = i / np.float32(cols)
    v[:, 4] = j / np.float32(rows)
    qi = np.arange(cols, dtype=np.uint32)
    qj = np.arange(rows, dtype=np.uint32)
    v00 = (qj[:, None] * nx + qi[None, :]).ravel()
    v10, v01 = v00 + 1, v00 + nx
    v11 = v01 + 1
    idx = np.stack([v00, v10, v11, v00, v11, v01], axis=1).ravel().astype(np.uint32)
    s = Scene("grid_%dx%d_%dx%d" % (cols, rows, width, height), width, height, VS_POS3_UV2, FS_TEXTURE, v, idx)
    s.texture = noise_texture(tex_size, seed + 1)
    s.blend = (1, BLEND_SRC_ALPHA, BLEND_ONE_MINUS_SRC_ALPHA, 0, BLEND_ONE, BLEND_ONE_MINUS_SRC_ALPHA, 0)
    s.clear_color = (0.0, 0.0, 0.0, 1.0)
    return s


WORKLOADS = {
    "triangles_256": lambda seed=rng.DEFAULT_SEED: random_triangles(1024, 256, 256, seed),
    "grid_1080p": lambda seed=rng.DEFAULT_SEED: jittered_grid(512, 1024, 1920, 1080, 1024, seed),
}
