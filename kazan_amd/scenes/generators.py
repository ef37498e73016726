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
FS_ALPHA_TEST = 3  # interpolated colour, discarded where alpha < 0.5

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
        self.render_area = None  # (x, y, w, h) of vkCmdBeginRenderPass; None: the whole framebuffer
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


def zipf_counts(n_draws, total, s=1.1, cap=65536, seed=rng.DEFAULT_SEED):
    """Rank-Zipf draw sizes: n_r = min(cap, max(1, round(K * r**-s))), K
    solved by bisection so that sum(n_r) is as close to `total` as the
    rounding allows; ranks are then permuted by the seed. (i.i.d. Zipf(1.1)
    draws clipped to 65,536 would total ~93M triangles for 4,096 draws, not
    ~4M; the rank form fixes the total the workload names.)"""
    r = np.arange(1, n_draws + 1, dtype=np.float64)

    def counts(K):
        return np.minimum(cap, np.maximum(1, np.rint(K * r ** -s))).astype(np.int64)

    lo, hi = 0.0, float(total) * 4
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if counts(mid).sum() < total:
            lo = mid
        else:
            hi = mid
    c = counts(hi)
    perm = np.argsort(rng.splitmix64(seed, n_draws))
    return c[perm]


def _perspective(fovy_deg, aspect, near, far):
    """Vulkan clip space (y down, z in [0, w]); column-major 4x4."""
    f = 1.0 / np.tan(np.radians(fovy_deg) / 2)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = -f
    m[2, 2] = far / (near - far)
    m[2, 3] = near * far / (near - far)
    m[3, 2] = -1.0
    return m


def _rotation(u):
    """A rotation matrix from three uniforms (a unit quaternion)."""
    u1, u2, u3 = u
    q = (np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
         np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3))
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def draw_mvp(seed, aspect):
    """Per-draw model-view-projection: the unit cube centred, randomly rotated,
    scaled by U[0.3, 1.0] and placed at x, y ~ U[-1.5, 1.5], depth U[4, 7] in
    front of a 60 degree perspective camera (near 0.5, far 20): every vertex
    lies inside the depth range, so no near/far clipping is needed."""
    u = rng.uniform_f32(seed, 8).astype(np.float64)
    model = np.eye(4)
    model[:3, :3] = _rotation(u[0:3]) * (0.3 + 0.7 * u[3])
    model[:3, 3] = [-1.5 + 3.0 * u[4], -1.5 + 3.0 * u[5], -(4.0 + 3.0 * u[6])]
    center = np.eye(4)
    center[:3, 3] = -0.5
    m = _perspective(60.0, aspect, 0.5, 20.0) @ model @ center
    return m.T.astype(np.float32).ravel()  # column-major


def cube_meshes(counts, size_cube=0.012, seed=rng.DEFAULT_SEED, with_normals=False):
    """Per draw, counts[d] small triangles in the unit cube: a centre
    ~ U[0,1]^3 plus per-vertex offsets ~ U[-size, size]^3. Colour ~ U[0,1]^4
    per vertex, or a unit normal (the triangle's centre direction from the
    cube centre, normalised) when with_normals."""
    n = int(np.sum(counts))
    u = rng.uniform_f32(seed, n * 12).reshape(n, 12).astype(np.float64)
    c = u[:, 0:3]
    off = (u[:, 3:12].reshape(n, 3, 3) * 2 - 1) * size_cube
    pos = (c[:, None, :] + off).reshape(n * 3, 3)
    if with_normals:
        d = pos - 0.5
        nrm = d / np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-6)
        v = np.concatenate([pos, nrm], axis=1).astype(np.float32)
    else:
        col = rng.uniform_f32(seed + 7, n * 3 * 4).reshape(n * 3, 4)
        v = np.concatenate([pos.astype(np.float32), col], axis=1).astype(np.float32)
    return v


def zipf_draws(n_draws=4096, total=4_194_304, width=2560, height=1440, seed=rng.DEFAULT_SEED, cap=65536):
    """Batching workload: n_draws draws with rank-Zipf(1.1) triangle counts
    (clipped to [1, cap]) summing to ~total, each its own seeded MVP
    (push constant) over a unit-cube mesh of small triangles. Depth LESS,
    no blend, interpolated colour."""
    counts = zipf_counts(n_draws, total, cap=cap, seed=seed)
    v = cube_meshes(counts, seed=seed + 11)
    draws, first = [], 0
    for d, n in enumerate(counts):
        draws.append(dict(count=int(3 * n), first=first, vertex_offset=0,
                          mvp=draw_mvp(seed + 1000 + d, width / height)))
        first += int(3 * n)
    s = Scene("zipf_%d_draws_%dx%d" % (n_draws, width, height), width, height, VS_MVP_POS3_COLOR4, FS_COLOR, v,
              draws=draws)
    return s


def gaussian_cluster(n_tris=4_194_304, width=1920, height=1080, cluster_frac=0.9, sigma_frac=0.05, radius=0.92,
                     seed=rng.DEFAULT_SEED, centre=(0.5, 0.5)):
    """Skewed workload: cluster_frac of the triangles centred ~ N(screen
    centre (or `centre`, fractions of W and H), sigma = sigma_frac * (W, H))
    and the rest ~ U(screen); vertices at
    per-vertex offsets ~ U[-radius, radius]^2 px from the centre (radius
    0.92 px gives ~30x overdraw at the cluster peak: peak density n_c /
    (2 pi sx sy) triangles per px^2 times the mean area 11/144 * (2r)^2).
    z ~ U[0,1], colour ~ U[0,1]^4, depth LESS + alpha blend."""
    nc = int(round(n_tris * cluster_frac))
    u = rng.uniform_f32(seed, n_tris * 4).reshape(n_tris, 4).astype(np.float64)
    cx = np.empty(n_tris)
    cy = np.empty(n_tris)
    # Box-Muller for the cluster, uniform for the rest
    r = np.sqrt(-2.0 * np.log(np.maximum(u[:nc, 0], 1e-12)))
    cx[:nc] = width * centre[0] + sigma_frac * width * r * np.cos(2 * np.pi * u[:nc, 1])
    cy[:nc] = height * centre[1] + sigma_frac * height * r * np.sin(2 * np.pi * u[:nc, 1])
    cx[nc:] = u[nc:, 2] * width
    cy[nc:] = u[nc:, 3] * height
    w = rng.uniform_f32(seed + 3, n_tris * 6).reshape(n_tris, 3, 2).astype(np.float64) * 2 - 1
    px = cx[:, None] + w[:, :, 0] * radius
    py = cy[:, None] + w[:, :, 1] * radius
    nv = n_tris * 3
    v = np.empty((nv, 8), np.float32)
    v[:, 0] = (px.ravel() * (2.0 / width) - 1.0).astype(np.float32)
    v[:, 1] = (py.ravel() * (2.0 / height) - 1.0).astype(np.float32)
    v[:, 2] = rng.uniform_f32(seed + 5, nv)
    v[:, 3] = 1.0
    v[:, 4:8] = rng.uniform_f32(seed + 6, nv * 4).reshape(nv, 4)
    s = Scene("cluster_%d_%dx%d" % (n_tris, width, height), width, height, VS_POS4_COLOR4, FS_COLOR, v)
    s.blend = (1, BLEND_SRC_ALPHA, BLEND_ONE_MINUS_SRC_ALPHA, 0, BLEND_ONE, BLEND_ONE_MINUS_SRC_ALPHA, 0)
    return s


def mixed_clusters(n_fine=3_000_000, n_coarse=30_000, width=1920, height=1080, seed=rng.DEFAULT_SEED):
    """Two Gaussian clusters side by side: n_fine triangles of radius 0.92 px
    (the skewed workload's size, ~0.3 fragments each) around (0.3 W, 0.5 H)
    and n_coarse of radius 10 px (~30 fragments each) around (0.7 W, 0.5 H).
    A tile in the fine cluster holds ten times the triangles of one in the
    coarse cluster but fewer fragments: binned triangle counts misjudge its
    raster cost, fragment counts do not."""
    a = gaussian_cluster(n_fine, width, height, 1.0, 0.05, 0.92, seed, (0.3, 0.5))
    b = gaussian_cluster(n_coarse, width, height, 1.0, 0.04, 10.0, seed + 7, (0.7, 0.5))
    s = Scene("mixed_%d_%d_%dx%d" % (n_fine, n_coarse, width, height), width, height, VS_POS4_COLOR4, FS_COLOR,
              np.concatenate([a.vertices, b.vertices]))
    s.blend = a.blend
    return s


def phong_lights(n=8, seed=rng.DEFAULT_SEED + 21):
    """The Phong uniform block {vec4 pos[n]; vec4 colour[n]; vec4 ambient}:
    lights around the unit cube, colours ~ U[0.2, 1]^3 / n."""
    u = rng.uniform_f32(seed, n * 6).reshape(n, 6).astype(np.float64)
    ubo = np.zeros((2 * n + 1, 4), np.float32)
    ubo[:n, 0:3] = u[:, 0:3] * 3.0 - 1.0
    ubo[n:2 * n, 0:3] = (0.2 + 0.8 * u[:, 3:6]) * (2.0 / n)
    ubo[2 * n, 0:3] = 0.05
    return ubo.ravel()


def stress_4k(n_draws=2048, total=8_388_608, width=3840, height=2160, samples=4, seed=rng.DEFAULT_SEED, cap=65536):
    """Combined stress: 4x MSAA at 3840x2160, total triangles in n_draws
    rank-Zipf draws (per-draw MVP), per-fragment Phong with 8 point lights
    and normals interpolated from per-vertex unit normals."""
    counts = zipf_counts(n_draws, total, cap=cap, seed=seed)
    v = cube_meshes(counts, size_cube=0.008, seed=seed + 13, with_normals=True)
    draws, first = [], 0
    for d, n in enumerate(counts):
        draws.append(dict(count=int(3 * n), first=first, vertex_offset=0,
                          mvp=draw_mvp(seed + 5000 + d, width / height)))
        first += int(3 * n)
    s = Scene("stress_%d_draws_%dx%d_x%d" % (n_draws, width, height, samples), width, height, VS_MVP_POS3_NORMAL3,
              FS_PHONG, v, draws=draws, samples=samples)
    s.ubo = phong_lights(8)
    s.num_lights = 8
    s.clear_color = (0.0, 0.0, 0.0, 1.0)
    return s


def camera_inside(n_small=4000, n_big=400, width=256, height=192, n_draws=4, samples=1, seed=rng.DEFAULT_SEED):
    """A camera standing inside the geometry: n_small small triangles (edges
    up to 4 % of their distance) and n_big large ones (edges up to 5 units)
    centred over a 60-degree perspective camera's view (near 0.5, far 20)
    from depth 0.3 to 16, drawn in n_draws draws
    with slightly different views (MVP push constants). Many large
    triangles cross the near plane or reach behind the eye, some leave the
    guard band or pass the far plane: the clipping workload. Small and large
    triangles are interleaved in primitive order. Depth LESS, alpha blend,
    interpolated colour."""
    n = n_small + n_big
    u = rng.uniform_f32(seed, n * 3).reshape(n, 3).astype(np.float64)
    # centres over the view frustum, widened by 20 %, from just behind the
    # eye to depth 16 (log-spaced, so that the near field is well populated)
    depth = np.exp(np.log(0.3) + u[:, 2] * (np.log(16.0) - np.log(0.3)))
    t = np.tan(np.radians(30.0)) * 1.2
    c = np.stack([(u[:, 0] * 2 - 1) * depth * t * width / height, (u[:, 1] * 2 - 1) * depth * t, -depth], axis=1)
    size = np.full(n, 0.02) * depth
    big = rng.splitmix64(seed + 1, n) % np.uint64(max(1, n // max(1, n_big))) == 0
    size[big] = 2.5
    off = (rng.uniform_f32(seed + 2, n * 9).reshape(n, 3, 3).astype(np.float64) * 2 - 1) * size[:, None, None]
    pos = (c[:, None, :] + off).reshape(n * 3, 3)
    col = rng.uniform_f32(seed + 3, n * 3 * 4).reshape(n * 3, 4)
    v = np.concatenate([pos.astype(np.float32), col], axis=1).astype(np.float32)
    proj = _perspective(60.0, width / height, 0.5, 20.0)
    draws, per = [], (n + n_draws - 1) // n_draws
    for d in range(n_draws):
        a, b = d * per, min(n, (d + 1) * per)
        view = np.eye(4)  # turned about the vertical axis by 0.06 rad per draw
        a_y = 0.06 * (d - (n_draws - 1) / 2.0)
        view[0, 0], view[0, 2], view[2, 0], view[2, 2] = np.cos(a_y), np.sin(a_y), -np.sin(a_y), np.cos(a_y)
        m = proj @ view
        draws.append(dict(count=3 * (b - a), first=3 * a, vertex_offset=0, mvp=m.T.astype(np.float32).ravel()))
    s = Scene("camera_inside_%d_%d_%dx%d" % (n_small, n_big, width, height), width, height, VS_MVP_POS3_COLOR4,
              FS_COLOR, v, draws=draws, samples=samples)
    s.blend = (1, BLEND_SRC_ALPHA, BLEND_ONE_MINUS_SRC_ALPHA, 0, BLEND_ONE, BLEND_ONE_MINUS_SRC_ALPHA, 0)
    return s


WORKLOADS = {
    "triangles_256": lambda seed=rng.DEFAULT_SEED: random_triangles(1024, 256, 256, seed),
    "grid_1080p": lambda seed=rng.DEFAULT_SEED: jittered_grid(512, 1024, 1920, 1080, 1024, seed),
    "zipf_4096_1440p": lambda seed=rng.DEFAULT_SEED: zipf_draws(4096, 4_194_304, 2560, 1440, seed),
    "cluster_4m_1080p": lambda seed=rng.DEFAULT_SEED: gaussian_cluster(4_194_304, 1920, 1080, seed=seed),
    "mixed_1080p": lambda seed=rng.DEFAULT_SEED: mixed_clusters(3_000_000, 30_000, 1920, 1080, seed),
    "stress_4k_msaa4": lambda seed=rng.DEFAULT_SEED: stress_4k(2048, 8_388_608, 3840, 2160, 4, seed),
}
