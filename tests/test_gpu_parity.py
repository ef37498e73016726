"""The HIP draw path, driven through the Vulkan C ABI, against the CPU oracle.

Every scene is rendered twice from the same input buffers: once through
libkazan_driver.so (vkCreateGraphicsPipelines -> vkCmdDraw* ->
vkQueueSubmit on the MI355X) and once by the oracle. The bar:
coverage and depth bit-exact, RGBA8 colour within 1 LSB (these kernels aim
for, and so far report, exact colour too), fragment counts equal.
"""
import numpy as np
import pytest

from kazan_amd.scenes import generators as g
from scene_kit import random_partition, rect_tris, scene_from_pixels

pytestmark = pytest.mark.gpu

COLOR_TOL_LSB = 1


_CTX = {}


@pytest.fixture(params=["balanced", "plain"])
def mctx(request, driver, monkeypatch):
    """A device in each pipeline form: 'balanced' (merged draws, cost-split
    work items, persistent raster) and 'plain' (per-draw batches, one raster
    workgroup per tile). Pixels must not depend on the form."""
    from kazan_amd.scenes.renderer import Context
    if request.param not in _CTX:
        monkeypatch.setenv("KAZAN_PIPELINE", request.param)
        _CTX[request.param] = Context(driver)
    return _CTX[request.param]


def render_gpu(ctx, scene, frames=1):
    from kazan_amd.scenes.renderer import SceneRenderer
    r = SceneRenderer(ctx, scene)
    st = r.render(frames)
    c, d, res = r.read_back()
    return c, d, res, st


def check_parity(ctx, oracle, scene, frames=1):
    print("parity case: %s %dx%d samples=%d tris=%d" % (scene.name, scene.width, scene.height, scene.samples,
                                                          scene.num_triangles), flush=True)
    c, d, res, st = render_gpu(ctx, scene, frames)
    oc, od, ores, ost = oracle.render(scene)
    assert st["fragments"] == frames * ost["fragments"], (st, ost)
    assert st["triangles_setup"] == frames * ost["triangles_setup"]
    assert st["triangles_clipped"] == frames * ost["triangles_clipped"]
    cov_gpu, cov_cpu = d != np.float32(scene.clear_depth), od != np.float32(scene.clear_depth)
    np.testing.assert_array_equal(cov_gpu, cov_cpu)
    np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))  # depth bit-exact
    diff = np.abs(c.astype(np.int16) - oc.astype(np.int16))
    assert diff.max() <= COLOR_TOL_LSB, "colour differs by %d LSB at %s" % (diff.max(), np.argwhere(diff > 1)[:4])
    if scene.samples > 1:
        assert np.abs(res.astype(np.int16) - ores.astype(np.int16)).max() <= COLOR_TOL_LSB
    return st, ost, int((diff > 0).sum())


def test_triangles_256_config(mctx, oracle):
    st, ost, n_off = check_parity(mctx, oracle, g.random_triangles(1024, 256, 256))
    assert ost["fragments"] > 5_000_000
    assert n_off == 0


@pytest.mark.parametrize("seed", [1, 2])
def test_random_triangles_odd_size(mctx, oracle, seed):
    check_parity(mctx, oracle, g.random_triangles(300, 100, 70, seed=seed))


def test_grid_small_textured_blended(mctx, oracle):
    check_parity(mctx, oracle, g.jittered_grid(64, 128, 320, 240, 64))


@pytest.mark.parametrize("tex_size", [64, 48])
@pytest.mark.parametrize("filt,addr", [(0, 0), (1, 0), (0, 2), (1, 2)])
def test_texture_filter_and_addressing(mctx, oracle, tex_size, filt, addr):
    """NEAREST (0) and LINEAR (1) filtering with REPEAT (0) and
    CLAMP_TO_EDGE (2) addressing, on a power-of-two texture and another
    size, with texture coordinates in [-1.5, 2.5] so that both wrap."""
    s = g.jittered_grid(48, 64, 200, 150, tex_size)
    s.vertices[:, 3:5] = s.vertices[:, 3:5] * np.float32(4.0) - np.float32(1.5)
    s.tex_filter = filt
    s.tex_addr = (addr, addr)
    s.name += "_f%d_a%d" % (filt, addr)
    check_parity(mctx, oracle, s)


def test_grid_1080p_full(mctx, oracle):
    st, ost, n_off = check_parity(mctx, oracle, g.jittered_grid(512, 1024, 1920, 1080, 1024))
    assert ost["fragments"] > 2_000_000


@pytest.mark.parametrize("seed", range(2))
def test_partition_count_image(mctx, oracle, seed):
    s = scene_from_pixels(random_partition(64, 64, 8, np.random.default_rng(seed)), count_image=True)
    check_parity(mctx, oracle, s)


def test_overlapping_alpha_order(mctx, oracle):
    # 3,000 small translucent triangles piled into a 48 px square: the result
    # depends on the blend order at every pixel
    r = np.random.default_rng(11)
    centres = r.uniform(8, 56, (3000, 1, 2))
    tris = np.round((centres + r.uniform(-5, 5, (3000, 3, 2))) * 16) / 16
    cols = r.uniform(0, 1, (3000, 3, 4)).astype(np.float32)
    s = scene_from_pixels(tris, colors=cols, z=r.uniform(0, 1, (3000, 3)))
    s.blend = (1, g.BLEND_SRC_ALPHA, g.BLEND_ONE_MINUS_SRC_ALPHA, 0, g.BLEND_ONE, g.BLEND_ONE_MINUS_SRC_ALPHA, 0)
    s.depth_test = 0
    check_parity(mctx, oracle, s)


def test_dense_depth_and_blend(mctx, oracle):
    r = np.random.default_rng(12)
    centres = r.uniform(0, 64, (4000, 1, 2))
    tris = np.round((centres + r.uniform(-3, 3, (4000, 3, 2))) * 16) / 16
    s = scene_from_pixels(tris, colors=r.uniform(0, 1, (4000, 3, 4)), z=r.uniform(0, 1, (4000, 3)))
    s.blend = (1, g.BLEND_SRC_ALPHA, g.BLEND_ONE_MINUS_SRC_ALPHA, 0, g.BLEND_ONE, g.BLEND_ONE_MINUS_SRC_ALPHA, 0)
    check_parity(mctx, oracle, s)


@pytest.fixture
def small_list_ctx(driver, monkeypatch):
    # a device whose tile lists hold at most 40,000 (triangle, tile) pairs
    from kazan_amd.scenes.renderer import Context
    monkeypatch.setenv("KAZAN_LIST_CAPACITY_LIMIT", "40000")
    return Context(driver)


def test_huge_triangles_list_overflow(small_list_ctx, oracle):
    # 24 triangles each spanning the whole 1080p target: ~196,000 (triangle,
    # tile) pairs, more than the device's lists hold, so most tiles take the
    # in-order scan path
    r = np.random.default_rng(13)
    tris = []
    for k in range(24):
        tris.append([(-100 + 20 * k, -50), (2100, -30 - 10 * k), (900, 1400 + 5 * k)])
    v = np.zeros((72, 8), np.float32)
    t = np.array(tris, np.float64).reshape(-1, 2)
    v[:, 0] = t[:, 0] / 960.0 - 1.0
    v[:, 1] = t[:, 1] / 540.0 - 1.0
    v[:, 2] = r.uniform(0, 1, 72)
    v[:, 3] = 1.0
    v[:, 4:8] = r.uniform(0, 1, (72, 4))
    s = g.Scene("huge_24", 1920, 1080, g.VS_POS4_COLOR4, g.FS_COLOR, v)
    s.blend = (1, g.BLEND_SRC_ALPHA, g.BLEND_ONE_MINUS_SRC_ALPHA, 0, g.BLEND_ONE, g.BLEND_ZERO, 0)
    st, _, _ = check_parity(small_list_ctx, oracle, s)
    assert st["list_overflow"] >= 1


@pytest.fixture(params=["balanced", "plain"])
def cut_list_ctx(driver, monkeypatch, request):
    # tile lists of at most 20,000 (triangle, tile) pairs, in both forms
    from kazan_amd.scenes.renderer import Context
    monkeypatch.setenv("KAZAN_LIST_CAPACITY_LIMIT", "20000")
    monkeypatch.setenv("KAZAN_PIPELINE", request.param)
    monkeypatch.setenv("KAZAN_AUTOTUNE", "0")  # 16x16 tiles: the small triangles' pairs fit, the large ones' do not
    return Context(driver)


def test_cut_after_the_first_chunks(cut_list_ctx, oracle):
    """8,192 small triangles (two pair chunks, one tile each) and then 64
    large ones (about a thousand tiles each) in a third chunk: the lists hold
    the small triangles' pairs, the third chunk is cut, and each tile it
    touches draws its list and then scans the triangles from the cut on, in
    primitive order (translucent, so order shows)."""
    r = np.random.default_rng(16)
    W = H = 512
    c = r.uniform(4, W - 4, (8192, 1, 2))
    small = np.round((c + r.uniform(-3, 3, (8192, 3, 2))) * 16) / 16
    big = np.round(r.uniform(-40, W + 40, (64, 3, 2)) * 16) / 16
    tris = np.concatenate([small, big])
    n = len(tris)
    s = scene_from_pixels(tris, colors=r.uniform(0, 1, (n, 3, 4)), z=r.uniform(0, 1, (n, 3)), width=W, height=H)
    s.blend = (1, g.BLEND_SRC_ALPHA, g.BLEND_ONE_MINUS_SRC_ALPHA, 0, g.BLEND_ONE, g.BLEND_ONE_MINUS_SRC_ALPHA, 0)
    s.depth_test = 0
    st, ost, _ = check_parity(cut_list_ctx, oracle, s)
    assert st["list_overflow"] >= 1


@pytest.fixture
def tiny_list_ctx(driver, monkeypatch):
    # balanced form; tile lists and sub-tile lists hold 10,000 entries
    from kazan_amd.scenes.renderer import Context
    monkeypatch.setenv("KAZAN_LIST_CAPACITY_LIMIT", "10000")
    monkeypatch.setenv("KAZAN_PIPELINE", "balanced")
    monkeypatch.setenv("KAZAN_AUTOTUNE", "0")  # 16x16 tiles, as the scene below assumes
    return Context(driver)


def test_split_tile_sub_list_overflow(tiny_list_ctx, oracle):
    # 7,000 translucent triangles inside tile (1, 1): its list (7,000
    # entries) fits, the tile is split into 16 sub-tiles, and their lists
    # (several entries per triangle) do not, so the sub-tiles past the
    # storage filter the tile's list instead
    r = np.random.default_rng(15)
    n = 7000
    tris = np.round(r.uniform(16.1, 31.9, (n, 3, 2)) * 16) / 16
    s = scene_from_pixels(tris, colors=r.uniform(0, 1, (n, 3, 4)), z=r.uniform(0, 1, (n, 3)), width=64, height=64)
    s.blend = (1, g.BLEND_SRC_ALPHA, g.BLEND_ONE_MINUS_SRC_ALPHA, 0, g.BLEND_ONE, g.BLEND_ONE_MINUS_SRC_ALPHA, 0)
    s.depth_test = 0
    st, ost, _ = check_parity(tiny_list_ctx, oracle, s)
    assert st["list_overflow"] == 0
    assert st["split_entries"] > 10000
    assert ost["fragments"] > 100000


@pytest.mark.parametrize("cull,front", [(1, 0), (2, 0), (2, 1), (3, 0)])
def test_cull_modes(mctx, oracle, cull, front):
    s = g.random_triangles(200, 64, 64, seed=5)
    s.cull_mode, s.front_face = cull, front
    check_parity(mctx, oracle, s)


def test_depth_greater_and_scissor(mctx, oracle):
    s = g.random_triangles(300, 96, 64, seed=6)
    s.depth_op = 4  # GREATER
    s.clear_depth = 0.0
    s.scissor = (10, 7, 60, 40)
    check_parity(mctx, oracle, s)


def test_partial_render_area_keeps_the_rest(mctx, oracle):
    """A render pass over part of the framebuffer clears and draws only
    there (its clear is a fill of the area, not folded into the raster,
    which then loads the tiles); the pixels outside keep the previous
    frame's values."""
    import copy
    from kazan_amd.scenes.renderer import SceneRenderer
    s = g.random_triangles(300, 96, 64, seed=11)
    r = SceneRenderer(mctx, s)
    r.render(1)
    c1, d1, _ = r.read_back()
    oc1, od1, _, _ = oracle.render(s)
    np.testing.assert_array_equal(d1.view(np.uint32), od1.view(np.uint32))
    s.clear_color = (0.25, 0.5, 0.75, 1.0)
    s.render_area = (16, 8, 48, 40)
    r._record()
    r.render(1)
    c2, d2, _ = r.read_back()
    s2 = copy.copy(s)
    s2.scissor = (16, 8, 48, 40)  # the area bounds the draws too
    oc2, od2, _, _ = oracle.render(s2)
    inside = np.zeros((s.height, s.width), bool)
    inside[8:48, 16:64] = True
    np.testing.assert_array_equal(d2[inside].view(np.uint32), od2[inside].view(np.uint32))
    assert np.abs(c2[inside].astype(np.int16) - oc2[inside].astype(np.int16)).max() <= COLOR_TOL_LSB
    np.testing.assert_array_equal(c2[~inside], c1[~inside])
    np.testing.assert_array_equal(d2[~inside].view(np.uint32), d1[~inside].view(np.uint32))


def test_msaa4_triangles(mctx, oracle):
    s = g.random_triangles(400, 128, 96, seed=7)
    s.samples = 4
    check_parity(mctx, oracle, s)


def test_offscreen_degenerate_and_clipped(mctx, oracle):
    r = np.random.default_rng(14)
    v = np.zeros((3 * 60, 8), np.float32)
    v[:, 0:2] = r.uniform(-3, 3, (180, 2))  # many outside the view
    v[:, 2] = r.uniform(-0.5, 1.5, 180)      # some need z clipping
    v[:, 3] = 1.0
    v[:, 4:8] = r.uniform(0, 1, (180, 4))
    v[3:6, 0:2] = v[3, 0:2]                  # a zero-area triangle
    s = g.Scene("mixed_clip", 64, 64, g.VS_POS4_COLOR4, g.FS_COLOR, v)
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["triangles_clipped"] > 0


def test_repeated_submissions_accumulate(mctx, oracle):
    s = g.random_triangles(256, 64, 64, seed=8)
    st, ost, _ = check_parity(mctx, oracle, s, frames=3)
    assert st["fragments"] == 3 * ost["fragments"]


# --- the batching, skew and stress workloads, at sizes the oracle finishes
# in well under a second ------------------------------------------------------

def test_zipf_draws_mvp_push_constants(mctx, oracle):
    # 96 draws, rank-Zipf sizes, each with its own MVP in push constants
    s = g.zipf_draws(96, 30000, 320, 180)
    s.vertices[:, 0:3] = s.vertices[:, 0:3]  # (positions as generated)
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["triangles_setup"] > 0


def test_zipf_draws_bigger_triangles(mctx, oracle):
    from kazan_amd.scenes import rng
    counts = g.zipf_counts(64, 6000)
    v = g.cube_meshes(counts, size_cube=0.15, seed=rng.DEFAULT_SEED + 4)
    draws, first = [], 0
    for d, n in enumerate(counts):
        draws.append(dict(count=int(3 * n), first=first, vertex_offset=0, mvp=g.draw_mvp(77 + d, 400 / 240)))
        first += int(3 * n)
    s = g.Scene("zipf_big", 400, 240, g.VS_MVP_POS3_COLOR4, g.FS_COLOR, v, draws=draws)
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["fragments"] > 10000


def test_gaussian_cluster_skew(mctx, oracle):
    # 200,000 ~1 px triangles, 90% in a Gaussian cluster: tiles at the centre
    # hold thousands of triangles (split into work items in balanced form)
    s = g.gaussian_cluster(200_000, 320, 240)
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["fragments"] > 10000


def test_mixed_triangle_sizes(mctx, oracle):
    # a fine cluster (40,000 ~1 px triangles) beside a coarse one (400 of
    # ~30 px): the balanced form's plan splits by predicted fragments here
    s = g.mixed_clusters(40_000, 400, 320, 180)
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["fragments"] > 20000


def test_gaussian_cluster_long_lists(mctx, oracle):
    # a tighter cluster: centre tiles exceed one LDS sort run (4,096 ids)
    s = g.gaussian_cluster(300_000, 256, 256, sigma_frac=0.02)
    check_parity(mctx, oracle, s)


def test_phong_msaa4_stress_small(mctx, oracle):
    s = g.stress_4k(48, 12000, 320, 200, 4)
    s.vertices = s.vertices.copy()
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["fragments"] > 0


@pytest.mark.parametrize("form", ["balanced", "plain"])
def test_sort_binning(driver, monkeypatch, oracle, form):
    # the other binning (KAZAN_BINNING=sort): per-tile counts, an unordered
    # scatter, then every list sorted into primitive order
    from kazan_amd.scenes.renderer import Context
    monkeypatch.setenv("KAZAN_PIPELINE", form)
    monkeypatch.setenv("KAZAN_BINNING", "sort")
    ctx = Context(driver)
    for s in (g.jittered_grid(32, 64, 160, 120, 64), g.gaussian_cluster(300_000, 256, 256, sigma_frac=0.02),
              g.zipf_draws(64, 20_000, 192, 128)):
        check_parity(ctx, oracle, s)


@pytest.mark.parametrize("form", ["balanced", "plain"])
def test_setup_records(driver, monkeypatch, oracle, form):
    # the opt-in path where binning stores each triangle's setup and the
    # raster loads it instead of running the vertex shader and setup again
    from kazan_amd.scenes.renderer import Context
    monkeypatch.setenv("KAZAN_PIPELINE", form)
    monkeypatch.setenv("KAZAN_SETUP_RECORDS", "1")
    ctx = Context(driver)
    phong = g.stress_4k(16, 4000, 160, 100, 4)
    phong.vertices = phong.vertices.copy()
    for s in (g.jittered_grid(32, 64, 160, 120, 64), g.gaussian_cluster(60_000, 160, 128), phong):
        check_parity(ctx, oracle, s)


def test_phong_big_triangles(mctx, oracle):
    from kazan_amd.scenes import rng
    counts = g.zipf_counts(16, 800)
    v = g.cube_meshes(counts, size_cube=0.2, seed=rng.DEFAULT_SEED + 9, with_normals=True)
    draws, first = [], 0
    for d, n in enumerate(counts):
        draws.append(dict(count=int(3 * n), first=first, vertex_offset=0, mvp=g.draw_mvp(300 + d, 1.5)))
        first += int(3 * n)
    s = g.Scene("phong_big", 240, 160, g.VS_MVP_POS3_NORMAL3, g.FS_PHONG, v, draws=draws)
    s.ubo = g.phong_lights(8)
    s.num_lights = 8
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["fragments"] > 2000


# --- fragment discard (OpKill): late depth ----------------------------------

@pytest.mark.parametrize("blend", [0, 1])
def test_alpha_test_discard(mctx, oracle, blend):
    """if (v_color.a < 0.5) discard: a discarded fragment writes no depth and
    no colour, so the triangles behind it show through (the oracle runs such
    a shader before its depth test)."""
    r = np.random.default_rng(15 + blend)
    centres = r.uniform(4, 60, (1500, 1, 2))
    tris = np.round((centres + r.uniform(-8, 8, (1500, 3, 2))) * 16) / 16
    s = scene_from_pixels(tris, colors=r.uniform(0, 1, (1500, 3, 4)), z=r.uniform(0, 1, (1500, 3)))
    s.fs = g.FS_ALPHA_TEST
    if blend:
        s.blend = (1, g.BLEND_SRC_ALPHA, g.BLEND_ONE_MINUS_SRC_ALPHA, 0, g.BLEND_ONE, g.BLEND_ONE_MINUS_SRC_ALPHA, 0)
    st, ost, _ = check_parity(mctx, oracle, s)
    assert 0 < ost["depth_passed"] < ost["fragments"]
    assert st["depth_passed"] == ost["depth_passed"]
