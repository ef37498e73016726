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
    assert st["triangles_unsupported"] == frames * ost["triangles_unsupported"]
    cov_gpu, cov_cpu = d != np.float32(scene.clear_depth), od != np.float32(scene.clear_depth)
    np.testing.assert_array_equal(cov_gpu, cov_cpu)
    np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))  # depth bit-exact
    diff = np.abs(c.astype(np.int16) - oc.astype(np.int16))
    assert diff.max() <= COLOR_TOL_LSB, "colour differs by %d LSB at %s" % (diff.max(), np.argwhere(diff > 1)[:4])
    if scene.samples > 1:
        assert np.abs(res.astype(np.int16) - ores.astype(np.int16)).max() <= COLOR_TOL_LSB
    return st, ost, int((diff > 0).sum())


def test_triangles_256_config(ctx, oracle):
    st, ost, n_off = check_parity(ctx, oracle, g.random_triangles(1024, 256, 256))
    assert ost["fragments"] > 5_000_000
    assert n_off == 0


@pytest.mark.parametrize("seed", [1, 2])
def test_random_triangles_odd_size(ctx, oracle, seed):
    check_parity(ctx, oracle, g.random_triangles(300, 100, 70, seed=seed))


def test_grid_small_textured_blended(ctx, oracle):
    check_parity(ctx, oracle, g.jittered_grid(64, 128, 320, 240, 64))


def test_grid_1080p_full(ctx, oracle):
    st, ost, n_off = check_parity(ctx, oracle, g.jittered_grid(512, 1024, 1920, 1080, 1024))
    assert ost["fragments"] > 2_000_000


@pytest.mark.parametrize("seed", range(2))
def test_partition_count_image(ctx, oracle, seed):
    s = scene_from_pixels(random_partition(64, 64, 8, np.random.default_rng(seed)), count_image=True)
    check_parity(ctx, oracle, s)


def test_overlapping_alpha_order(ctx, oracle):
    # 3,000 small translucent triangles piled into a 48 px square: the result
    # depends on the blend order at every pixel
    r = np.random.default_rng(11)
    centres = r.uniform(8, 56, (3000, 1, 2))
    tris = np.round((centres + r.uniform(-5, 5, (3000, 3, 2))) * 16) / 16
    cols = r.uniform(0, 1, (3000, 3, 4)).astype(np.float32)
    s = scene_from_pixels(tris, colors=cols, z=r.uniform(0, 1, (3000, 3)))
    s.blend = (1, g.BLEND_SRC_ALPHA, g.BLEND_ONE_MINUS_SRC_ALPHA, 0, g.BLEND_ONE, g.BLEND_ONE_MINUS_SRC_ALPHA, 0)
    s.depth_test = 0
    check_parity(ctx, oracle, s)


def test_dense_depth_and_blend(ctx, oracle):
    r = np.random.default_rng(12)
    centres = r.uniform(0, 64, (4000, 1, 2))
    tris = np.round((centres + r.uniform(-3, 3, (4000, 3, 2))) * 16) / 16
    s = scene_from_pixels(tris, colors=r.uniform(0, 1, (4000, 3, 4)), z=r.uniform(0, 1, (4000, 3)))
    s.blend = (1, g.BLEND_SRC_ALPHA, g.BLEND_ONE_MINUS_SRC_ALPHA, 0, g.BLEND_ONE, g.BLEND_ONE_MINUS_SRC_ALPHA, 0)
    check_parity(ctx, oracle, s)


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
    ctx = small_list_ctx
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
    st, _, _ = check_parity(ctx, oracle, s)
    assert st["list_overflow"] >= 1


@pytest.mark.parametrize("cull,front", [(1, 0), (2, 0), (2, 1), (3, 0)])
def test_cull_modes(ctx, oracle, cull, front):
    s = g.random_triangles(200, 64, 64, seed=5)
    s.cull_mode, s.front_face = cull, front
    check_parity(ctx, oracle, s)


def test_depth_greater_and_scissor(ctx, oracle):
    s = g.random_triangles(300, 96, 64, seed=6)
    s.depth_op = 4  # GREATER
    s.clear_depth = 0.0
    s.scissor = (10, 7, 60, 40)
    check_parity(ctx, oracle, s)


def test_msaa4_triangles(ctx, oracle):
    s = g.random_triangles(400, 128, 96, seed=7)
    s.samples = 4
    check_parity(ctx, oracle, s)


def test_offscreen_degenerate_and_clipped(ctx, oracle):
    r = np.random.default_rng(14)
    v = np.zeros((3 * 60, 8), np.float32)
    v[:, 0:2] = r.uniform(-3, 3, (180, 2))  # many outside the view
    v[:, 2] = r.uniform(-0.5, 1.5, 180)      # some need z clipping
    v[:, 3] = 1.0
    v[:, 4:8] = r.uniform(0, 1, (180, 4))
    v[3:6, 0:2] = v[3, 0:2]                  # a zero-area triangle
    s = g.Scene("mixed_clip", 64, 64, g.VS_POS4_COLOR4, g.FS_COLOR, v)
    st, ost, _ = check_parity(ctx, oracle, s)
    assert ost["triangles_unsupported"] > 0


def test_repeated_submissions_accumulate(ctx, oracle):
    s = g.random_triangles(256, 64, 64, seed=8)
    st, ost, _ = check_parity(ctx, oracle, s, frames=3)
    assert st["fragments"] == 3 * ost["fragments"]
