"""Clipping on the GPU (kz_clip_polygon in the binning and raster kernels)
against the oracle's R-CLIP, through the Vulkan C ABI.

Triangles that cross the near or far plane, reach behind the eye or leave
the guard band are clipped into a fan of pieces drawn in their place: the
binning kernel bins a clipped triangle by the union of its pieces' boxes,
and the raster draws the pieces as a batch of their own. Cases: a camera
inside the geometry (both pipeline forms, 1x and 4x samples), clipped
triangles interleaved with thousands of small ones in long tile lists,
the setup-record and sort-binning paths, and the list-overflow path (a
first frame whose pairs exceed the list storage). The bar is the parity
suite's: fragment, set-up and clipped counts equal, coverage and depth
bit-exact, colour within 1 LSB.
"""
import numpy as np
import pytest

from kazan_amd.scenes import generators as g

from test_gpu_parity import check_parity, mctx  # noqa: F401  (the per-form device fixture)

pytestmark = pytest.mark.gpu


def test_camera_inside_scene(mctx, oracle):
    s = g.camera_inside(4000, 400, 256, 192)
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["triangles_clipped"] > 50
    assert st["triangles_clipped"] == ost["triangles_clipped"]


def test_camera_inside_scene_msaa4(mctx, oracle):
    s = g.camera_inside(2000, 200, 160, 128, samples=4)
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["triangles_clipped"] > 20


def test_clipped_in_long_lists(mctx, oracle):
    """200,000 ~1 px triangles around the screen centre with 600 large
    near-plane-crossing ones interleaved: tiles hold thousands of entries,
    and pieces are drawn mid-list, after and before small triangles."""
    small = g.gaussian_cluster(200_000, 320, 240)
    big = g.camera_inside(0, 600, 320, 240, n_draws=1)
    # the big triangles' clip positions, written out (the small scene's
    # vertex shader passes clip positions through)
    m = np.asarray(big.draws[0]["mvp"], np.float32).reshape(4, 4).T
    pos = np.c_[big.vertices[:, 0:3], np.ones(len(big.vertices), np.float32)] @ m.T
    vb = np.c_[pos.astype(np.float32), big.vertices[:, 3:7]]
    vs = small.vertices.reshape(-1, 3, 8)
    vb = vb.reshape(-1, 3, 8)
    at = np.sort(np.random.default_rng(5).choice(len(vs), len(vb), replace=False))
    v = np.insert(vs, at, vb, axis=0).reshape(-1, 8)
    s = g.Scene("clipped_in_long_lists", 320, 240, g.VS_POS4_COLOR4, g.FS_COLOR, v)
    s.blend = small.blend
    st, ost, _ = check_parity(mctx, oracle, s)
    assert ost["triangles_clipped"] > 100


@pytest.mark.parametrize("form", ["balanced", "plain"])
def test_clip_list_overflow(driver, monkeypatch, oracle, form):
    """2,000 large clipped triangles over a 256x256 target on a device whose
    tile lists hold at most 20,000 (triangle, tile) pairs: the pairs overflow
    the list storage, and the tiles past it scan every triangle's box (the
    filter path), pieces included."""
    from kazan_amd.scenes.renderer import Context
    monkeypatch.setenv("KAZAN_PIPELINE", form)
    monkeypatch.setenv("KAZAN_LIST_CAPACITY_LIMIT", "20000")
    ctx = Context(driver)
    s = g.camera_inside(0, 2000, 256, 256, n_draws=2)
    st, ost, _ = check_parity(ctx, oracle, s, frames=2)
    assert ost["triangles_clipped"] > 200
    assert st["list_overflow"] >= 1


@pytest.mark.parametrize("form", ["balanced", "plain"])
@pytest.mark.parametrize("env", [("KAZAN_SETUP_RECORDS", "1"), ("KAZAN_BINNING", "sort")])
def test_clip_other_paths(driver, monkeypatch, oracle, form, env):
    from kazan_amd.scenes.renderer import Context
    monkeypatch.setenv("KAZAN_PIPELINE", form)
    monkeypatch.setenv(*env)
    ctx = Context(driver)
    for s in (g.camera_inside(3000, 300, 192, 128), g.camera_inside(1000, 100, 128, 96, samples=4)):
        check_parity(ctx, oracle, s)
