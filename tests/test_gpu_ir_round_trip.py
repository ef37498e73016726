"""Pipelines built from IR text parsed back (KAZAN_IR_ROUND_TRIP=1): the
translator hands the backend the module that its text form parses to, and
the scenes render as the oracle says, as they do from the translator's own
module. Shaders with textures, uniform-buffer loops (Phong), discard and
the control-flow templates' branches and switches."""
import pytest

from kazan_amd.scenes import generators as g

from test_gpu_parity import check_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt_ctx(driver):
    from kazan_amd.scenes.renderer import Context
    mp = pytest.MonkeyPatch()
    mp.setenv("KAZAN_IR_ROUND_TRIP", "1")
    ctx = Context(driver)
    yield ctx
    mp.undo()


def test_round_trip_scenes(rt_ctx, oracle):
    phong = g.stress_4k(16, 3000, 160, 100, 4)
    phong.vertices = phong.vertices.copy()
    for s in (g.jittered_grid(32, 64, 160, 120, 64), g.zipf_draws(32, 8000, 192, 128), phong,
              g.camera_inside(1500, 150, 128, 96)):
        check_parity(rt_ctx, oracle, s)


def test_round_trip_alpha_test(rt_ctx, oracle):
    import numpy as np
    from scene_kit import scene_from_pixels
    r = np.random.default_rng(21)
    centres = r.uniform(4, 60, (800, 1, 2))
    tris = np.round((centres + r.uniform(-8, 8, (800, 3, 2))) * 16) / 16
    s = scene_from_pixels(tris, colors=r.uniform(0, 1, (800, 3, 4)), z=r.uniform(0, 1, (800, 3)))
    s.fs = g.FS_ALPHA_TEST
    st, ost, _ = check_parity(rt_ctx, oracle, s)
    assert 0 < ost["depth_passed"] < ost["fragments"]
