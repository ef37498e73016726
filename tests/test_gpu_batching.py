"""Batched launch: every tile size the autotuner may pick, the autotuner's
trials, and hipGraph replay of a recorded command buffer, each against the
CPU oracle (pixels must not depend on tile size, grid or launch path)."""
import ctypes

import numpy as np
import pytest

from kazan_amd.scenes import generators as g
from test_gpu_parity import COLOR_TOL_LSB, check_parity

pytestmark = pytest.mark.gpu


def _ctx(driver, monkeypatch, **env):
    from kazan_amd.scenes.renderer import Context
    for k, v in env.items():
        monkeypatch.setenv(k, str(v))
    return Context(driver)


@pytest.mark.parametrize("shift", [3, 5, 6])
@pytest.mark.parametrize("form", ["balanced", "plain"])
def test_fixed_tile_size(driver, monkeypatch, oracle, shift, form):
    ctx = _ctx(driver, monkeypatch, KAZAN_PIPELINE=form, KAZAN_AUTOTUNE=0, KAZAN_TILE_SHIFT=shift,
               KAZAN_COST_MODEL="none")
    check_parity(ctx, oracle, g.jittered_grid(48, 96, 200, 150, 64))
    check_parity(ctx, oracle, g.gaussian_cluster(80_000, 200, 136))
    # 4x MSAA at 64x64 tiles does not fit LDS: the pipeline uses 32x32
    s = g.random_triangles(300, 130, 90, seed=9)
    s.samples = 4
    check_parity(ctx, oracle, s)


def _graph_stats(ctx):
    c, n = ctypes.c_uint32(), ctypes.c_uint32()
    ctx.d.vkGetGraphStatsKAZAN(ctx.device, ctypes.byref(c), ctypes.byref(n))
    return c.value, n.value


def test_autotune_trials_then_graph_replays(driver, monkeypatch, oracle):
    from kazan_amd.scenes.renderer import SceneRenderer
    ctx = _ctx(driver, monkeypatch, KAZAN_PIPELINE="balanced", KAZAN_AUTOTUNE=1, KAZAN_GRAPHS=1)
    s = g.zipf_draws(96, 30000, 320, 180)
    oc, od, _, ost = oracle.render(s)
    r = SceneRenderer(ctx, s)
    assert ctx.d.vkHasCostModelKAZAN(ctx.device) == 1  # the shipped model
    frames = 32
    for frame in range(frames):
        print("frame %d" % frame, flush=True)
        st = r.render(1)
        c, d, _ = r.read_back()  # eager work between graph replays
        assert st["fragments"] == ost["fragments"], frame
        np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))
        assert np.abs(c.astype(np.int16) - oc.astype(np.int16)).max() <= COLOR_TOL_LSB, frame
    ts, gd, cm, pr, ml, fs = (ctypes.c_uint32() for _ in range(6))
    ctx.d.vkGetTuningKAZAN(ctx.device, *(ctypes.byref(x) for x in (ts, gd, cm, pr, ml, fs)))
    assert ts.value in (8, 16, 32, 64) and gd.value in (1, 2, 4) and cm.value in (0, 1) and pr.value in (0, 1)
    assert ml.value in (0, 1, 2) and fs.value in (1, 4, 16)
    # timed trials: the 4 tile sizes 8..64, two grid sizes, the plan without
    # the model and the plain raster; when a planned raster is fastest, two
    # split depths, two fragment weights (with the model) and the 11 other
    # pairs of tile size and grid size; the three fastest twice more each;
    # then the choice, a capture and graph launches
    trials = 8 + (15 if pr.value else 0) + 6
    captures, launches = _graph_stats(ctx)
    assert captures == 1 and launches == frames - trials - 1, (captures, launches, trials)


def test_rerecording_recaptures(driver, monkeypatch, oracle):
    from kazan_amd.scenes.renderer import SceneRenderer
    ctx = _ctx(driver, monkeypatch, KAZAN_PIPELINE="balanced", KAZAN_AUTOTUNE=0, KAZAN_GRAPHS=1)
    s = g.random_triangles(200, 96, 64, seed=12)
    r = SceneRenderer(ctx, s)
    r.render(3)
    c0, _ = _graph_stats(ctx)
    r._record()  # a new command buffer: a new capture
    r.render(2)
    c1, n1 = _graph_stats(ctx)
    assert c1 == c0 + 1
    c, d, _ = r.read_back()
    oc, od, _, _ = oracle.render(s)
    np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))
    assert np.abs(c.astype(np.int16) - oc.astype(np.int16)).max() <= COLOR_TOL_LSB
