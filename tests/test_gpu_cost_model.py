"""The learned tile-cost model on the GPU: inference and training kernels
against numpy restatements, the tile profile against the rasteriser's own
counts, the device-side training-data generator against the host one, and
pixel neutrality of the balanced form whatever the model predicts."""
import ctypes

import numpy as np
import pytest

from kazan_amd import vk
from kazan_amd.balancer import model as M
from kazan_amd.scenes import generators as g

from test_gpu_parity import check_parity

pytestmark = pytest.mark.gpu

GAMMA = 0x9E3779B97F4A7C15


def _eval(driver, blob, feat):
    fn = driver.lib.kazan_cost_model_eval
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    feat = np.ascontiguousarray(feat, np.float32)
    out = np.zeros(len(feat), np.float32)
    assert fn(blob, len(blob), feat.ctypes.data, len(feat), out.ctypes.data) == 0
    return out


def _train(driver, x, y, state, steps, step0, batch, lr, seed):
    fn = driver.lib.kazan_cost_model_train
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64, ctypes.c_void_p]
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    state = np.ascontiguousarray(state, np.float32).copy()
    loss = np.zeros(steps, np.float32)
    assert fn(x.ctypes.data, y.ctypes.data, y.size, state.ctypes.data, steps, step0, batch, lr, seed,
              loss.ctypes.data) == 0
    return state, loss


def _random_blob(seed, b4=None, frag_weight=0.5):
    r = np.random.default_rng(seed)
    p = M.init_params(seed)
    p[M.P_B1:M.P_W4] = r.normal(0, 0.1, M.P_W4 - M.P_B1)
    p[M.P_B4] = 4.0 if b4 is None else b4
    if b4 is not None:
        p[M.P_W1:M.P_B4] = 0.0  # constant output b4
    return M.pack(p, r.normal(0, 1, 16), r.uniform(0.5, 2, 16), frag_weight)


def test_inference_kernel_matches_numpy(driver):
    blob = _random_blob(3)
    feat = np.random.default_rng(4).normal(0, 2, (1000, 16)).astype(np.float32)  # not a multiple of 16 rows
    gpu = _eval(driver, blob, feat)
    ref = M.predict_reference(blob, feat)
    d = np.abs(np.log2(1 + gpu) - np.log2(1 + ref))
    # bf16 activations: an fp32 sum in another order can round a hidden unit
    # to the neighbouring bf16 value, so agreement is to ~1e-2 in log2 units
    assert np.percentile(d, 99) < 0.02 and d.max() < 0.1, (np.percentile(d, 99), d.max())


def _splitmix_idx(seed, k, n):
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) ^ (k.astype(np.uint64) * np.uint64(0x632BE59BD9B4E019))) + np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z % np.uint64(n)).astype(np.int64)


def _reference_gradient(p, x, t):
    """The training kernel's arithmetic: bf16 weights, activations and
    backpropagated errors, fp32 (here f64) sums, mean squared error."""
    b = M.bf16
    w1 = b(p[M.P_W1:M.P_W2].reshape(64, 16)).astype(np.float64)
    w2 = b(p[M.P_W2:M.P_W3].reshape(64, 64)).astype(np.float64)
    w3 = b(p[M.P_W3:M.P_B1].reshape(64, 64)).astype(np.float64)
    b1, b2, b3 = (p[o:o + 64].astype(np.float64) for o in (M.P_B1, M.P_B2, M.P_B3))
    w4, b4 = p[M.P_W4:M.P_B4].astype(np.float64), float(p[M.P_B4])
    h0 = b(x).astype(np.float64)
    h1 = b(np.maximum(h0 @ w1.T + b1, 0).astype(np.float32)).astype(np.float64)
    h2 = b(np.maximum(h1 @ w2.T + b2, 0).astype(np.float32)).astype(np.float64)
    h3 = b(np.maximum(h2 @ w3.T + b3, 0).astype(np.float32)).astype(np.float64)
    y = h3 @ w4 + b4
    dy = 2 * (y - t) / len(t)
    g = np.zeros(M.PARAMS)
    g[M.P_W4:M.P_B4] = dy @ h3
    g[M.P_B4] = dy.sum()
    dz3 = np.where(h3 > 0, dy[:, None] * w4[None, :], 0.0)
    g[M.P_B3:M.P_B3 + 64] = dz3.sum(0)
    dz3b = b(dz3.astype(np.float32)).astype(np.float64)
    g[M.P_W3:M.P_B1] = (dz3b.T @ h2).ravel()
    dz2 = np.where(h2 > 0, dz3b @ w3, 0.0)
    g[M.P_B2:M.P_B2 + 64] = dz2.sum(0)
    dz2b = b(dz2.astype(np.float32)).astype(np.float64)
    g[M.P_W2:M.P_W3] = (dz2b.T @ h1).ravel()
    dz1 = np.where(h1 > 0, dz2b @ w2, 0.0)
    g[M.P_B1:M.P_B1 + 64] = dz1.sum(0)
    dz1b = b(dz1.astype(np.float32)).astype(np.float64)
    g[M.P_W1:M.P_W2] = (dz1b.T @ h0).ravel()
    return g


def test_training_kernel_gradient_matches_numpy(driver):
    r = np.random.default_rng(5)
    n, batch, seed = 3000, 256, 11
    x = r.normal(0, 1, (n, 16)).astype(np.float32)
    t = (2 + x[:, 0] - 0.5 * x[:, 1] ** 2).astype(np.float32)
    p = M.init_params(2)
    p[M.P_B1:M.P_W4] = r.normal(0, 0.05, M.P_W4 - M.P_B1)
    state0 = np.concatenate([p, np.zeros(2 * M.PARAMS, np.float32)])
    state, loss = _train(driver, x, t, state0, 1, 0, batch, 1e-3, seed)
    g_gpu = state[M.PARAMS:2 * M.PARAMS].astype(np.float64) / 0.1  # first moment after one step = 0.1 g
    idx = _splitmix_idx(seed, np.arange(batch), n)
    g_ref = _reference_gradient(p, x[idx], t[idx].astype(np.float64))
    for name, a, b in (("w1", M.P_W1, M.P_W2), ("w2", M.P_W2, M.P_W3), ("w3", M.P_W3, M.P_B1),
                       ("biases", M.P_B1, M.P_W4), ("w4", M.P_W4, M.P_B4), ("b4", M.P_B4, M.PARAMS)):
        err = np.linalg.norm(g_gpu[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-12)
        assert err < 2e-2, (name, err)
    # the reported loss is the minibatch mean squared error before the step
    y_ref = _reference_loss(p, x[idx], t[idx])
    assert abs(loss[0] - y_ref) < 1e-3 * max(1.0, y_ref)


def _reference_loss(p, x, t):
    b = M.bf16
    h = b(x).astype(np.float64)
    for wo, bo, k in ((M.P_W1, M.P_B1, 16), (M.P_W2, M.P_B2, 64), (M.P_W3, M.P_B3, 64)):
        w = b(p[wo:wo + 64 * k].reshape(64, k)).astype(np.float64)
        h = b(np.maximum(h @ w.T + p[bo:bo + 64], 0).astype(np.float32)).astype(np.float64)
    y = h @ p[M.P_W4:M.P_B4].astype(np.float64) + p[M.P_B4]
    return float(np.mean((y - t) ** 2))


def test_training_kernel_learns(driver):
    r = np.random.default_rng(6)
    x = r.normal(0, 1, (8192, 16)).astype(np.float32)
    t = (3 + x[:, 0] - 0.5 * x[:, 1] ** 2 + np.maximum(x[:, 2], 0)).astype(np.float32)
    state0 = np.concatenate([M.init_params(3), np.zeros(2 * M.PARAMS, np.float32)])
    state, loss = _train(driver, x, t, state0, 1500, 0, 256, 2e-3, 9)
    assert loss[-100:].mean() < 0.05 * loss[:20].mean(), (loss[:20].mean(), loss[-100:].mean())
    # the trained weights evaluate through the inference kernel as trained
    blob = M.pack(state[:M.PARAMS], np.zeros(16), np.ones(16), 1.0)
    y = np.log2(1 + _eval(driver, blob, x[:512]))
    assert np.mean((y - t[:512]) ** 2) < 0.2


def _profile(ctx):
    n = ctypes.c_uint32(0)
    ctx.d.check("vkGetTileProfileKAZAN", ctx.device, ctypes.byref(n), None, None, None, None)
    t = n.value
    arrs = [np.zeros((t, 16), np.float32)] + [np.zeros(t, np.float32) for _ in range(3)]
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    ctx.d.check("vkGetTileProfileKAZAN", ctx.device, ctypes.byref(n), *[fp(a) for a in arrs])
    return arrs


def _ctx(driver, monkeypatch, form, model="none"):
    from kazan_amd.scenes.renderer import Context
    monkeypatch.setenv("KAZAN_PIPELINE", form)
    monkeypatch.setenv("KAZAN_COST_MODEL", model)
    return Context(driver)


def test_tile_profile_counts(driver, monkeypatch, oracle):
    from kazan_amd.scenes.renderer import SceneRenderer
    ctx = _ctx(driver, monkeypatch, "plain")
    ctx.d.vkSetTileProfileKAZAN(ctx.device, 1)
    s = g.gaussian_cluster(100_000, 320, 240)
    r = SceneRenderer(ctx, s)
    st = r.render(1)
    feat, frags, cycles, pred = _profile(ctx)
    assert feat.shape == (20 * 15, 16)
    # labels: per-tile fragments sum to the frame's count; binned triangles
    # (feature 0) sum to the list entries
    assert frags.sum() == st["fragments"]
    assert np.rint(np.exp2(feat[:, 0].astype(np.float64)) - 1).sum() == st["list_entries"]
    assert (cycles[frags > 0] > 0).all() and (pred == 0).all()
    # histogram fractions sum to one on busy tiles; pipeline bits: blend on,
    # depth on, 4 varyings, 1x samples
    busy = feat[:, 0] > 0
    np.testing.assert_allclose(feat[busy, 2:7].sum(1), 1.0, atol=1e-5)
    assert (feat[:, 10] == 1).all() and (feat[:, 11] == 1).all() and (feat[:, 9] == 0).all()
    np.testing.assert_allclose(feat[:, 12], 4 / 16)
    assert (feat[:, 8] >= 0).all() and (feat[:, 8] <= 1).all()


def test_features_do_not_depend_on_the_binning(driver, monkeypatch):
    # the pair statistics gathered by the radix sort's last pass equal those
    # of the per-tile count kernel of the other binning, tile for tile
    from kazan_amd.scenes.renderer import SceneRenderer
    s = g.gaussian_cluster(100_000, 320, 240)
    feats = []
    for binning in ("radix", "sort"):
        monkeypatch.setenv("KAZAN_BINNING", binning)
        ctx = _ctx(driver, monkeypatch, "plain")
        ctx.d.vkSetTileProfileKAZAN(ctx.device, 1)
        SceneRenderer(ctx, s).render(1)
        feats.append(_profile(ctx)[:2])
    np.testing.assert_array_equal(feats[0][0], feats[1][0])
    np.testing.assert_array_equal(feats[0][1], feats[1][1])


def test_fused_cost_pass_matches_the_separate_kernels(driver, monkeypatch):
    # the balanced form's one-launch cost pass (features, predictions and the
    # plan's cost parts) against the plain form's feature and MLP kernels on
    # the same frame, and the predictions against numpy
    from kazan_amd.scenes.renderer import SceneRenderer
    monkeypatch.setenv("KAZAN_AUTOTUNE", "0")
    blob = _random_blob(5)
    s = g.gaussian_cluster(100_000, 320, 240)
    out = {}
    for form in ("plain", "balanced"):
        ctx = _ctx(driver, monkeypatch, form)
        ctx.d.check("vkSetCostModelKAZAN", ctx.device, blob, len(blob))
        ctx.d.vkSetTileProfileKAZAN(ctx.device, 1)
        SceneRenderer(ctx, s).render(1)
        out[form] = _profile(ctx)
    fp, pp = out["plain"][0], out["plain"][3]
    fb, pb = out["balanced"][0], out["balanced"][3]
    assert fb.shape == (20 * 15, 16) and (fb[:, 0] > 0).any()
    np.testing.assert_allclose(fb, fp, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(pb, pp, rtol=1e-5, atol=1e-6)
    assert (pb > 0).any()
    ref = M.predict_reference(blob, fb)
    d = np.abs(np.log2(1 + pb) - np.log2(1 + ref))
    assert np.percentile(d, 99) < 0.02 and d.max() < 0.1, (np.percentile(d, 99), d.max())


@pytest.mark.parametrize("b4", [20.0, 200.0, -20.0])
def test_balanced_pixels_do_not_depend_on_the_model(driver, monkeypatch, oracle, b4):
    # b4 = 20: every tile predicted at a million fragments, so every tile is
    # cut into 16 sub-tiles; b4 = 200: a prediction beyond fp32 (exp2 -> inf),
    # which the plan caps at the tile's samples x triangles, so tiles still
    # split; b4 = -20: no fragments predicted, so tiles split by their
    # binned-triangle count alone, as without a model
    ctx = _ctx(driver, monkeypatch, "balanced")
    blob = _random_blob(8, b4=b4, frag_weight=1.0)
    ctx.d.check("vkSetCostModelKAZAN", ctx.device, blob, len(blob))
    assert ctx.d.vkHasCostModelKAZAN(ctx.device) == 1
    static = _ctx(driver, monkeypatch, "balanced")
    for s in (g.gaussian_cluster(60_000, 160, 128), g.jittered_grid(32, 64, 160, 120, 64),
              g.random_triangles(200, 96, 80, seed=4)):
        st, _, _ = check_parity(ctx, oracle, s)
        st0, _, _ = check_parity(static, oracle, s)
        if b4 > 0:
            assert st["split_entries"] > st0["split_entries"]
        else:
            assert st["split_entries"] == st0["split_entries"]


def test_invalid_model_blob_is_rejected(driver, monkeypatch):
    ctx = _ctx(driver, monkeypatch, "balanced")
    bad = bytearray(_random_blob(1))
    bad[0] ^= 0xFF
    r = ctx.d.vkSetCostModelKAZAN(ctx.device, bytes(bad), len(bad))
    assert r == vk.VK_ERROR_INITIALIZATION_FAILED
    assert ctx.d.vkHasCostModelKAZAN(ctx.device) == 0


def _read_buffer(ctx, buf, nbytes):
    stage = ctx.buffer(nbytes, True)
    cb = ctx.command_buffer()
    ctx.begin(cb)
    region = vk.VkBufferCopy(0, 0, nbytes)
    ctx.d.vkCmdCopyBuffer(cb, buf, stage[0], 1, ctypes.byref(region))
    ctx.d.check("vkEndCommandBuffer", cb)
    ctx.submit(cb)
    out = np.frombuffer(bytes((ctypes.c_uint8 * nbytes).from_address(ctx.map(stage[1], nbytes))), np.uint8).copy()
    ctx.destroy_buffer(stage)
    return out


def test_device_cluster_generator_matches_host(driver, monkeypatch):
    ctx = _ctx(driver, monkeypatch, "plain")
    n, seed = 5000, 1234
    host = g.gaussian_cluster(n, 640, 480, cluster_frac=0.8, sigma_frac=0.07, radius=1.5, seed=seed).vertices
    buf = ctx.device_buffer_with(np.zeros_like(host))
    p = vk.KazanClusterScene(seed=seed, triangles=n, width=640, height=480, cluster_frac=0.8, sigma_frac=0.07,
                             radius=1.5)
    ctx.d.check("vkFillClusterSceneKAZAN", ctx.device, buf[0], 0, ctypes.byref(p))
    dev = _read_buffer(ctx, buf[0], host.nbytes).view(np.float32).reshape(host.shape)
    # the device's double-precision log/sin/cos may differ from the host's
    # in the last bit, which rarely survives the rounding to f32
    same = dev == host
    assert same.mean() > 0.999, same.mean()
    assert np.abs(dev - host).max() < 1e-5
    # uniform fields (z, colour, w) are exact
    assert (dev[:, 2:] == host[:, 2:]).all()
    # out-of-range requests are refused
    p.triangles = n + 1
    assert ctx.d.vkFillClusterSceneKAZAN(ctx.device, buf[0], 0, ctypes.byref(p)) == vk.VK_ERROR_INITIALIZATION_FAILED
