"""Host-side pieces of the cost model, on the CPU: the blob format, the
numpy restatement of inference, and the training script's fits."""
import numpy as np

from kazan_amd.balancer import model as M
from kazan_amd.balancer import train as T


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between the bf16 neighbours 1 and 1 + 2^-7: ties
    # to even (1); 1 + 3 * 2^-9 is above halfway: up
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0], np.float32)
    np.testing.assert_array_equal(M.bf16(x), np.array([1.0, 1.0, 1.0 + 2 ** -7, -2.5, 0.0], np.float32))
    assert M.to_bf16_bits(np.array([1.0], np.float32))[0] == 0x3F80
    big = M.bf16(np.array([3.0e38], np.float32))[0]
    assert abs(big - 3.0e38) / 3.0e38 < 2 ** -8


def test_pack_unpack_round_trip():
    p = M.init_params(3)
    blob = M.pack(p, np.arange(16), np.ones(16) * 2, 0.25)
    m = M.unpack(blob)
    np.testing.assert_array_equal(m["w2"], M.bf16(p[M.P_W2:M.P_W3]).reshape(64, 64))
    np.testing.assert_array_equal(m["b3"], p[M.P_B3:M.P_B3 + 64])
    assert m["b4"] == p[M.P_B4] and m["frag_weight"] == 0.25
    np.testing.assert_array_equal(m["feat_mean"], np.arange(16, dtype=np.float32))


def test_reference_inference_of_a_constant_model():
    p = np.zeros(M.PARAMS, np.float32)
    p[M.P_B4] = 10.0  # output log2(1 + f) = 10 whatever the input
    blob = M.pack(p, np.zeros(16), np.ones(16), 1.0)
    pred = M.predict_reference(blob, np.random.default_rng(0).normal(size=(5, 16)))
    np.testing.assert_allclose(pred, 1023.0)


def test_reference_inference_of_a_linear_path():
    # one hidden path: y = relu(relu(relu(x0) ) ) with unit weights -> log2(1 + pred) = max(x0, 0)
    p = np.zeros(M.PARAMS, np.float32)
    p[M.P_W1 + 0 * 16 + 0] = 1.0
    p[M.P_W2 + 0 * 64 + 0] = 1.0
    p[M.P_W3 + 0 * 64 + 0] = 1.0
    p[M.P_W4 + 0] = 1.0
    blob = M.pack(p, np.zeros(16), np.ones(16), 1.0)
    x = np.zeros((3, 16), np.float32)
    x[:, 0] = [-1.0, 2.0, 5.0]
    np.testing.assert_allclose(np.log2(1 + M.predict_reference(blob, x)), [0.0, 2.0, 5.0], rtol=1e-6)


def test_frag_weight_fit_recovers_coefficients():
    r = np.random.default_rng(1)
    tris = r.integers(1, 5000, 4000).astype(np.float64)
    frags = tris * r.uniform(0.2, 8.0, 4000)
    cycles = 3000 + 40 * tris + 10 * frags + r.normal(0, 50, 4000)
    feat = np.zeros((4000, 16), np.float32)
    feat[:, 0] = np.log2(1 + tris)
    w, coef = T.fit_frag_weight(feat, frags.astype(np.float32), cycles.astype(np.float32))
    assert abs(w - 0.25) < 0.01, (w, coef)


def test_frag_weight_fit_stays_non_negative():
    # fragments proportional to triangles and cycles driven by fragments:
    # an unconstrained fit can return a negative triangle coefficient; the
    # fit keeps both coefficients non-negative and the weight at its bound
    r = np.random.default_rng(3)
    tris = r.integers(1, 5000, 4000).astype(np.float64)
    frags = 1.3 * tris
    cycles = 1000 + 25 * frags + r.normal(0, 200, 4000)
    feat = np.zeros((4000, 16), np.float32)
    feat[:, 0] = np.log2(1 + tris)
    w, coef = T.fit_frag_weight(feat, frags.astype(np.float32), cycles.astype(np.float32))
    assert min(coef) >= 0, coef
    assert T.W_MIN <= w <= T.W_MAX


def test_keep_tiles_keeps_every_busy_tile():
    r = np.random.default_rng(2)
    tris = np.concatenate([np.zeros(100), r.integers(1, 255, 2000), r.integers(256, 9000, 50)])
    feat = np.zeros((tris.size, 16), np.float32)
    feat[:, 0] = np.log2(1 + tris)
    idx = T.keep_tiles(feat, r, min_tris=256, random_tiles=512)
    assert set(np.flatnonzero(tris >= 256)) <= set(idx)
    assert idx.size == 50 + 512 and (tris[idx] > 0).all()


def test_error_stats():
    f = np.array([0, 50, 100, 200, 400], np.float64)
    s = T.error_stats(f * 1.1, f)
    assert s["tiles_ge_100_frags"] == 3 and abs(s["median_rel_err_ge_100"] - 0.1) < 1e-9
    assert T.error_stats(f, f)["r2_log2"] == 1.0


def _feat_of(tris):
    f = np.zeros((len(tris), M.FEATURES), np.float32)
    f[:, 0] = np.log2(1.0 + np.asarray(tris, np.float64))
    return f


def test_tile_study_when_fragments_follow_triangles():
    """Fragments proportional to triangles: correlation 1, the two plans
    rank the tiles alike, cycles follow both."""
    from kazan_amd.balancer.tile_study import study
    r = np.random.default_rng(4)
    tris = r.integers(1, 2000, 3000).astype(np.float64)
    frags = 0.5 * tris
    cycles = 1000 + 300 * frags
    res = study(_feat_of(tris), frags, cycles, frags.copy(), 1.0)
    assert res["tiles"] == 3000
    assert abs(res["fragments_per_triangle"] - 0.5) < 1e-3
    assert res["pearson_frags_tris"] > 0.999 and res["spearman_frags_tris"] > 0.999
    assert res["r2_frags_from_tris_through_origin"] > 0.999
    assert res["r2_cycles_from_tris"] > 0.999 and res["r2_model_pred_vs_frags"] > 0.999
    assert res["spearman_static_vs_model_cost"] > 0.999
    assert res["top5%_overlap_static_model"] > 0.95 and res["top2%_overlap_static_cycles"] > 0.95


def test_tile_study_when_fragments_do_not_follow_triangles():
    """Two populations: many tiny triangles with few fragments, few big ones
    with many. The triangle count misranks the expensive tiles, the model's
    cost (given its predictions) does not."""
    from kazan_amd.balancer.tile_study import study
    r = np.random.default_rng(5)
    fine = r.integers(800, 1200, 1500).astype(np.float64)    # many triangles, ~0.3 px each
    coarse = r.integers(20, 60, 1500).astype(np.float64)     # few triangles, ~60 px each
    tris = np.concatenate([fine, coarse])
    frags = np.concatenate([0.3 * fine, 60.0 * coarse])
    cycles = 500 + 350 * frags
    res = study(_feat_of(tris), frags, cycles, frags.copy(), 4.0)
    assert res["pearson_frags_tris"] < 0.0
    assert res["r2_cycles_from_frags"] > 0.999 and res["r2_cycles_from_tris"] < 0.9
    assert res["top5%_overlap_static_cycles"] < 0.1
    assert res["top5%_overlap_model_cycles"] > 0.9
    # tiles without triangles are left out
    res2 = study(_feat_of(np.r_[tris, 0, 0]), np.r_[frags, 0, 0], np.r_[cycles, 0, 0], np.r_[frags, 0, 0], 4.0)
    assert res2["tiles"] == res["tiles"]
