"""Trains the balanced form's tile-cost model (cost_model.bin) on a GPU.

  python -m kazan_amd.balancer.train [--frames 2000] [--out kazan_amd/cost_model.bin] [--report report.json]

Data. Plain-form frames with tile profiling: per tile, the 16 features the
binning kernels compute, the fragments the plain rasteriser counts (the
label) and the raster workgroup's cycles. Most frames are the
Gaussian-cluster workload (4,194,304 triangles at 1920x1080, 90% in the
cluster) at consecutive seeds, written on the device by
vkFillClusterSceneKAZAN; a quarter of them vary the cluster's width and the
triangle size, and a few frames of the other generators (textured grid,
large random triangles) show the model other pipelines and triangle sizes.
Every tile with at least 256 binned triangles is kept, plus 512 random tiles
per frame.

Training. The target is log2(1 + fragments); features are standardised;
the driver's MFMA training kernel (kazan_cost_model_train) runs minibatch
Adam from He-initialised weights. frag_weight, the list-entry equivalent of
one fragment in raster time, is the ratio of the fragment and triangle
coefficients of a non-negative least-squares fit of the measured cycles,
clipped to [W_MIN, W_MAX].

Evaluation. The last tenth of the frames is held out; the report gives the
model's error on their measured fragment counts beside that of a
per-triangle linear fit (fragments ~ binned triangles).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from .. import vk
from ..scenes import generators as g
from ..scenes import rng
from ..scenes.renderer import Context, SceneRenderer
from . import model as M

CLUSTER_TRIS = 4_194_304
WIDTH, HEIGHT = 1920, 1080


def profile(ctx):
    n = ctypes.c_uint32(0)
    ctx.d.check("vkGetTileProfileKAZAN", ctx.device, ctypes.byref(n), None, None, None, None)
    t = n.value
    feat = np.zeros((t, M.FEATURES), np.float32)
    frags = np.zeros(t, np.float32)
    cycles = np.zeros(t, np.float32)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    ctx.d.check("vkGetTileProfileKAZAN", ctx.device, ctypes.byref(n), fp(feat), fp(frags), fp(cycles), None)
    return feat, frags, cycles


def keep_tiles(feat, r, min_tris=256, random_tiles=512):
    tris = np.exp2(feat[:, 0]) - 1.0
    busy = np.flatnonzero(tris >= min_tris - 0.5)
    rest = np.flatnonzero((tris >= 0.5) & (tris < min_tris - 0.5))
    pick = r.choice(rest, min(random_tiles, rest.size), replace=False) if rest.size else rest
    return np.concatenate([busy, pick])


def collect(ctx, frames, log):
    """(features, fragments, cycles, frame id) of the kept tiles."""
    r = np.random.default_rng(7)
    rows = []
    base = g.gaussian_cluster(CLUSTER_TRIS, WIDTH, HEIGHT)
    sr = SceneRenderer(ctx, base)
    t0 = time.time()
    for k in range(frames):
        p = vk.KazanClusterScene(seed=rng.DEFAULT_SEED + 1000 + k, triangles=CLUSTER_TRIS, width=WIDTH,
                                 height=HEIGHT, cluster_frac=0.9, sigma_frac=0.05, radius=0.92)
        if k % 4 == 3:  # variations of the cluster's width and the triangle size
            p.sigma_frac = float(r.uniform(0.015, 0.15))
            p.radius = float(np.exp(r.uniform(np.log(0.4), np.log(4.0))))
            p.cluster_frac = float(r.uniform(0.5, 0.95))
        ctx.d.check("vkFillClusterSceneKAZAN", ctx.device, sr.vbuf[0], 0, ctypes.byref(p))
        sr.render(1)
        feat, frags, cycles = profile(ctx)
        idx = keep_tiles(feat, r)
        rows.append((feat[idx], frags[idx], cycles[idx], np.full(idx.size, k, np.int32)))
        if k % 100 == 0:
            log("frame %d/%d: %d tiles kept, %.1f s" % (k, frames, idx.size, time.time() - t0))
    # other generators: a few frames each
    extra = [g.jittered_grid(512, 1024, WIDTH, HEIGHT, 1024, seed=rng.DEFAULT_SEED + 50 + i) for i in range(3)]
    extra += [g.random_triangles(4000, WIDTH, HEIGHT, seed=rng.DEFAULT_SEED + 60 + i) for i in range(3)]
    for i, s in enumerate(extra):
        er = SceneRenderer(ctx, s)
        er.render(1)
        feat, frags, cycles = profile(ctx)
        idx = keep_tiles(feat, r)
        rows.append((feat[idx], frags[idx], cycles[idx], np.full(idx.size, frames + i, np.int32)))
        log("extra frame %s: %d tiles kept" % (s.name, idx.size))
    cols = list(zip(*rows))
    return tuple(np.concatenate(c) for c in cols) + (frames + len(extra),)


W_MIN, W_MAX = 0.02, 4.0


def fit_frag_weight(feat, frags, cycles):
    """cycles ~ c0 + c_tri * tris + c_frag * fragments with non-negative
    coefficients (tile cycles cannot fall with more work; when fragments
    and triangles are nearly proportional an unconstrained fit trades one
    coefficient against the other); returns c_frag / c_tri, clipped to
    [W_MIN, W_MAX] (W_MAX when only fragments explain the cycles)."""
    from scipy.optimize import nnls
    tris = np.exp2(feat[:, 0].astype(np.float64)) - 1.0
    ok = (tris > 0) & (cycles > 0)
    A = np.stack([np.ones(ok.sum()), tris[ok], frags[ok].astype(np.float64)], 1)
    scale = A.max(0)
    scale[scale == 0] = 1.0
    coef, _ = nnls(A / scale, cycles[ok].astype(np.float64))
    coef = coef / scale
    w = coef[2] / coef[1] if coef[1] > 0 else (W_MAX if coef[2] > 0 else W_MIN)
    return float(np.clip(w, W_MIN, W_MAX)), [float(c) for c in coef]


def train(driver, x, y, steps, batch, lr, seed, state=None, step0=0):
    state = np.concatenate([M.init_params(seed), np.zeros(2 * M.PARAMS, np.float32)]) if state is None else state
    state = np.ascontiguousarray(state, np.float32)
    loss = np.zeros(steps, np.float32)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    fn = driver.lib.kazan_cost_model_train
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64, ctypes.c_void_p]
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    rc = fn(x.ctypes.data, y.ctypes.data, y.size, state.ctypes.data, steps, step0, batch, lr, seed, loss.ctypes.data)
    if rc != 0:
        raise RuntimeError("kazan_cost_model_train failed (%d)" % rc)
    return state, loss


def evaluate(driver, blob, feat):
    fn = driver.lib.kazan_cost_model_eval
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    feat = np.ascontiguousarray(feat, np.float32)
    out = np.zeros(len(feat), np.float32)
    rc = fn(blob, len(blob), feat.ctypes.data, len(feat), out.ctypes.data)
    if rc != 0:
        raise RuntimeError("kazan_cost_model_eval failed (%d)" % rc)
    return out


def error_stats(pred, frags):
    big = frags >= 100
    rel = np.abs(pred[big] - frags[big]) / frags[big]
    lp, lf = np.log2(1 + pred), np.log2(1 + frags)
    r2 = 1.0 - np.sum((lp - lf) ** 2) / max(np.sum((lf - lf.mean()) ** 2), 1e-12)
    return {"tiles": int(frags.size), "tiles_ge_100_frags": int(big.sum()),
            "median_rel_err_ge_100": float(np.median(rel)) if rel.size else None,
            "p90_rel_err_ge_100": float(np.percentile(rel, 90)) if rel.size else None,
            "r2_log2": float(r2)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(__file__)), "cost_model.bin"))
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    log = lambda m: print("train: " + m, file=sys.stderr, flush=True)

    os.environ["KAZAN_PIPELINE"] = "plain"
    os.environ["KAZAN_COST_MODEL"] = "none"
    ctx = Context()
    ctx.d.vkSetTileProfileKAZAN(ctx.device, 1)
    feat, frags, cycles, frame, nframes = collect(ctx, a.frames, log)
    log("%d tiles from %d frames" % (frags.size, nframes))

    held = frame >= int(a.frames * 0.9)
    held &= frame < a.frames  # held-out cluster frames; the extra frames train
    tr = ~held
    frag_weight, coef = fit_frag_weight(feat[tr], frags[tr], cycles[tr])
    mean = feat[tr].mean(0)
    std = feat[tr].std(0)
    scale = (1.0 / np.maximum(std, 1e-3)).astype(np.float32)
    x = (feat - mean) * scale
    y = np.log2(1.0 + frags)

    t0 = time.time()
    state, loss1 = train(ctx.d, x[tr], y[tr], a.steps, a.batch, a.lr, a.seed)
    state, loss2 = train(ctx.d, x[tr], y[tr], a.steps // 4, a.batch, a.lr * 0.2, a.seed, state, a.steps)
    t_train = time.time() - t0
    blob = M.pack(state[:M.PARAMS], mean, scale, frag_weight)
    with open(a.out, "wb") as f:
        f.write(blob)
    log("model written to %s (train %.2f s, final loss %.4f)" % (a.out, t_train, float(loss2[-50:].mean())))

    pred = evaluate(ctx.d, blob, feat[held])
    tris = np.exp2(feat[:, 0]) - 1.0
    k = np.sum(tris[tr] * frags[tr]) / max(np.sum(tris[tr] ** 2), 1e-12)  # frags ~ k * tris
    report = {
        "frames": nframes, "tiles": int(frags.size), "train_tiles": int(tr.sum()), "held_out_tiles": int(held.sum()),
        "steps": a.steps + a.steps // 4, "batch": a.batch, "lr": a.lr, "train_seconds": t_train,
        "loss_first": float(loss1[:50].mean()), "loss_last": float(loss2[-50:].mean()),
        "frag_weight": frag_weight, "cycle_fit_c0_ctri_cfrag": coef,
        "model_held_out": error_stats(pred, frags[held]),
        "linear_in_tris_held_out": error_stats(k * tris[held], frags[held]),
        "feature_mean": [float(v) for v in mean], "feature_scale": [float(v) for v in scale],
    }
    log(json.dumps(report["model_held_out"]) + " vs linear " + json.dumps(report["linear_in_tris_held_out"]))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
