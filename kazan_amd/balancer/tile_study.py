"""Does a tile's raster work follow its binned triangle count? The question
behind the balanced form's cost model, answered from the tile profile.

  python -m kazan_amd.balancer.tile_study --workload cluster_4m_1080p [--frames 3] [--json out.json]

Renders the workload in the plain form (one raster workgroup per tile) with
tile profiling on (vkSetTileProfileKAZAN) and reports over the tiles with
binned triangles:
- how well binned triangles predict fragments: Pearson r, Spearman rank
  correlation, and the R^2 of the through-origin fit fragments ~ a * tris;
- how much of the raster cycles' variance triangles alone, fragments alone,
  and both explain (least squares with an intercept);
- the shipped model (cost_model.bin beside the library): the R^2 of its
  predicted fragments against the measured ones;
- the two plans' tile costs, the static one (binned triangles) and the
  model's (triangles + frag_weight x predicted fragments): their rank
  correlation and the overlap of the tiles each ranks most expensive (the
  top 2 % and 5 %, the tiles a plan splits first).
When fragments follow triangles (r near 1, the top sets nearly equal) the
static plan already orders the tiles as the model would, and the model's
own kernels can only add time. The plain form profiles its last batch, one
per draw: study single-draw workloads (a multi-draw one shows its last draw).
"""
import argparse
import json
import os
import sys

import numpy as np

from ..scenes import generators as g
from ..scenes.renderer import Context, SceneRenderer
from . import model as M
from .train import evaluate, profile


def _ranks(x):
    r = np.empty(len(x), np.float64)
    r[np.argsort(x, kind="stable")] = np.arange(len(x), dtype=np.float64)
    return r


def _r2(A, y):
    coef, *_ = np.linalg.lstsq(A, y, rcond=None)
    res = y - A @ coef
    return float(1.0 - res @ res / max(((y - y.mean()) ** 2).sum(), 1e-30))


def study(feat, frags, cycles, pred, frag_weight):
    tris = np.exp2(feat[:, 0].astype(np.float64)) - 1.0
    keep = tris > 0.5
    tris, frags, cycles = tris[keep], frags[keep].astype(np.float64), cycles[keep].astype(np.float64)
    pred = pred[keep].astype(np.float64)
    one = np.ones_like(tris)
    a = float(tris @ frags / (tris @ tris))
    out = {
        "tiles": int(keep.sum()),
        "fragments_per_triangle": a,
        "pearson_frags_tris": float(np.corrcoef(tris, frags)[0, 1]),
        "spearman_frags_tris": float(np.corrcoef(_ranks(tris), _ranks(frags))[0, 1]),
        "r2_frags_from_tris_through_origin": float(1.0 - ((frags - a * tris) ** 2).sum() / ((frags - frags.mean()) ** 2).sum()),
        "r2_cycles_from_tris": _r2(np.stack([one, tris], 1), cycles),
        "r2_cycles_from_frags": _r2(np.stack([one, frags], 1), cycles),
        "r2_cycles_from_both": _r2(np.stack([one, tris, frags], 1), cycles),
        "r2_model_pred_vs_frags": float(1.0 - ((frags - pred) ** 2).sum() / ((frags - frags.mean()) ** 2).sum()),
        "frag_weight": frag_weight,
    }
    static_cost = tris
    model_cost = tris + frag_weight * np.minimum(pred, tris * 16.0)
    out["spearman_static_vs_model_cost"] = float(np.corrcoef(_ranks(static_cost), _ranks(model_cost))[0, 1])
    for q in (0.02, 0.05):
        k = max(1, int(round(q * len(tris))))
        top_s = set(np.argsort(-static_cost, kind="stable")[:k])
        top_m = set(np.argsort(-model_cost, kind="stable")[:k])
        top_f = set(np.argsort(-cycles, kind="stable")[:k])
        out["top%d%%_overlap_static_model" % int(q * 100)] = len(top_s & top_m) / k
        out["top%d%%_overlap_static_cycles" % int(q * 100)] = len(top_s & top_f) / k
        out["top%d%%_overlap_model_cycles" % int(q * 100)] = len(top_m & top_f) / k
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="cluster_4m_1080p", choices=sorted(g.WORKLOADS))
    ap.add_argument("--frames", type=int, default=3)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    log = lambda m: print("tile_study: " + m, file=sys.stderr, flush=True)
    os.environ["KAZAN_PIPELINE"] = "plain"
    ctx = Context()
    ctx.d.vkSetTileProfileKAZAN(ctx.device, 1)
    scene = g.WORKLOADS[a.workload]()
    r = SceneRenderer(ctx, scene)
    rows = []
    for k in range(a.frames):
        r.render(1)
        rows.append(profile(ctx))
        log("frame %d profiled: %d tiles" % (k, len(rows[-1][1])))
    # the tiles' features and fragments repeat frame to frame; cycles vary: take their median
    feat, frags = rows[-1][0], rows[-1][1]
    cycles = np.median(np.stack([c for _, _, c in rows]), axis=0)
    blob_path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cost_model.bin")
    with open(blob_path, "rb") as f:
        blob = f.read()
    pred = evaluate(ctx.d, blob, feat)
    res = {"workload": a.workload, "frames": a.frames}
    res.update(study(feat, frags, cycles, pred, M.unpack(blob)["frag_weight"]))
    text = json.dumps(res, indent=1)
    print(text)
    if a.json:
        with open(a.json, "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
