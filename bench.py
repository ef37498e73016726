"""Draw-path throughput on one MI355X.

A step is one frame: the recorded command buffer (attachment clears + all
draws of the workload) submitted once through vkQueueSubmit. Geometry,
indices and textures are resident in HBM before the timed region; the timed
region is K back-to-back submissions bracketed by vkQueueWaitIdle.

Prints one JSON line: value = fragments/s (covered (triangle, sample) pairs
before the depth test, the count the CPU oracle reproduces), ms_per_step =
frame time, a roofline object for the dominant kernel (HIP-event timed on the
queue's stream) and a cpu_baseline object (the C++ CPU restatement, run on
this host's cores on the same input buffers).

Each pipeline form in --forms is measured on its own device: "learned"
(balanced form planned by the cost model), "balanced_static" (balanced form,
split by binned counts only) and "plain" (one workgroup per tile). The first
form is the reported value; config.forms holds all of them and
config.<form>_vs_plain the throughput ratios.

  python bench.py [--gpus 1] [--steps 50] [--warmup 30] [--workload grid_1080p]
                  [--forms learned,balanced_static,plain]
  python bench.py --copy-probe 1073741824   # HBM counter calibration only
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fragments/sec + frame ms at 1080p, learned/batched vs plain GPU pipeline"


def algorithmic_bytes(scene, fragments):
    """Geometry and texture read once + per fragment D32 read+write and RGBA8
    write (+ RGBA8 read when blending)."""
    geo = scene.vertices.nbytes + (scene.indices.nbytes if scene.indices is not None else 0)
    tex = scene.texture.nbytes if scene.texture is not None else 0
    per_frag = 16 if scene.blend[0] else 12
    return geo + tex + fragments * per_frag, per_frag


def host_threads():
    """Host cores this process may use: the CPU affinity set, bounded by the
    share the machine grants one GPU's job when it states one
    (OMP_NUM_THREADS); returns (threads, visible cores)."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(visible, int(share)) if share.isdigit() and int(share) > 0 else visible
    return max(1, threads), visible


def cpu_baseline(scene, fragments, seconds):
    from oracle import bridge  # the CPU restatement: measured beside the GPU, never the product
    threads, visible = host_threads()
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < 2:
        t0 = time.perf_counter()
        _, _, _, st = bridge.render(scene, threads=threads)
        times.append(time.perf_counter() - t0)
        if len(times) >= 200:
            break
    assert st["fragments"] == fragments, "CPU and GPU fragment counts differ"
    med = statistics.median(times)
    return {"value": fragments / med, "unit": "fragments/s", "cores": threads, "kind": "port",
            "label": "build CPU restatement, not Kazan (Kazan has no CPU draw path)",
            "host_cores_visible": visible, "frame_ms": med * 1e3,
            "sample": "full %s frame (%d fragments), %d repetitions, median" % (scene.name, fragments, len(times))}


FORMS = {
    # name: (KAZAN_PIPELINE, KAZAN_COST_MODEL or None for the default model)
    "learned": ("balanced", None),
    "balanced_static": ("balanced", "none"),
    "plain": ("plain", "none"),
}


def measure(scene, form, args):
    """Times K frames of `scene` on a device created in pipeline form `form`
    (the environment is read at device creation), then one pass with
    per-kernel HIP events."""
    from kazan_amd.scenes.renderer import Context, SceneRenderer
    pipeline, model = FORMS[form]
    os.environ["KAZAN_PIPELINE"] = pipeline
    if model is None:
        os.environ.pop("KAZAN_COST_MODEL", None)
    else:
        os.environ["KAZAN_COST_MODEL"] = model
    ctx = Context()
    has_model = bool(ctx.d.vkHasCostModelKAZAN(ctx.device))
    if form == "learned" and not has_model:
        print("bench: no cost model (kazan_amd/cost_model.bin); 'learned' runs the static split rule",
              file=sys.stderr, flush=True)
    r = SceneRenderer(ctx, scene)
    print("bench[%s]: %s on %s, %d triangles" % (form, scene.name, ctx.device_name, scene.num_triangles),
          file=sys.stderr, flush=True)
    # steady state first: the balanced form times one tuning candidate per
    # submission, and every form runs its first submissions eagerly before
    # the command buffer is captured into a hipGraph. Submit until the
    # engine reports the graph captured (so tuning is over), then warm up.
    caps, launches = ctypes.c_uint32(), ctypes.c_uint32()
    tuning_frames = 0
    while tuning_frames < 200:
        ctx.d.vkGetGraphStatsKAZAN(ctx.device, ctypes.byref(caps), ctypes.byref(launches))
        if caps.value >= 1:
            break
        r.render(1)
        tuning_frames += 1
    steady = caps.value >= 1
    if args.warmup:
        r.render(args.warmup)
    ctx.stats()  # drops the counters of the tuning and warm-up frames
    ctx.d.vkGetGraphStatsKAZAN(ctx.device, ctypes.byref(caps), ctypes.byref(launches))
    caps0 = caps.value
    # timed region: K frames back to back, bracketed by queue waits
    ctx.wait_idle()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.submit(r.cb, wait=False)
    ctx.wait_idle()
    t1 = time.perf_counter()
    st = ctx.stats()
    ctx.d.vkGetGraphStatsKAZAN(ctx.device, ctypes.byref(caps), ctypes.byref(launches))
    recaptures = caps.value - caps0
    print("bench[%s]: %d frames in %.3f s, gpu %.3f ms total" % (form, args.steps, t1 - t0, st["gpu_ms"]),
          file=sys.stderr, flush=True)
    ctx.set_kernel_timing(True)
    n_prof = max(3, min(args.steps, 10))
    r.render(n_prof)
    ktimes = ctx.kernel_times()
    ctx.set_kernel_timing(False)
    ts, gd, cm, pr, ml, fs = (ctypes.c_uint32() for _ in range(6))
    ctx.d.vkGetTuningKAZAN(ctx.device, *(ctypes.byref(x) for x in (ts, gd, cm, pr, ml, fs)))
    return {"stats": st, "wall": t1 - t0, "ktimes": ktimes, "n_prof": n_prof, "ctx": ctx, "renderer": r,
            "cost_model": has_model, "tuning_frames": tuning_frames, "steady": steady,
            "recaptures_in_timed_region": recaptures,
            "tuned": {"tile": ts.value, "grid_divisor": gd.value, "cost_model": bool(cm.value),
                      "persistent_raster": bool(pr.value), "max_split_level": ml.value,
                      "frag_weight_scale": fs.value}}


def skewed_summary(scene, order, args):
    """Frame time and fragments/s of each form on a skewed workload (the
    case the balanced form and its cost model exist for), and the ratios to
    the plain form."""
    res = {}
    for form in order:
        m = measure(scene, form, args)
        res[form] = {"ms_per_frame": m["wall"] * 1e3 / args.steps,
                     "fragments_per_s": m["stats"]["fragments"] / m["wall"],
                     "fragments_per_frame": m["stats"]["fragments"] // args.steps,
                     "tuning_frames": m["tuning_frames"], "tuned": m["tuned"],
                     "recaptures_in_timed_region": m["recaptures_in_timed_region"]}
        del m
    out = {"workload": scene.name, "triangles": scene.num_triangles, "forms": res}
    if "plain" in res:
        for f in res:
            if f != "plain":
                out["%s_vs_plain" % f] = res[f]["fragments_per_s"] / res["plain"]["fragments_per_s"]
    return out


def copy_probe(nbytes, reps=3):
    """A plain device-to-device copy of known size, `reps` times: the scale
    for rocprofv3's FETCH_SIZE / WRITE_SIZE counters (run it under the same
    counter as the frame and divide by reps x nbytes)."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 4, "reps": reps}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--forms", default="learned,balanced_static,plain",
                    help="pipeline forms to measure, first = the reported one (%s)" % ",".join(FORMS))
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # the balanced form autotunes over its first (at most) 12 submissions
    # (one timed candidate each), runs the choice eagerly on the next and
    # captures the command buffer into a hipGraph on the one after: warm up
    # past that
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="grid_1080p")
    ap.add_argument("--skewed", default="cluster_4m_1080p",
                    help="skewed workload whose frame times and ratios per form go into config.skewed "
                         "('none' skips it)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes 3 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print("bench.py measures one GPU: the draw path runs on a single MI355X (multi-GPU sharding is a "
              "design note, not a measured configuration); run with --gpus 1", file=sys.stderr)
        return 2

    from kazan_amd.scenes import generators as g

    scene = g.WORKLOADS[args.workload]()
    order = [f for f in args.forms.split(",") if f]
    for f in order:
        if f not in FORMS:
            print("unknown form %s (known: %s)" % (f, ", ".join(FORMS)), file=sys.stderr)
            return 2
    forms = {}
    for form in order:
        forms[form] = measure(scene, form, args)
    main_form = order[0]
    m = forms[main_form]
    st, wall, ktimes, n_prof = m["stats"], m["wall"], m["ktimes"], m["n_prof"]
    frags_per_frame = st["fragments"] // args.steps
    dom = max(ktimes, key=lambda k: ktimes[k][0])
    dom_ms, dom_n = ktimes[dom]
    avg_launch_s = dom_ms / max(dom_n, 1) / 1e3
    alg_bytes, per_frag = algorithmic_bytes(scene, frags_per_frame)
    achieved = alg_bytes / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("raster_hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": frags_per_frame * args.steps / wall,
        "unit": "fragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene generator)",
        "config": {
            "workload": scene.name,
            "width": scene.width,
            "height": scene.height,
            "samples": scene.samples,
            "triangles": scene.num_triangles,
            "fragments_per_frame": frags_per_frame,
            "gpu_ms_per_frame": st["gpu_ms"] / args.steps,
            "batches_per_frame": st["batches"] / args.steps,
            "kernel_ms_per_frame": {k: v[0] / n_prof for k, v in ktimes.items()},
            "pipeline": main_form,
            "forms": {f: {"fragments_per_s": v["stats"]["fragments"] / v["wall"],
                          "ms_per_frame": v["wall"] * 1e3 / args.steps,
                          "gpu_ms_per_frame": v["stats"]["gpu_ms"] / args.steps,
                          "launches_per_frame": v["stats"]["launches"] / args.steps,
                          "split_entries_per_frame": v["stats"]["split_entries"] / args.steps,
                          "cost_model": v["cost_model"],
                          "tuned": v["tuned"],
                          "tuning_frames": v["tuning_frames"],
                          "steady_state": v["steady"],
                          "recaptures_in_timed_region": v["recaptures_in_timed_region"],
                          "kernel_ms_per_frame": {k: t[0] / v["n_prof"] for k, t in v["ktimes"].items()}}
                      for f, v in forms.items()},
        },
        "roofline": {
            "kernel": "kz_raster" if dom == "raster" else "kz_" + dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": alg_bytes,
            "bytes_per_fragment": per_frag,
            "avg_launch_ms": avg_launch_s * 1e3,
        },
    }
    if "plain" in forms:
        rate = lambda f: forms[f]["stats"]["fragments"] / forms[f]["wall"]
        for f in forms:
            if f != "plain":
                out["config"]["%s_vs_plain" % f] = rate(f) / rate("plain")
    if args.skewed != "none" and args.skewed != args.workload:
        out["config"]["skewed"] = skewed_summary(g.WORKLOADS[args.skewed](), order, args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, frags_per_frame, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
