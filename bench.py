#!/usr/bin/env python3
"""Benchmark of the MI355X tiled rasterizer on the 4K, 2M-triangle scene.

Workload: a 1024x1024-quad fBm terrain (2,097,152 triangles) seen at a 10
degree grazing angle plus 8 screen-covering backdrop triangles, 3840x2160,
four 2048^2 RGBA8 maps, fp32 depth.  A step is one frame over the next
camera of a seeded 100-frame path (disjoint from the 200-frame path the
bin-capacity model is trained on), enqueued back to back stage by stage
(--graph: replayed from a captured graph; the other form is timed beside the
headline one and reported with it).  Scene data is uploaded once, so the
inputs are resident in HBM when the timed region starts.  Binning parameters come from the autotuner's table
(custom-gpgpu-rasterizer_amd/tuning.json) unless given on the command line.

Prints one JSON line: whole-frame Mtris/s (value) for learned single-pass
binning, the same for exact two-pass binning and for the previous-frame
heuristic, the plain configuration (every triangle binned, no setup-time
raster) for the binner comparison, shaded Gpix/s, the shading pass against
the HBM roofline (algorithmic bytes from the committed fixture
tests/golden/terrain_shading_bytes.json, HBM traffic from the committed
rocprofv3 counter summary), the 256-job batched workload, and the CPU oracle
timed on the host cores (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "custom-gpgpu-rasterizer_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
SHADING_FIXTURE = os.path.join(ROOT, "tests", "golden", "terrain_shading_bytes.json")
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_shade_terrain4k_traffic.txt")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--tile", type=int, default=None)
    ap.add_argument("--big", type=int, default=None, help="tile span from which a triangle goes to the screen list")
    ap.add_argument("--micro", type=int, default=None, help="bbox pixel limit of the setup-time raster")
    ap.add_argument("--small", type=int, default=None, help="tile coverage from which a triangle is rasterized per pixel")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-binner comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-job batched workload")
    ap.add_argument("--only-batch", action="store_true", help="run only the 256-job batched workload")
    ap.add_argument("--no-flythrough", action="store_true", help="skip the 7-view flythrough workload")
    ap.add_argument("--only-flythrough", action="store_true", help="run only the 7-view flythrough workload")
    ap.add_argument("--fly-frames", type=int, default=300)
    ap.add_argument("--fly-train-frames", type=int, default=100)
    ap.add_argument("--graph", action="store_true",
                    help="headline frames replayed from captured graphs (default: enqueued stage by stage; "
                         "graph replay is timed beside it either way)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--train-frames", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--margin", type=float, default=None,
                    help="learned queue depth margin (default: the tuning table's for the workload's bucket)")
    ap.add_argument("--slack", type=int, default=None, help="learned queue depth slack (default: as --margin)")
    ap.add_argument("--cpu-frames", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--policies", default="exact,learned,prevframe")
    ap.add_argument("--resolve-mode", type=int, default=0,
                    help="resolve launch form: 0 one workgroup per tile, 1 persistent, 2 persistent with shading")
    ap.add_argument("--no-resolve-ab", dest="resolve_ab", action="store_false",
                    help="skip timing the three resolve launch forms")
    return ap.parse_args()


def time_frames(R, r, cams, opts, policy, warmup, steps, graph=False, start=0):
    """Warm up, then time exactly `steps` frames of the path (graph replay or
    stage by stage),
    bracketed by stream synchronizations: the frames are enqueued back to
    back (the host does not read them), as a renderer runs.  Returns
    seconds."""
    flags = R.FLAG_GRAPH if graph else 0
    for i in range(warmup):
        r.render(cams[(start + i) % len(cams)], opts, policy, flags)
    r.sync()
    t0 = time.perf_counter()
    for i in range(steps):
        r.render(cams[(start + warmup + i) % len(cams)], opts, policy, flags)
    r.sync()
    return time.perf_counter() - t0


def stage_profile(R, r, cams, opts, policy, warmup, steps):
    """The same frames once more, one at a time with per-stage HIP events
    (stages one after the other), read back after every frame: per-stage
    milliseconds and the binning statistics per frame."""
    keys = ("ms_vertex", "ms_model", "ms_setup", "ms_scan", "ms_fill", "ms_fallback", "ms_raster", "ms_shade",
            "ms_total")
    acc = {k: 0.0 for k in keys}
    ovf_tiles = cap_total = pairs = 0
    # the path frame before the first profiled one, so single-pass capacities
    # come from the path's previous frame as in the timed rounds
    r.render(cams[(warmup - 1) % len(cams)], opts, policy, R.FLAG_TIMING)
    rerenders0 = r.stats()["rerenders"]
    for i in range(steps):
        r.render(cams[(warmup + i) % len(cams)], opts, policy, R.FLAG_TIMING)
        st = r.stats()
        for k in keys:
            acc[k] += st[k]
        ovf_tiles += st["overflow_tiles"]
        cap_total += st["capacity_total"]
        pairs += st["num_bin_pairs"]
    extra = {"overflow_tiles_per_frame": ovf_tiles / steps, "capacity_entries_per_frame": cap_total / steps,
             "bin_pairs_per_frame": pairs / steps, "rerenders": r.stats()["rerenders"] - rerenders0}
    return {k: v / steps for k, v in acc.items()}, extra


def queue_of(autotune, args, width, height, tris, table=None):
    """Queue depth (margin, slack) per single-pass policy: the tuning
    table's for the frame's bucket; --margin / --slack replace the learned
    policy's."""
    q = {p: autotune.queue_depth(width, height, tris, p, table) for p in ("learned", "prevframe")}
    m, s = q["learned"]
    q["learned"] = (m if args.margin is None else args.margin, s if args.slack is None else args.slack)
    return q


def train_model(R, r, w, train_path, args):
    model = R.BinModel(r, hidden=64, seed=7)
    model.attach(r, *args.queue["learned"])
    for i, cam in enumerate(train_path):
        r.render(cam, w.opts, R.BIN_EXACT, R.FLAG_NO_SHADE)
        if i > 0:
            model.record(r)
    model.train(epochs=args.epochs, batch=8192, lr=2e-3, tau=0.95)
    return model


ROUNDS = 3


def run_policies(R, r, w, train_path, test_path, policies, args):
    """Trains the capacity model (if the learned policy is asked for) and
    times every policy on the test path: ROUNDS rounds, each timing every
    policy in turn over the same `steps` frames (so drift of the device
    clock or temperature affects all alike); a policy's figure is its median
    round.  Returns (results, train seconds, the last round's seconds per
    policy)."""
    results = {}
    model = None
    train_s = 0.0
    if "learned" in policies:
        t0 = time.perf_counter()
        model = train_model(R, r, w, train_path, args)
        train_s = time.perf_counter() - t0
    code = {"exact": R.BIN_EXACT, "learned": R.BIN_LEARNED, "prevframe": R.BIN_PREVFRAME}

    def select(p):
        if model is not None and p == "learned":
            model.attach(r, *args.queue["learned"])
        else:  # detaches the model
            r.set_queue_depth(*args.queue["prevframe"])

    rounds = {p: [] for p in policies}
    rounds_other = {p: [] for p in policies}  # the other launch form (graph replay by default)
    for _ in range(ROUNDS):
        for p in policies:
            select(p)
            rounds[p].append(time_frames(R, r, test_path, w.opts, code[p], args.warmup, args.steps,
                                         graph=args.graph))
            rounds_other[p].append(time_frames(R, r, test_path, w.opts, code[p], args.warmup, args.steps,
                                               graph=not args.graph))
    resolve_ab = None
    if getattr(args, "resolve_ab", False) and policies:
        # the resolve stage's launch forms on the headline policy, interleaved rounds
        p = "learned" if "learned" in policies else policies[0]
        select(p)
        names = {R.RESOLVE_PER_TILE: "per_tile", R.RESOLVE_PERSISTENT: "persistent",
                 R.RESOLVE_PERSISTENT_SHADE: "persistent_shade"}
        ab = {m: [] for m in names}
        for _ in range(ROUNDS):
            for m in names:
                r.set_resolve_mode(m)
                ab[m].append(time_frames(R, r, test_path, w.opts, code[p], args.warmup, args.steps,
                                         graph=args.graph))
        r.set_resolve_mode(args.resolve_mode)
        resolve_ab = {"policy": p, **{names[m]: round(sorted(v)[len(v) // 2] * 1e3 / args.steps, 4)
                                      for m, v in ab.items()}}
    for p in policies:
        select(p)
        stages, extra = stage_profile(R, r, test_path, w.opts, code[p], args.warmup, args.steps)
        secs = sorted(rounds[p])[len(rounds[p]) // 2]
        other = sorted(rounds_other[p])[len(rounds_other[p]) // 2]
        results[p] = {"ms_per_frame": secs * 1e3 / args.steps,
                      "rounds_ms_per_frame": [round(t * 1e3 / args.steps, 4) for t in rounds[p]],
                      ("eager_ms_per_frame" if args.graph else "graph_replay_ms_per_frame"):
                          round(other * 1e3 / args.steps, 4),
                      "stages_ms": {k: round(v, 4) for k, v in stages.items()}, **extra}
    if model is not None:
        model.close()
    if resolve_ab:
        results["_resolve_modes"] = resolve_ab
    return results, train_s, {p: rounds[p][-1] for p in policies}


def batch_workload(R, S, autotune, args):
    """The 256-job batch of uneven frames: each job's parameters from the
    tuning table (its size bucket), all jobs in one batched launch."""
    jobs = S.config4()
    table = autotune.load_table()
    rs = []
    spec = []
    for j in jobs:
        r = R.Renderer(j.width, j.height)
        autotune.choose(j.width, j.height, j.scene.num_triangles, table).apply(r)
        r.add_scene(j.scene)
        rs.append(r)
        spec.append((r, j.camera, j.opts, R.BIN_EXACT, 0))
    steps = max(args.steps, 10)

    def time_batch():
        for _ in range(max(args.warmup, 2)):
            R.render_batch(spec)
        for r in rs:
            r.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            R.render_batch(spec)
        for r in rs:
            r.sync()
        return (time.perf_counter() - t0) * 1e3 / steps

    # the resolve stage's launch forms (a batch takes its first job's), interleaved rounds
    forms = {R.RESOLVE_PER_TILE: "per_tile", R.RESOLVE_PERSISTENT: "persistent",
             R.RESOLVE_PERSISTENT_SHADE: "persistent_shade"}
    times = {m: [] for m in forms}
    for _ in range(ROUNDS):
        for m in forms:
            rs[0].set_resolve_mode(m)
            times[m].append(time_batch())
    med = {m: sorted(v)[len(v) // 2] for m, v in times.items()}
    rs[0].set_resolve_mode(args.resolve_mode)
    ms = med[args.resolve_mode]
    for r in rs:
        r.close()
    tris = sum(j.scene.num_triangles for j in jobs)
    px = sum(j.pixels for j in jobs)
    return {"workload": "256 jobs, 64..1024 px per axis (log-uniform), icospheres and terrains of 1k..200k tris",
            "jobs": len(jobs), "triangles": tris, "pixels": px, "ms_per_batch": round(ms, 4),
            "mtris_per_s": round(tris / (ms * 1e-3) / 1e6, 2), "gpix_per_s": round(px / (ms * 1e-3) / 1e9, 3),
            "resolve_form": forms[args.resolve_mode],
            "resolve_forms_ms_per_batch": {forms[m]: round(v, 4) for m, v in med.items()},
            "tuning": "per job-size bucket from tuning.json" if table else "defaults"}


def flythrough_workload(R, S, autotune, args):
    """The combined workload: terrain + 64 instanced icospheres (4,001,920
    tris), every frame one batched launch of the 3840x2160 main view and six
    1024^2 cube-map faces (views of one scene) with learned single-pass
    binning.  A capacity model per view is trained on a separate path first.
    A second pass over the path compares every 10th frame with the exact
    binner (vis and depth bit for bit)."""
    import numpy as np

    fly = S.config5(frames=args.fly_frames)
    ntri = fly.scene.num_triangles
    table = autotune.load_table()

    def views_for():
        owner = R.Renderer(*fly.main)
        autotune.choose(fly.main[0], fly.main[1], ntri, table).apply(owner)
        owner.add_scene(fly.scene)
        vs = [owner]
        for _ in range(6):
            v = owner.view(fly.face, fly.face)
            autotune.choose(fly.face, fly.face, ntri, table).apply(v)
            vs.append(v)
        return vs

    learned, exact = views_for(), views_for()
    models = [R.BinModel(v, hidden=64, seed=31 + k) for k, v in enumerate(learned)]
    for m, v, (_, W, H) in zip(models, learned, fly.views(0)):
        m.attach(v, *queue_of(autotune, args, W, H, ntri, table)["learned"])
    train = S.Flythrough(fly.scene, S.camera_path(S.SEED_BASE + 0x505, args.fly_train_frames), fly.main, fly.face)
    t0 = time.perf_counter()
    for f in range(len(train.path)):
        R.render_batch([(v, cam, fly.opts, R.BIN_EXACT, R.FLAG_NO_SHADE) for v, (cam, _, _) in zip(learned, train.views(f))])
        if f > 0:
            for m, v in zip(models, learned):
                m.record(v)
    for m in models:
        m.train(epochs=args.epochs, batch=8192, lr=2e-3, tau=0.95)
    train_s = time.perf_counter() - t0

    def jobs(vs, f, policy):
        return [(v, cam, fly.opts, policy, 0) for v, (cam, _, _) in zip(vs, fly.views(f))]

    for f in range(3):
        R.render_batch(jobs(learned, f, R.BIN_LEARNED))
    for v in learned:
        v.sync()
    t0 = time.perf_counter()
    for f in range(len(fly.path)):  # frames back to back, as a renderer runs
        R.render_batch(jobs(learned, f, R.BIN_LEARNED))
    for v in learned:
        v.sync()
    secs = time.perf_counter() - t0
    # second pass over the path: every frame's statistics, and every 10th
    # frame against the exact binner
    ovf = tiles = 0
    checked = mismatched = 0
    for f in range(len(fly.path)):
        R.render_batch(jobs(learned, f, R.BIN_LEARNED))
        for v in learned:
            st = v.stats()
            ovf += st["overflow_tiles"]
            tiles += st["num_tiles"]
        if f % 10 == 0:
            R.render_batch(jobs(exact, f, R.BIN_EXACT))
            for a, b in zip(learned, exact):
                va, da, _ = a.readback(rgba=False)
                vb, db, _ = b.readback(rgba=False)
                mismatched += int(not (np.array_equal(va, vb) and np.array_equal(da.view(np.uint32), db.view(np.uint32))))
            checked += 1
    for m in models:
        m.close()
    for vs in (learned, exact):
        for v in vs[1:]:
            v.close()
        vs[0].close()
    frames = len(fly.path)
    px = fly.main[0] * fly.main[1] + 6 * fly.face * fly.face
    ms = secs * 1e3 / frames
    return {"workload": f"terrain {1160}^2 + 64 icospheres s5, main 3840x2160 + six {fly.face}^2 cube faces per frame",
            "frames": frames, "views_per_frame": 7, "triangles": ntri, "pixels_per_frame": px,
            "ms_per_frame": round(ms, 4), "mtris_per_s": round(ntri * 7 / (ms * 1e-3) / 1e6, 2),
            "gpix_per_s": round(px / (ms * 1e-3) / 1e9, 3), "binning": "learned (one model per view)",
            "fallback_tiles_per_frame": round(ovf / frames, 2), "fallback_rate": round(ovf / max(tiles, 1), 5),
            "checked_frames": checked, "views_checked": checked * 7, "mismatched_views": mismatched,
            "train_seconds": round(train_s, 2)}


def read_traffic():
    """HBM bytes per k_shade launch from the committed counter summary:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by the factor
    that maps the same counter onto a plain copy of known size."""
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    vals = {}
    with open(TRAFFIC_SUMMARY) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    need = ("fetch_kib_per_dispatch", "write_kib_per_dispatch", "copy_scale_fetch", "copy_scale_write")
    if not all(k in vals for k in need):
        return None
    return int((vals["fetch_kib_per_dispatch"] * vals["copy_scale_fetch"] +
                vals["write_kib_per_dispatch"] * vals["copy_scale_write"]) * 1024)


def copy_calibration(mib):
    """Plain device copies of a known size, for scaling the HBM counters."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 5}))


def cpu_share():
    """Host threads this process may use: the GPU box allots a fixed CPU share
    per GPU and exports it as OMP_NUM_THREADS; elsewhere every core."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(int(env), n) if env.isdigit() and int(env) > 0 else n


def cpu_baseline(w, cams, frames):
    """The restated reference CPU path (the C oracle: vertex transform, setup,
    the sequential ERenderer raster loop split over row bands, shading) timed
    on the host: the median of `frames` frames of the test path on every host
    thread this process may use, and one frame on one thread."""
    import statistics

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = cpu_share()
    O.render(w.scene, cams[0], w.opts, w.width, w.height, threads=threads)  # warm caches
    times = []
    for i in range(frames):
        t0 = time.perf_counter()
        O.render(w.scene, cams[i % len(cams)], w.opts, w.width, w.height, threads=threads)
        times.append(time.perf_counter() - t0)
    dt = statistics.median(times)
    t1 = time.perf_counter()
    O.render(w.scene, cams[0], w.opts, w.width, w.height, threads=1)
    dt1 = time.perf_counter() - t1
    ntri = w.scene.num_triangles
    return {"value": round(ntri / dt / 1e6, 3), "unit": "Mtris/s", "cores": threads, "kind": "port",
            "sample": f"median of {frames} frames of the terrain test path rendered by the C oracle (restated "
                      f"reference CPU path, vis+depth+shading) on {threads} threads",
            "nproc": os.cpu_count(), "threads": threads, "median_of": frames,
            "frame_seconds": [round(t, 3) for t in times],
            "single_thread_mtris_s": round(ntri / dt1 / 1e6, 3), "cpu_model": _cpu_model()}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build renders on one GPU only; run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    import autotune
    import rast as R
    import scenes as S

    if args.only_batch:
        print(json.dumps({"batched": batch_workload(R, S, autotune, args)}))
        return
    if args.only_flythrough:
        print(json.dumps({"flythrough": flythrough_workload(R, S, autotune, args)}))
        return
    w = S.config3()
    train_path, test_path = S.train_test_paths(args.train_frames, 100)
    ntri = w.scene.num_triangles
    P = w.width * w.height
    tuned = autotune.choose(w.width, w.height, ntri)
    tuned = autotune.Params(tile=args.tile or tuned.tile, micro=tuned.micro if args.micro is None else args.micro,
                            small=tuned.small if args.small is None else args.small, big=args.big or tuned.big)
    args.queue = queue_of(autotune, args, w.width, w.height, ntri)
    r = R.Renderer(w.width, w.height)
    tuned.apply(r)
    r.set_resolve_mode(args.resolve_mode)
    r.add_scene(w.scene)
    policies = [p for p in args.policies.split(",") if p]
    results, train_s, _ = run_policies(R, r, w, train_path, test_path, policies, args)
    resolve_modes = results.pop("_resolve_modes", None)
    r.close()

    plain = None
    if not args.no_plain and "exact" in policies and "learned" in policies:
        # every triangle binned (no setup-time raster), default tile: the plain binner comparison
        pp = autotune.Params(tile=16, micro=0, small=16, big=256)
        r2 = R.Renderer(w.width, w.height)
        pp.apply(r2)
        r2.add_scene(w.scene)
        res2, _, _ = run_policies(R, r2, w, train_path, test_path, ["exact", "learned"],
                                  argparse.Namespace(**{**vars(args), "resolve_ab": False}))
        r2.close()
        plain = {"params": vars(pp), "exact_ms_per_frame": round(res2["exact"]["ms_per_frame"], 4),
                 "learned_ms_per_frame": round(res2["learned"]["ms_per_frame"], 4),
                 "learned_over_exact": round(res2["exact"]["ms_per_frame"] / res2["learned"]["ms_per_frame"], 3),
                 "stages_ms": {k: v["stages_ms"] for k, v in res2.items()},
                 "overflow_tiles_per_frame": res2["learned"]["overflow_tiles_per_frame"]}

    head = "learned" if "learned" in results else policies[0]
    ms = results[head]["ms_per_frame"]
    value = ntri / (ms * 1e-3) / 1e6

    roofline = None
    if os.path.exists(SHADING_FIXTURE):
        with open(SHADING_FIXTURE) as f:
            fx = json.load(f)
        frames = fx["frames"]
        used = [(args.warmup + i) % len(test_path) for i in range(args.steps)]
        b_shade = sum(frames[str(i)]["bytes"] for i in used) / len(used)
        t_shade = results[head]["stages_ms"]["ms_shade"] * 1e-3
        if t_shade > 0:
            ach = b_shade / t_shade / 1e9
            roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": read_traffic(), "kernel": "k_shade",
                        "algorithmic_bytes_per_launch": int(b_shade),
                        "avg_launch_ms": round(results[head]["stages_ms"]["ms_shade"], 4)}

    batch = None if args.no_batch else batch_workload(R, S, autotune, args)
    fly = None if args.no_flythrough else flythrough_workload(R, S, autotune, args)
    cpu = None if args.no_cpu else cpu_baseline(w, test_path, args.cpu_frames)
    line = {
        "metric": "Mtris/s + shaded Gpix/s at 4K, 2M-tri scene; learned vs exact binning",
        "value": round(value, 2),
        "unit": "Mtris/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural terrain, textures and camera path)",
        "config": {"workload": w.name, "triangles": ntri, "width": w.width, "height": w.height,
                   "binning": head, "params": vars(tuned),
                   "queue_depth": {k: {"margin": m, "slack": s} for k, (m, s) in args.queue.items()},
                   "camera_path_frames": 100,
                   "train_frames": args.train_frames,
                   "launch": "graph replay" if args.graph else "stage by stage (graph replay timed beside it)",
                   "timing": f"median of {ROUNDS} rounds of {args.steps} back-to-back frames, policies interleaved"},
        "gpix_per_s": round(P / (ms * 1e-3) / 1e9, 3),
        "policies": {k: {"mtris_per_s": round(ntri / (v["ms_per_frame"] * 1e-3) / 1e6, 2), **v}
                     for k, v in results.items()},
        "resolve_modes_ms_per_frame": resolve_modes,
        "plain_binner": plain,
        "batched": batch,
        "flythrough": fly,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "train_seconds": round(train_s, 2),
    }
    if "exact" in results and "learned" in results:
        line["learned_over_exact"] = round(results["exact"]["ms_per_frame"] / results["learned"]["ms_per_frame"], 3)
    if "prevframe" in results and "learned" in results:
        line["learned_over_prevframe"] = round(
            results["prevframe"]["ms_per_frame"] / results["learned"]["ms_per_frame"], 3)
    if plain and "learned" in results:
        # the learned single-pass frame against the plain exact two-pass binner
        # (every triangle binned, no setup-time raster, default tile)
        line["learned_over_plain_exact"] = round(plain["exact_ms_per_frame"] / results["learned"]["ms_per_frame"], 3)
    if roofline:
        line["shading_gpix_per_s"] = round(P / (results[head]["stages_ms"]["ms_shade"] * 1e-3) / 1e9, 2)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
