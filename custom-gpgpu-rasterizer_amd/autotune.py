"""Autotuner of the binning parameters per job-size bucket.

The reference fixes its bin grid and queue depth at compile time
(CCR/DEFINES.h:3-35, BINMULTIPLIER / BINQUEUEMAXSIZE; the bin grid derived
from the resolution in CCR/main.cpp:372-389).  Here they are run-time
parameters of the C ABI, and this module measures which values render a
given class of frames fastest on the device at hand:

  tile        screen tile edge (8, 16 or 32 pixels)           rast_setup_bins
  big         tile span from which a triangle goes to the
              screen-wide list instead of per-tile queues     rast_setup_bins
  micro       bounding-box pixel count up to which a
              triangle is rasterized at setup, unbinned       rast_set_tuning
  small       tile-clipped pixel count from which the
              resolve pass rasterizes a triangle per pixel    rast_set_tuning

and, for the single-pass policies, the per-tile queue depth, capacity =
ceil(predicted * margin) + slack (rast_renderer_set_binmodel), swept over
margin x slack for the learned model and for the previous-frame heuristic
once the parameters above are fixed.  Queue depth matters only when the
frame differs from the one the capacities were predicted from, so it is
timed over consecutive frames of a camera path: the workload's own path
where it has one, else a short sideways pan from each job's camera.

Jobs are bucketed by log2 of their pixel count and triangle count; every
candidate is timed on all jobs of a bucket rendered as one batch (captured
graph, so host launch costs are excluded), and the fastest wins.  The
result is a JSON table that `choose()` and `queue_depth()` read back.
Output never depends on these parameters (the GPU tests check tile, micro
and small invariance, and single-pass frames against exact ones).
"""
from __future__ import annotations

import argparse
import itertools
import json
import math
import os
import sys
import time
from dataclasses import asdict, dataclass
from typing import Dict, List, Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(HERE, "tuning.json")

TILES = (8, 16, 32)
MICROS = (0, 16, 32, 64, 128, 256)
SMALLS = (16, 64, 192, 512, 1024)
BIGS = (64, 256, 1024)
QUEUE_MARGINS = (1.0, 1.1, 1.25, 1.5)
QUEUE_SLACKS = (0, 4, 16)
DEFAULT_QUEUE = {"learned": (1.25, 4), "prevframe": (1.0, 0)}


@dataclass
class Params:
    tile: int = 16
    micro: int = 16
    small: int = 16
    big: int = 256

    def apply(self, r) -> None:
        r.setup_bins(self.tile, self.big)
        r.set_tuning(self.micro, self.small)


def bucket_of(pixels: int, tris: int) -> str:
    """Bucket key: log2 of the pixel count in steps of 2 (4x area), log2 of
    the triangle count in steps of 2."""
    pb = max(0, int(math.log2(max(pixels, 1))) // 2)
    tb = max(0, int(math.log2(max(tris, 1))) // 2)
    return f"p{pb}_t{tb}"


def load_table(path: str = TABLE_PATH) -> Dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def choose(width: int, height: int, tris: int, table: Optional[Dict] = None) -> Params:
    """Tuned parameters of the bucket of a width x height frame of `tris`
    triangles; the nearest tuned bucket when this one was not tuned, the
    defaults when nothing was."""
    table = load_table() if table is None else table
    buckets = table.get("buckets", {})
    if not buckets:
        return Params()
    key = bucket_of(width * height, tris)
    if key in buckets:
        return Params(**buckets[key]["params"])

    def dist(k: str) -> int:
        p, t = k[1:].split("_t")
        q, u = key[1:].split("_t")
        return abs(int(p) - int(q)) + abs(int(t) - int(u))

    return Params(**buckets[min(buckets, key=dist)]["params"])


def queue_depth(width: int, height: int, tris: int, policy: str = "learned",
                table: Optional[Dict] = None) -> tuple:
    """(margin, slack) of a single-pass policy ("learned" or "prevframe")
    for the bucket of a width x height frame of `tris` triangles: the tuned
    pair of that bucket, else of the nearest bucket that has one, else the
    default."""
    table = load_table() if table is None else table
    tuned = {k: b["queue"][policy] for k, b in table.get("buckets", {}).items()
             if policy in b.get("queue", {})}
    if not tuned:
        return DEFAULT_QUEUE[policy]
    key = bucket_of(width * height, tris)

    def dist(k: str) -> int:
        p, t = k[1:].split("_t")
        q, u = key[1:].split("_t")
        return abs(int(p) - int(q)) + abs(int(t) - int(u))

    q = tuned[key] if key in tuned else tuned[min(tuned, key=dist)]
    return float(q["margin"]), int(q["slack"])


def candidates(tiles=TILES, micros=MICROS, smalls=SMALLS, bigs=BIGS) -> List[Params]:
    return [Params(t, m, s, b) for t, m, s, b in itertools.product(tiles, micros, smalls, bigs)]


def time_batch(R, jobs, reps: int) -> float:
    """Seconds per batch of `jobs` [(renderer, camera, opts, policy, flags)]."""
    R.render_batch(jobs)  # capture (or re-capture) and first run
    for r, *_ in jobs:
        r.sync()
    best = float("inf")
    for _ in range(2):
        t0 = time.perf_counter()
        for _ in range(reps):
            R.render_batch(jobs)
        for r, *_ in jobs:
            r.sync()
        best = min(best, (time.perf_counter() - t0) / reps)
    return best


def tune_bucket(R, jobs, cands: Sequence[Params], reps: int = 5, log=None):
    """Fastest candidate for one bucket's jobs; returns (params, ms, all timings)."""
    timings = []
    for p in cands:
        for r, *_ in jobs:
            p.apply(r)
        ms = time_batch(R, jobs, reps) * 1e3
        timings.append((ms, p))
        if log:
            log(f"  {asdict(p)}: {ms:.3f} ms")
    timings.sort(key=lambda x: x[0])
    return timings[0][1], timings[0][0], timings


def device_label(R) -> Dict:
    info = R.device_info(0)
    # the marketing name is empty on some ROCm releases; the ISA name is not
    return {"device": info["name"] or info["arch"], "arch": info["arch"], "compute_units": info["compute_units"]}


def pan_path(cam, frames: int = 6, step: float = 0.01) -> list:
    """A sideways pan from `cam`: each frame moves the camera by `step`
    times its distance from the origin, perpendicular to its view."""
    import numpy as np
    from dataclasses import replace

    p = np.array(cam.position, np.float64)
    f = np.array(cam.forward, np.float64)
    right = np.cross(np.array([0.0, 1.0, 0.0]), f)
    if np.linalg.norm(right) < 1e-6:
        right = np.array([1.0, 0.0, 0.0])
    right /= np.linalg.norm(right)
    d = step * max(np.linalg.norm(p), 1e-3)
    return [replace(cam, position=tuple(float(v) for v in p + k * d * right)) for k in range(frames)]


def time_path(R, jobs, paths, policy: int, frames: int, reps: int) -> float:
    """Milliseconds per batch over consecutive frames of each job's path
    (every job in one batched launch per frame, frames back to back)."""
    def frame(k):
        return [(r, path[k % len(path)], opts, policy, 0) for (r, _, opts, *_), path in zip(jobs, paths)]

    R.render_batch(frame(0))  # capture
    best = float("inf")
    for _ in range(reps):
        R.render_batch(frame(1))  # the previous frame of the first timed one
        for r, *_ in jobs:
            r.sync()
        t0 = time.perf_counter()
        for k in range(frames):
            R.render_batch(frame(2 + k))
        for r, *_ in jobs:
            r.sync()
        best = min(best, (time.perf_counter() - t0) / frames)
    return best * 1e3


def train_bucket_model(R, jobs, paths, frames: int = 24, log=None):
    """One capacity model for a bucket's jobs, trained on their exact
    counts along their paths (a model serves every renderer it is attached
    to)."""
    m = R.BinModel(jobs[0][0], hidden=64, seed=11)
    for r, *_ in jobs:
        m.attach(r, *DEFAULT_QUEUE["learned"])
    for k in range(frames):
        R.render_batch([(r, path[k % len(path)], opts, R.BIN_EXACT, R.FLAG_NO_SHADE)
                        for (r, _, opts, *_), path in zip(jobs, paths)])
        if k > 0:
            for r, *_ in jobs:
                m.record(r)
    loss = m.train(epochs=4, batch=8192, lr=2e-3, tau=0.95)
    if log:
        log(f"  model trained, loss {loss:.4f}")
    return m


def tune_queue(R, jobs, paths, frames: int = 16, reps: int = 5, log=None) -> Dict:
    """Fastest (margin, slack) per single-pass policy for one bucket;
    exact two-pass binning timed alongside on the same frames."""
    out = {"exact_ms_per_batch": round(time_path(R, jobs, paths, R.BIN_EXACT, frames, reps), 4)}
    model = train_bucket_model(R, jobs, paths, log=log)
    try:
        for name, code in (("learned", R.BIN_LEARNED), ("prevframe", R.BIN_PREVFRAME)):
            timings = []
            for margin, slack in itertools.product(QUEUE_MARGINS, QUEUE_SLACKS):
                for r, *_ in jobs:
                    if name == "learned":
                        model.attach(r, margin, slack)
                    else:
                        r.set_queue_depth(margin, slack)
                ms = time_path(R, jobs, paths, code, frames, reps)
                timings.append((ms, margin, slack))
                if log:
                    log(f"  {name} margin {margin} slack {slack}: {ms:.3f} ms")
            ms, margin, slack = min(timings)
            out[name] = {"margin": margin, "slack": slack, "ms_per_batch": round(ms, 4)}
    finally:
        for r, *_ in jobs:
            model.detach(r)
        model.close()
    return out


def tune(R, jobs_by_bucket: Dict[str, list], cands: Sequence[Params], reps: int = 5, log=None) -> Dict:
    out = {**device_label(R), "buckets": {}}
    for key, jobs in sorted(jobs_by_bucket.items()):
        if log:
            log(f"bucket {key}: {len(jobs)} jobs")
        best, ms, timings = tune_bucket(R, jobs, cands, reps, log)
        default_ms = next((t for t, p in timings if p == Params()), None)
        out["buckets"][key] = {"params": asdict(best), "ms_per_batch": round(ms, 4), "jobs": len(jobs),
                               "default_ms_per_batch": round(default_ms, 4) if default_ms else None}
    return out


def _main():
    sys.path.insert(0, HERE)
    import rast as R
    import scenes as S

    ap = argparse.ArgumentParser(description="tune binning parameters on the GPU and write the table")
    ap.add_argument("--workloads", default="terrain4k,batch256,flythrough")
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="a reduced candidate set")
    ap.add_argument("--queue-only", action="store_true",
                    help="keep the table's binning parameters, sweep only the queue depth")
    ap.add_argument("--no-queue", action="store_true", help="skip the queue-depth sweep")
    args = ap.parse_args()
    cands = candidates(micros=(0, 64, 128), smalls=(16, 192), bigs=(256,)) if args.quick else candidates()
    log = lambda s: print(s, flush=True)  # noqa: E731
    by_bucket: Dict[str, list] = {}
    paths: Dict[str, list] = {}  # per bucket, each job's camera path (queue-depth sweep)
    keep = []

    def add(key, job, path):
        by_bucket.setdefault(key, []).append(job)
        paths.setdefault(key, []).append(path)

    names = args.workloads.split(",")
    if "terrain4k" in names:
        w = S.config3()
        r = R.Renderer(w.width, w.height)
        r.add_scene(w.scene)
        keep.append(r)
        _, path = S.train_test_paths(1, 24)
        add(bucket_of(w.width * w.height, w.scene.num_triangles), (r, path[0], w.opts, R.BIN_EXACT, 0), path)
    if "batch256" in names:
        for job in S.config4():
            r = R.Renderer(job.width, job.height)
            r.add_scene(job.scene)
            keep.append(r)
            add(bucket_of(job.pixels, job.scene.num_triangles), (r, job.camera, job.opts, R.BIN_EXACT, 0),
                pan_path(job.camera, 24))
    if "flythrough" in names:
        # the main view and the six cube faces of the combined workload (views of one scene)
        fly = S.config5(frames=24)
        owner = R.Renderer(*fly.main)
        owner.add_scene(fly.scene)
        keep.append(owner)
        views = [owner] + [owner.view(fly.face, fly.face) for _ in range(6)]
        keep[len(keep):len(keep)] = views[1:]
        per_frame = [fly.views(f) for f in range(len(fly.path))]
        for k, (v, (cam, W, H)) in enumerate(zip(views, per_frame[0])):
            add(bucket_of(W * H, fly.scene.num_triangles), (v, cam, fly.opts, R.BIN_EXACT, 0),
                [pf[k][0] for pf in per_frame])
    t0 = time.time()
    if args.queue_only:
        table = load_table(args.out)
        table.update(device_label(R))
    else:
        table = tune(R, by_bucket, cands, args.reps, log)
        table["candidates"] = len(cands)
    if not args.no_queue:
        for key, jobs in sorted(by_bucket.items()):
            p = Params(**table["buckets"][key]["params"]) if key in table["buckets"] else Params()
            for r, *_ in jobs:
                p.apply(r)
            log(f"bucket {key}: queue depth, {len(jobs)} jobs, {asdict(p)}")
            q = tune_queue(R, jobs, paths[key], log=log)
            table["buckets"].setdefault(key, {"params": asdict(p)})["queue"] = q
        table["queue_candidates"] = {"margin": list(QUEUE_MARGINS), "slack": list(QUEUE_SLACKS)}
    table["seconds"] = round(time.time() - t0, 1)
    for r in reversed(keep):  # views before the renderer whose meshes they draw
        r.close()
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    _main()
