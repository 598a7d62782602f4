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

and, for single-pass binning, the per-tile queue depth, capacity =
ceil(predicted * margin) + slack (rast_renderer_set_binmodel).

Jobs are bucketed by log2 of their pixel count and triangle count; every
candidate is timed on all jobs of a bucket rendered as one batch (captured
graph, so host launch costs are excluded), and the fastest wins.  The
result is a JSON table that `choose()` reads back.  Output never depends on
these parameters (the GPU tests check tile, micro and small invariance).
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
MICROS = (0, 16, 64, 128, 256)
SMALLS = (16, 64, 192)
BIGS = (64, 256, 1024)


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


def tune(R, jobs_by_bucket: Dict[str, list], cands: Sequence[Params], reps: int = 5, log=None) -> Dict:
    out = {"device": R.device_info(0)["name"], "buckets": {}}
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
    args = ap.parse_args()
    cands = candidates(micros=(0, 64, 128), smalls=(16, 192), bigs=(256,)) if args.quick else candidates()
    log = lambda s: print(s, flush=True)  # noqa: E731
    by_bucket: Dict[str, list] = {}
    keep = []
    names = args.workloads.split(",")
    if "terrain4k" in names:
        w = S.config3()
        r = R.Renderer(w.width, w.height)
        r.add_scene(w.scene)
        keep.append(r)
        _, path = S.train_test_paths(1, 4)
        by_bucket.setdefault(bucket_of(w.width * w.height, w.scene.num_triangles), []).append(
            (r, path[0], w.opts, R.BIN_EXACT, 0))
    if "batch256" in names:
        for job in S.config4():
            r = R.Renderer(job.width, job.height)
            r.add_scene(job.scene)
            keep.append(r)
            by_bucket.setdefault(bucket_of(job.pixels, job.scene.num_triangles), []).append(
                (r, job.camera, job.opts, R.BIN_EXACT, 0))
    if "flythrough" in names:
        # the main view and the six cube faces of the combined workload (views of one scene)
        fly = S.config5(frames=1)
        owner = R.Renderer(*fly.main)
        owner.add_scene(fly.scene)
        keep.append(owner)
        views = [owner] + [owner.view(fly.face, fly.face) for _ in range(6)]
        keep[len(keep):len(keep)] = views[1:]
        for v, (cam, W, H) in zip(views, fly.views(0)):
            by_bucket.setdefault(bucket_of(W * H, fly.scene.num_triangles), []).append(
                (v, cam, fly.opts, R.BIN_EXACT, 0))
    t0 = time.time()
    table = tune(R, by_bucket, cands, args.reps, log)
    table["candidates"] = len(cands)
    table["seconds"] = round(time.time() - t0, 1)
    for r in reversed(keep):  # views before the renderer whose meshes they draw
        r.close()
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    _main()
