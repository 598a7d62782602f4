"""Writes tests/golden/golden.json: SHA-256 digests of the oracle's frames on
the 640x480 UV-sphere workload and on seeded random-triangle scenes, plus
digests of the deterministic scene generators' outputs.  The digests pin the
oracle and the generators against drift (compiler flags, refactors, another
host).  Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "custom-gpgpu-rasterizer_amd"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
import scenes as S  # noqa: E402
from scene_util import random_triangle_scene, screen_camera  # noqa: E402


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def frames():
    out = {}
    w = S.config1()
    for vis_mode, name in ((S.VIS_PBR, "pbr"), (S.VIS_DEPTH, "depth"), (S.VIS_NORMAL, "normal")):
        for sample, sname in ((S.SAMPLE_POINT, "point"), (S.SAMPLE_LINEAR, "linear")):
            if vis_mode != S.VIS_PBR and sample == S.SAMPLE_LINEAR:
                continue
            opts = S.Opts(visualisation=vis_mode, sample_state=sample)
            vis, depth, rgba, st = O.render(w.scene, w.camera, opts, w.width, w.height, threads=4, want_stats=True)
            out[f"uvsphere_{name}_{sname}"] = {"vis": digest(vis), "depth": digest(depth), "rgba": digest(rgba),
                                               "stats": st}
    for seed in range(3):
        vis, depth, rgba, _ = O.render(random_triangle_scene(seed, 200), screen_camera(), S.Opts(cull_mode=S.CULL_NONE),
                                       160, 96, threads=2)
        out[f"random{seed}"] = {"vis": digest(vis), "depth": digest(depth), "rgba": digest(rgba)}
    return out


def generators():
    out = {}
    m = S.uv_sphere(64, 32)
    out["uv_sphere_64x32"] = {"verts": digest(m.verts), "indices": digest(m.indices)}
    m = S.icosphere(5)
    out["icosphere_5"] = {"verts": digest(m.verts), "indices": digest(m.indices)}
    m = S.terrain(64)
    out["terrain_64"] = {"verts": digest(m.verts), "indices": digest(m.indices)}
    out["checker_256"] = digest(S.checker())
    out["noise_128"] = [digest(S.noise_texture(128, 128, 7, k)) for k in range(4)]
    path = S.camera_path(S.SEED_BASE + 0x301, 10)
    out["camera_path"] = digest(np.array([c.position + c.forward for c in path], np.float32))
    out["world_matrix"] = digest(S.world_matrix((1, 2, 3), (0.1, 0.2, 0.3), 1.25))
    return out


def main():
    data = {"frames": frames(), "generators": generators()}
    with open(os.path.join(ROOT, "tests", "golden", "golden.json"), "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print("wrote tests/golden/golden.json")


if __name__ == "__main__":
    main()
