"""GPU parity at the benchmark configurations themselves, against the CPU
oracle (visibility and depth bit for bit, RGBA8 within 1/255):

  * the 256-job batch of uneven requests (64..1024 px per axis, icospheres
    and terrains of 1k..200k triangles, 512^2 textures), rendered as one
    batched launch with each job's parameters from the tuning table;
  * the combined flythrough at full size (4,001,920 triangles; a 3840x2160
    main view and six 1024^2 cube faces per frame in one batch) with learned
    single-pass binning, three frames of the path;
  * the 4K terrain with the tuned parameters and a trained learned model,
    three frames of the test path (the headline configuration of the bench).

The batched path realises the per-mesh host loop and the "persistent kernel"
TODO of CCR/CUDARenderer.cu:206-354."""
from concurrent.futures import ThreadPoolExecutor
import os

import numpy as np
import pytest

import autotune
import oracle as O
import rast as R
import scenes as S

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def assert_frame(tag, got, want):
    vis, depth, rgba = got
    ovis, odepth, orgba = want
    bad = np.count_nonzero(vis != ovis)
    assert bad == 0, f"{tag}: vis differs at {bad} pixels"
    bad = np.count_nonzero(depth.view(np.uint32) != odepth.view(np.uint32))
    assert bad == 0, f"{tag}: depth differs at {bad} pixels"
    d = np.abs(rgba.view(np.uint8).astype(np.int16) - orgba.view(np.uint8).astype(np.int16))
    if d.max() > 1:
        bad_px = np.argwhere(d > 1)[:8]
        detail = [(int(x) // 4, int(y), f"{int(rgba[y, x // 4]):08x}", f"{int(orgba[y, x // 4]):08x}",
                   int(vis[y, x // 4])) for y, x in bad_px]
        raise AssertionError(f"{tag}: rgba off by {d.max()} at {(d > 1).sum()} channels; "
                             f"(x, y, got, want, triangle): {detail}")


def test_config4_batch_of_256_uneven_jobs_matches_the_oracle():
    jobs = S.config4()
    table = autotune.load_table()
    rs = []
    try:
        spec = []
        for j in jobs:
            r = R.Renderer(j.width, j.height)
            autotune.choose(j.width, j.height, j.scene.num_triangles, table).apply(r)
            r.add_scene(j.scene)
            rs.append(r)
            spec.append((r, j.camera, j.opts, R.BIN_EXACT, 0))
        R.render_batch(spec)
        R.render_batch(spec)  # the replay of the captured batch graph
        got = [r.readback() for r in rs]
        tiles = {r.tile for r in rs}
        assert len(tiles) > 1  # the tuning table gives different tile sizes per bucket

        def oracle(j):
            v, d, c, _ = O.render(j.scene, j.camera, j.opts, j.width, j.height, threads=1)
            return v, d, c

        with ThreadPoolExecutor(THREADS) as ex:
            want = list(ex.map(oracle, jobs))
        for k, (j, g, w) in enumerate(zip(jobs, got, want)):
            assert_frame(f"job {k} ({j.name})", g, w)
    finally:
        for r in rs:
            r.close()


def make_views(fly, table, ntri):
    owner = R.Renderer(*fly.main)
    autotune.choose(fly.main[0], fly.main[1], ntri, table).apply(owner)
    owner.add_scene(fly.scene)
    vs = [owner]
    for _ in range(6):
        v = owner.view(fly.face, fly.face)
        autotune.choose(fly.face, fly.face, ntri, table).apply(v)
        vs.append(v)
    return vs


def test_config5_full_size_flythrough_learned_matches_the_oracle():
    fly = S.config5(frames=3)
    ntri = fly.scene.num_triangles
    assert ntri == 4_001_920
    table = autotune.load_table()
    vs = make_views(fly, table, ntri)
    models = [R.BinModel(v, hidden=64, seed=41 + k) for k, v in enumerate(vs)]
    try:
        for m, v in zip(models, vs):
            m.attach(v, 1.25, 4)
        train = S.Flythrough(fly.scene, S.camera_path(S.SEED_BASE + 0x505, 8), fly.main, fly.face)
        for f in range(len(train.path)):
            R.render_batch([(v, cam, fly.opts, R.BIN_EXACT, R.FLAG_NO_SHADE)
                            for v, (cam, _, _) in zip(vs, train.views(f))])
            if f > 0:
                for m, v in zip(models, vs):
                    m.record(v)
        for m in models:
            m.train(epochs=4, batch=8192, lr=2e-3, tau=0.95)
        fallback = 0
        for f in range(len(fly.path)):
            views = fly.views(f)
            R.render_batch([(v, cam, fly.opts, R.BIN_LEARNED, 0) for v, (cam, _, _) in zip(vs, views)])
            got = [v.readback() for v in vs]
            for k, ((cam, W, H), g) in enumerate(zip(views, got)):
                if f == 0 and k == 0:
                    continue  # the first frame has no previous counts: exact binning
                ovis, odepth, orgba, _ = O.render(fly.scene, cam, fly.opts, W, H, threads=THREADS)
                assert_frame(f"frame {f} view {k}", g, (ovis, odepth, orgba))
                fallback += vs[k].stats()["overflow_tiles"]
        print("fallback tiles over the checked views:", fallback)
    finally:
        for m in models:
            m.close()
        for v in vs[1:]:
            v.close()
        vs[0].close()


def test_config3_tuned_learned_headline_matches_the_oracle():
    w = S.config3()
    train_path, test_path = S.train_test_paths(24, 100)
    p = autotune.choose(w.width, w.height, w.scene.num_triangles)
    with R.Renderer(w.width, w.height) as r:
        p.apply(r)
        assert r.tile == p.tile
        r.add_scene(w.scene)
        m = R.BinModel(r, hidden=64, seed=7)
        m.attach(r, 1.25, 4)
        for i, cam in enumerate(train_path):
            r.render(cam, w.opts, R.BIN_EXACT, R.FLAG_NO_SHADE)
            if i > 0:
                m.record(r)
        m.train(epochs=8, batch=8192, lr=2e-3, tau=0.95)
        r.render(test_path[9], w.opts, R.BIN_LEARNED, R.FLAG_GRAPH)  # previous counts for the first checked frame
        single = 0
        for i in (10, 11, 12):
            cam = test_path[i]
            r.render(cam, w.opts, R.BIN_LEARNED, R.FLAG_GRAPH)
            g = r.readback()
            st = r.stats()
            single += st["capacity_total"] > 0
            ovis, odepth, orgba, _ = O.render(w.scene, cam, w.opts, w.width, w.height, threads=THREADS)
            assert_frame(f"test-path frame {i}", g, (ovis, odepth, orgba))
        assert single == 3  # every checked frame was binned in a single pass
        m.close()
