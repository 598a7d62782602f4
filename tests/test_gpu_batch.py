"""GPU: captured frame graphs, batched rendering and shared-scene views give
the frames of plain per-frame rendering, bit for bit (all three buffers),
and the plain path itself matches the oracle (tests in test_gpu_parity)."""
import numpy as np
import pytest

import oracle as O
import rast as R
import scenes as S
from scene_util import random_triangle_scene, screen_camera

pytestmark = pytest.mark.gpu


def same_frames(a, b, what):
    for name, x, y in zip(("vis", "depth", "rgba"), a, b):
        bad = np.count_nonzero(x.view(np.uint32) != y.view(np.uint32))
        assert bad == 0, f"{what}: {name} differs at {bad} pixels"


def small_terrain():
    return S.config3(n=160, tex_size=256, width=480, height=272)


@pytest.mark.parametrize("policy", [R.BIN_EXACT, R.BIN_PREVFRAME, R.BIN_LEARNED])
def test_graph_replay_matches_eager_over_moving_camera(policy):
    """One captured graph per count-buffer parity is replayed while the
    camera moves (the camera lives in the uploaded mesh table, not in the
    graph); eager work runs between the replays."""
    w = small_terrain()
    _, path = S.train_test_paths(2, 8)
    with R.Renderer(w.width, w.height) as eager, R.Renderer(w.width, w.height) as graph:
        eager.add_scene(w.scene)
        graph.add_scene(w.scene)
        models = []
        for r in (eager, graph):
            m = R.BinModel(r, hidden=16, seed=5)
            m.attach(r, 1.1, 2)
            models.append(m)
        for i, cam in enumerate(path):
            eager.render(cam, w.opts, policy)
            graph.render(cam, w.opts, policy, R.FLAG_GRAPH)
            a, b = eager.readback(), graph.readback()
            same_frames(a, b, f"frame {i}")
            sa, sb = eager.stats(), graph.stats()
            for k in ("num_visible", "num_bin_pairs", "num_big", "overflow_tiles", "overflow_entries"):
                assert sa[k] == sb[k], (i, k, sa[k], sb[k])
        for m in models:
            m.close()


def test_graph_recaptures_after_option_and_size_changes():
    w = S.config1()
    with R.Renderer(w.width, w.height) as eager, R.Renderer(w.width, w.height) as graph:
        eager.add_scene(w.scene)
        graph.add_scene(w.scene)
        steps = [
            ("plain", lambda r: None, S.Opts()),
            ("no cull", lambda r: None, S.Opts(cull_mode=S.CULL_NONE)),
            ("tile 32", lambda r: r.setup_bins(32), S.Opts()),
            ("micro 0", lambda r: r.set_tuning(0, 16), S.Opts(visualisation=S.VIS_DEPTH)),
            ("linear", lambda r: None, S.Opts(sample_state=S.SAMPLE_LINEAR)),
            ("again", lambda r: None, S.Opts()),
        ]
        for name, change, opts in steps:
            for r in (eager, graph):
                change(r)
            for _ in range(3):  # replays of the same graph, both parities
                eager.render(w.camera, opts)
                graph.render(w.camera, opts, R.BIN_EXACT, R.FLAG_GRAPH)
                same_frames(eager.readback(), graph.readback(), name)


def test_graph_frame_matches_oracle():
    w = S.config1()
    ovis, odepth, orgba, _ = O.render(w.scene, w.camera, w.opts, w.width, w.height, threads=4)
    with R.Renderer(w.width, w.height) as r:
        r.add_scene(w.scene)
        for _ in range(4):
            r.render(w.camera, w.opts, R.BIN_EXACT, R.FLAG_GRAPH)
            vis, depth, rgba = r.readback()
            np.testing.assert_array_equal(vis, ovis)
            np.testing.assert_array_equal(depth.view(np.uint32), odepth.view(np.uint32))
            d = np.abs(rgba.view(np.uint8).astype(np.int16) - orgba.view(np.uint8).astype(np.int16))
            assert d.max() <= 1


def test_view_renders_the_owners_scene():
    w = S.config1()
    cam2 = S.Camera(position=(1.5, 0.4, 2.5), forward=(-0.5, -0.1, -0.8))
    with R.Renderer(w.width, w.height) as owner, R.Renderer(320, 200) as solo:
        owner.add_scene(w.scene)
        solo.add_scene(w.scene)
        view = owner.view(320, 200)
        try:
            owner.render(w.camera, w.opts)
            view.render(cam2, w.opts)
            solo.render(cam2, w.opts)
            same_frames(view.readback(), solo.readback(), "view vs own scene")
            ovis, odepth, _, _ = O.render(w.scene, w.camera, w.opts, w.width, w.height, threads=4, want_rgba=False)
            vis, depth, _ = owner.readback()
            np.testing.assert_array_equal(vis, ovis)
            # a view has no meshes of its own, and keeps its owner alive
            with pytest.raises(R.RastError):
                view.add_mesh(w.scene.meshes[0])
            with pytest.raises(R.RastError):
                R.check(R.lib().rast_renderer_destroy(owner.h), "destroy owner")
            # meshes added to the owner later show up in the view
            extra = random_triangle_scene(3, 40).meshes[0]
            owner.add_mesh(extra)
            solo.add_mesh(extra)
            view.render(screen_camera(), S.Opts(cull_mode=S.CULL_NONE))
            solo.render(screen_camera(), S.Opts(cull_mode=S.CULL_NONE))
            same_frames(view.readback(), solo.readback(), "view after owner grew")
        finally:
            view.close()


def batch_jobs(k, n_jobs):
    """Renderers of uneven size over their own scenes plus views of one scene."""
    rs, jobs = [], []
    shared = R.Renderer(256, 160)
    shared.add_scene(S.config1().scene)
    # and a mesh without maps, in random vertex colours (the views' colour path)
    coloured = random_triangle_scene(77, 60).meshes[0]
    coloured.verts[:, 11:14] = np.random.default_rng(77).uniform(0, 1, (coloured.verts.shape[0], 3)).astype(np.float32)
    shared.add_mesh(coloured)
    rs.append(shared)
    for j in range(n_jobs):
        W = int(S.uniform(S.SEED_BASE + 0x400, 2 * j, 48, 400))
        H = int(S.uniform(S.SEED_BASE + 0x400, 2 * j + 1, 40, 300))
        if j % 3 == 0:
            r = shared.view(W, H)
            cam = S.Camera(position=(S.uniform(9, j, -2, 2), S.uniform(10, j, -1, 1), 3.0 + k * 0.1),
                           forward=(0.0, 0.0, -1.0))
        else:
            r = R.Renderer(W, H, tile=(8, 16, 32)[j % 3])
            r.add_scene(random_triangle_scene(100 + j, 50 + 40 * j))
            cam = screen_camera()
        # both sampling modes in one batch: the shading stage runs one kernel instance per mode
        opts = S.Opts(cull_mode=(S.CULL_BACK, S.CULL_NONE)[j % 2],
                      sample_state=(S.SAMPLE_POINT, S.SAMPLE_LINEAR)[(j // 2) % 2])
        rs.append(r)
        jobs.append((r, cam, opts, R.BIN_EXACT, 0))
    return rs, jobs


def test_batch_matches_one_by_one():
    rs, jobs = batch_jobs(0, 9)
    try:
        for k in range(4):  # repeated batches: cached graphs for both parities, moving cameras
            jobs = [(r, S.Camera(position=(cam.position[0], cam.position[1], cam.position[2] + 0.05 * k),
                                 forward=cam.forward), o, p, f) for r, cam, o, p, f in jobs]
            R.render_batch(jobs)
            got = [r.readback() for r, *_ in jobs]
            for (r, cam, o, p, f), g in zip(jobs, got):
                if r.owner is not None:
                    # views of one scene in a batch store its view-independent
                    # attributes once: the per-view attribute read is refused
                    with pytest.raises(R.RastError):
                        r.debug_read(R.DEBUG_VERTEX_ATTR, np.float32)
                r.render(cam, o, p)  # the same frame, one renderer at a time
                same_frames(r.readback(), g, f"batch {k}, {r.width}x{r.height}")
                if r.owner is not None:  # a frame of its own stores them whole again
                    assert r.debug_read(R.DEBUG_VERTEX_ATTR, np.float32).size > 0
    finally:
        for r, *_ in jobs:
            r.close()
        rs[0].close()


def test_batch_rejects_shared_targets():
    with R.Renderer(64, 64) as r:
        r.add_scene(random_triangle_scene(1, 10))
        with pytest.raises(R.RastError):
            R.render_batch([(r, screen_camera(), S.Opts(), R.BIN_EXACT, 0)] * 2)


def test_back_to_back_launches_keep_their_own_staging():
    """Frames and batches are enqueued without waiting for the previous one:
    each launch stages its cameras in one of two pinned blocks in turn.  Two
    sets of renderers with the same batch shape share one launch plan; three
    batches in a row (sets A, B, then A again) must each render their own
    cameras, and a renderer's unread frames are simply replaced."""
    scene = random_triangle_scene(41, 200)
    W, H = 96, 64
    cams = [S.Camera(position=(0.03 * k, -0.02 * k, 0.05 * k), forward=(0.0, 0.0, 1.0), fov_deg=90.0,
                     near_z=0.1, far_z=100.0) for k in range(6)]
    opts = S.Opts(cull_mode=S.CULL_NONE)
    A = [R.Renderer(W, H) for _ in range(2)]
    B = [R.Renderer(W, H) for _ in range(2)]
    try:
        for r in A + B:
            r.add_scene(scene)
        R.render_batch([(A[0], cams[0], opts, R.BIN_EXACT, 0), (A[1], cams[1], opts, R.BIN_EXACT, 0)])
        R.render_batch([(B[0], cams[2], opts, R.BIN_EXACT, 0), (B[1], cams[3], opts, R.BIN_EXACT, 0)])
        R.render_batch([(A[0], cams[4], opts, R.BIN_EXACT, 0), (A[1], cams[5], opts, R.BIN_EXACT, 0)])
        for r, cam in zip(A + B, (cams[4], cams[5], cams[2], cams[3])):
            vis, depth, rgba = r.readback()
            ovis, odepth, orgba, _ = O.render(scene, cam, opts, W, H, threads=4)
            np.testing.assert_array_equal(vis, ovis)
            np.testing.assert_array_equal(depth.view(np.uint32), odepth.view(np.uint32))
        # single frames: five in a row (both staging blocks twice), only the last read
        r = A[0]
        for k in range(5):
            r.render(cams[k], opts, R.BIN_EXACT, R.FLAG_GRAPH if k % 2 else 0)
        vis, depth, _ = r.readback()
        ovis, odepth, _, _ = O.render(scene, cams[4], opts, W, H, threads=4, want_rgba=False)
        np.testing.assert_array_equal(vis, ovis)
        np.testing.assert_array_equal(depth.view(np.uint32), odepth.view(np.uint32))
    finally:
        for r in A + B:
            r.close()


@pytest.mark.parametrize("mode", [R.RESOLVE_PERSISTENT, R.RESOLVE_PERSISTENT_SHADE])
def test_persistent_resolve_gives_the_per_tile_frames(mode):
    """The persistent resolve kernels (one counter over every job's tiles;
    with shading in the resolve pass) render the frames of the one-workgroup-
    per-tile launch, bit for bit: a batch of uneven jobs of all three tile
    sizes (both sampling modes: shading stays a pass of its own) and of one
    sampling mode without views of one scene (shaded in the resolve pass;
    views sharing a scene keep a shading pass of their own), then single
    frames of the terrain with every binning policy over a moving camera."""
    for mixed in (True, False):
        rs, all_jobs = batch_jobs(1, 9)
        jobs = all_jobs
        try:
            if not mixed:
                jobs = [(r, c, S.Opts(cull_mode=o.cull_mode, sample_state=S.SAMPLE_LINEAR), p, f)
                        for r, c, o, p, f in jobs if r.owner is None]
            R.render_batch(jobs)
            want = [r.readback() for r, *_ in jobs]
            jobs[0][0].set_resolve_mode(mode)  # a batch takes its first job's form
            for k in range(2):
                R.render_batch(jobs)
                for (r, *_), w in zip(jobs, want):
                    same_frames(r.readback(), w, f"mixed={mixed} batch {k}, {r.width}x{r.height}")
        finally:
            for r, *_ in all_jobs:
                r.close()
            rs[0].close()
    w = small_terrain()
    _, path = S.train_test_paths(2, 4)
    for tile in (8, 16, 32):
        with R.Renderer(w.width, w.height, tile=tile) as a, R.Renderer(w.width, w.height, tile=tile) as b:
            a.add_scene(w.scene)
            b.add_scene(w.scene)
            b.set_resolve_mode(mode)
            models = []
            for r in (a, b):
                m = R.BinModel(r, hidden=16, seed=5)
                m.attach(r, 1.1, 2)
                models.append(m)
            for policy in (R.BIN_EXACT, R.BIN_LEARNED, R.BIN_PREVFRAME):
                for i, cam in enumerate(path):
                    a.render(cam, w.opts, policy)
                    b.render(cam, w.opts, policy, R.FLAG_GRAPH if i % 2 else 0)
                    same_frames(a.readback(), b.readback(), f"tile {tile} policy {policy} frame {i}")
                    sa, sb = a.stats(), b.stats()
                    for key in ("num_visible", "num_bin_pairs", "overflow_tiles"):
                        assert sa[key] == sb[key], (tile, policy, i, key)
            for m in models:
                m.close()
