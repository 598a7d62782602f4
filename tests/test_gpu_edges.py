"""GPU edge cases against the oracle, through the C ABI: frames with nothing
to draw, non-finite, degenerate and behind-the-camera vertices, framebuffers
of one pixel, extreme aspect ratios and sizes that are not multiples of the
tile or of the 64x4 shading strip, and exact depth ties (D2: the later
triangle wins).  Visibility and depth bitwise, RGBA8 within 1/255."""
import numpy as np
import pytest

import oracle as O
import rast as R
import scenes as S
from scene_util import mesh, random_triangle_scene, screen_camera
from test_gpu_parity import assert_bitwise, assert_rgba_close, check_frame

pytestmark = pytest.mark.gpu


def test_scene_outside_the_frustum_gives_a_clear_frame():
    # every triangle right of the view (x > w after projection): all culled
    far_right = [(40.0 + i, 0.1 * i, -3.0 - 0.01 * i) for i in range(30)]
    scene = S.Scene([mesh(far_right, [(3 * t, 3 * t + 1, 3 * t + 2) for t in range(10)])])
    st = check_frame(scene, screen_camera(), S.Opts(), 96, 64)
    assert st["num_visible"] == 0


def test_culling_every_triangle_gives_a_clear_frame():
    scene = random_triangle_scene(3, 40)  # these triangles all face the camera
    ovis = {c: O.render(scene, screen_camera(), S.Opts(cull_mode=c), 96, 64, threads=4, want_rgba=False)[0]
            for c in (S.CULL_BACK, S.CULL_FRONT)}
    assert (ovis[S.CULL_BACK] != R.VIS_NONE).any() and (ovis[S.CULL_FRONT] == R.VIS_NONE).all()
    for cull in (S.CULL_BACK, S.CULL_FRONT):
        check_frame(scene, screen_camera(), S.Opts(cull_mode=cull), 96, 64)


def test_non_finite_degenerate_and_behind_camera_vertices():
    base = random_triangle_scene(5, 120).meshes[0]
    v = base.verts.copy()
    v[0, 0] = np.nan                   # triangle 0: NaN x
    v[4, 1] = np.inf                   # triangle 1: infinite y
    v[8, 2] = -np.inf                  # triangle 2: infinite z
    v[9:12, :3] = v[9, :3]             # triangle 3: all three vertices equal (zero area)
    v[13, :3] = 0.5 * (v[12, :3] + v[14, :3])  # triangle 4: collinear (zero area up to rounding)
    v[15, 2] = 2.0                     # triangle 5: one vertex behind the camera
    v[18:21, 2] = [1.5, -2.0, 3.0]     # triangle 6: two behind, one in front
    v[21, :2] = [3e30, -3e30]          # triangle 7: huge but finite coordinates
    scene = S.Scene([S.Mesh(verts=v, indices=base.indices)])
    for cull in (S.CULL_BACK, S.CULL_NONE):
        check_frame(scene, screen_camera(), S.Opts(cull_mode=cull), 160, 96)


@pytest.mark.parametrize("size", [(1, 1), (1, 37), (37, 1), (65, 5), (127, 131), (16384, 4), (4, 4100)])
def test_ragged_and_extreme_framebuffer_sizes(size):
    W, H = size
    # random triangles in front of a backdrop quad that covers every pixel
    backdrop = mesh([(-1e6, -1e6, -20), (1e6, -1e6, -20), (1e6, 1e6, -20), (-1e6, 1e6, -20)], [(0, 1, 2), (0, 2, 3)])
    scene = S.Scene(random_triangle_scene(9, 200, spread=1.4).meshes + [backdrop])
    for tile in (8, 32):
        check_frame(scene, screen_camera(), S.Opts(cull_mode=S.CULL_NONE), W, H, tile=tile)


def test_exact_depth_ties_go_to_the_later_triangle():
    # three copies of the same triangles: equal depth at every pixel, the last copy wins
    one = random_triangle_scene(12, 30).meshes[0]
    n = one.verts.shape[0]
    verts = np.concatenate([one.verts] * 3)
    idx = np.concatenate([one.indices + k * n for k in range(3)]).astype(np.uint32)
    scene = S.Scene([S.Mesh(verts=verts, indices=idx)])
    ovis, _, _, _ = O.render(scene, screen_camera(), S.Opts(cull_mode=S.CULL_NONE), 128, 96, threads=4,
                             want_rgba=False)
    shown = ovis[ovis != R.VIS_NONE]
    assert shown.size and (shown >= 2 * (one.indices.size // 3)).all()  # only the last copy is visible
    for tile in (8, 16, 32):
        for micro in (0, 64):
            with R.Renderer(128, 96, tile=tile) as r:
                r.set_tuning(micro, 16)
                r.add_scene(scene)
                r.render(screen_camera(), S.Opts(cull_mode=S.CULL_NONE))
                vis, depth, _ = r.readback()
            assert_bitwise(f"vis tile {tile} micro {micro}", vis, ovis)


def test_learned_binning_with_non_finite_vertices_gives_the_exact_frame():
    """The projected-vertex density count (learned capacities) skips
    vertices off screen, behind the camera or non-finite; the frames equal
    the exact binner's bit for bit."""
    base = random_triangle_scene(21, 400).meshes[0]
    v = base.verts.copy()
    v[::37, 0] = np.nan
    v[5::41, 1] = np.inf
    v[7::43, 2] = 2.5  # behind the camera
    scene = S.Scene([S.Mesh(verts=v, indices=base.indices)])
    cams = [S.Camera(position=(0.02 * k, 0.0, 0.0), forward=(0.0, 0.0, 1.0), fov_deg=90.0, near_z=0.1, far_z=100.0)
            for k in range(4)]
    opts = S.Opts(cull_mode=S.CULL_NONE)
    with R.Renderer(200, 120) as ex, R.Renderer(200, 120) as lr:
        for r in (ex, lr):
            r.set_tuning(0, 16)  # bin every triangle
            r.add_scene(scene)
        model = R.BinModel(lr, hidden=16, seed=4)
        model.attach(lr, 1.0, 0)
        for i, cam in enumerate(cams):
            lr.render(cam, opts, R.BIN_EXACT)
            if i > 0:  # features need a previous frame
                model.record(lr)
        model.train(epochs=2, batch=256)
        for cam in cams:
            ex.render(cam, opts, R.BIN_EXACT)
            lr.render(cam, opts, R.BIN_LEARNED)
            a, b = ex.readback(), lr.readback()
            assert_bitwise("vis", b[0], a[0])
            assert_bitwise("depth", b[1], a[1])
            np.testing.assert_array_equal(b[2], a[2])
        model.close()


def test_culling_survivor_counts_match_the_oracle_in_frames_and_batches():
    """num_visible is summed from per-setup-block counts when the counters
    are published (by the shading kernel, or by k_publish for unshaded
    frames): check it against the oracle's survivors for multi-mesh frames,
    shaded and unshaded, alone and in a batch with a view."""
    def scene(seed):
        return S.Scene([random_triangle_scene(seed + k, 300 + 157 * k).meshes[0] for k in range(3)])

    def want(sc, cam, cull, W, H):
        flags, _ = O.setup(sc, cam, cull, W, H)
        return int(flags.astype(bool).sum())

    cam = screen_camera()
    a, b = scene(31), scene(41)
    with R.Renderer(160, 96) as ra, R.Renderer(200, 120) as rb:
        ra.add_scene(a)
        rb.add_scene(b)
        for flags in (0, R.FLAG_NO_SHADE, R.FLAG_GRAPH):
            for cull in (S.CULL_BACK, S.CULL_NONE):
                ra.render(cam, S.Opts(cull_mode=cull), R.BIN_EXACT, flags)
                assert ra.stats()["num_visible"] == want(a, cam, cull, 160, 96), (flags, cull)
        v = ra.view(96, 64)
        try:
            jobs = [(ra, cam, S.Opts(), R.BIN_EXACT, 0), (rb, cam, S.Opts(cull_mode=S.CULL_NONE), R.BIN_EXACT, 0),
                    (v, cam, S.Opts(), R.BIN_EXACT, R.FLAG_NO_SHADE)]
            R.render_batch(jobs)
            assert ra.stats()["num_visible"] == want(a, cam, S.CULL_BACK, 160, 96)
            assert rb.stats()["num_visible"] == want(b, cam, S.CULL_NONE, 200, 120)
            assert v.stats()["num_visible"] == want(a, cam, S.CULL_BACK, 96, 64)
        finally:
            v.close()
