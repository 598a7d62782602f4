"""CPU checks that pin the oracle (the C restatement of the reference's CPU
path) without the reference running: closed forms, invariants, brute force
on tiny inputs and library-routine sanity bounds.  No GPU needed."""
import math

import numpy as np
import pytest

import oracle as O
import scenes as S
from scene_util import mesh, random_triangle_scene, screen_camera

W = H = 64


def exact_pixel_camera():
    """Camera at the origin looking down -z with FOV 90 on a square frame:
    proj[0][0] = proj[1][1] = 1 exactly, and a vertex at view depth z = -1
    has clip w = 1 - z = 2 (the reference projection keeps [3][3] = 1), so
    screen x = 16 * x_view + 32 exactly for x_view = k / 16 (64x64 frame)."""
    return S.Camera(position=(0.0, 0.0, 0.0), forward=(0.0, 0.0, 1.0), fov_deg=90.0, near_z=0.1, far_z=100.0)


def at_pixel(sx, sy, z=-1.0):
    """World point that projects exactly onto screen pixel (sx, sy) of the 64x64 frame."""
    return ((sx - 32) / 16.0, (32 - sy) / 16.0, z)


def test_exact_pixel_camera_projection():
    fr = O.camera_frame(exact_pixel_camera(), W, H)
    assert fr.proj[0] == 1.0 and fr.proj[5] == 1.0
    assert fr.proj[15] == 1.0  # Identity()'s [3][3] survives (CCR/Camera.cpp:97-108)
    flags, rec = O.setup(S.Scene([mesh([at_pixel(10, 10), at_pixel(10, 30), at_pixel(30, 10)], [(0, 1, 2)])]),
                         exact_pixel_camera(), S.CULL_NONE, W, H)
    sx = rec[0, :6].view(np.float32)
    assert list(sx) == [10.0, 10.0, 10.0, 30.0, 30.0, 10.0]


@pytest.mark.parametrize("n", [1, 2, 5, 17, 30])
def test_right_triangle_covers_closed_form_pixel_count(n):
    # legs of n pixels along +x and +y (screen y down), counter-clockwise on screen:
    # inclusive edges -> the (n+1)(n+2)/2 lattice points of the closed triangle
    x0, y0 = 12, 12
    tri = mesh([at_pixel(x0, y0), at_pixel(x0, y0 + n), at_pixel(x0 + n, y0)], [(0, 1, 2)])
    vis, depth, _, _ = O.render(S.Scene([tri]), exact_pixel_camera(), S.Opts(cull_mode=S.CULL_NONE), W, H,
                                want_rgba=False)
    covered = np.argwhere(vis != O.VIS_NONE)
    assert len(covered) == (n + 1) * (n + 2) // 2
    ys, xs = covered[:, 0] - y0, covered[:, 1] - x0
    assert (xs >= 0).all() and (ys >= 0).all() and (xs + ys <= n).all()


def test_opposite_winding_covers_nothing():
    # all three edge functions must be >= 0 (CCR/CUDAROPs.cu:128-145): the
    # clockwise copy of a covering triangle covers no pixel
    tri = mesh([at_pixel(12, 12), at_pixel(30, 12), at_pixel(12, 30)], [(0, 1, 2)])
    vis, _, _, _ = O.render(S.Scene([tri]), exact_pixel_camera(), S.Opts(cull_mode=S.CULL_NONE), W, H,
                            want_rgba=False)
    assert (vis == O.VIS_NONE).all()


def test_full_screen_quad_covers_every_pixel_and_depth_is_planar():
    quad = mesh([at_pixel(-4, -4), at_pixel(-4, 70), at_pixel(70, 70), at_pixel(70, -4)], [(0, 1, 2), (0, 2, 3)])
    vis, depth, _, _ = O.render(S.Scene([quad]), exact_pixel_camera(), S.Opts(cull_mode=S.CULL_NONE), W, H,
                                want_rgba=False)
    assert (vis != O.VIS_NONE).all()
    A, B = 100.0 / (0.1 - 100.0), (100.0 * 0.1) / (0.1 - 100.0)
    z = (A * -1.0 + B) / 2.0  # (A z + B) / (1 - z) at view depth z = -1
    assert np.allclose(depth, z, rtol=0, atol=4e-7)


def test_constant_depth_triangle_stays_within_rounding_of_its_depth():
    tri = mesh([at_pixel(5, 5), at_pixel(5, 60), at_pixel(60, 5)], [(0, 1, 2)])
    vis, depth, _, _ = O.render(S.Scene([tri]), exact_pixel_camera(), S.Opts(cull_mode=S.CULL_NONE), W, H,
                                want_rgba=False)
    fr_flags, rec = O.setup(S.Scene([tri]), exact_pixel_camera(), S.CULL_NONE, W, H)
    z0 = rec[0, 6:9].view(np.float32)
    assert z0[0] == z0[1] == z0[2]
    d = depth[vis != O.VIS_NONE]
    # three weighted terms: at most a few ulp from the plane value
    assert np.abs(d - z0[0]).max() <= 4 * np.spacing(np.float32(z0[0]))


def test_equal_depth_tie_goes_to_the_later_triangle():
    pts = [at_pixel(8, 8), at_pixel(8, 40), at_pixel(40, 8)]
    twice = mesh(pts + pts, [(0, 1, 2), (3, 4, 5)])
    vis, _, _, _ = O.render(S.Scene([twice]), exact_pixel_camera(), S.Opts(cull_mode=S.CULL_NONE), W, H,
                            want_rgba=False)
    assert set(np.unique(vis).tolist()) == {1, O.VIS_NONE}


def test_nearer_triangle_wins_regardless_of_order():
    near = [at_pixel(8, 8, -1.0), at_pixel(8, 40, -1.0), at_pixel(40, 8, -1.0)]
    far = [at_pixel(8, 8, -1.5), at_pixel(8, 40, -1.5), at_pixel(40, 8, -1.5)]
    for first, second, want in ((near, far, 0), (far, near, 1)):
        m = mesh(first + second, [(0, 1, 2), (3, 4, 5)])
        vis, _, _, _ = O.render(S.Scene([m]), exact_pixel_camera(), S.Opts(cull_mode=S.CULL_NONE), W, H,
                                want_rgba=False)
        covered = vis[vis != O.VIS_NONE]
        # the near triangle projects smaller (w = 2 vs 2.5), so every pixel it covers is its own
        assert (covered == want).sum() > 0
        nv, _, _, _ = O.render(S.Scene([mesh(near, [(0, 1, 2)])]), exact_pixel_camera(),
                               S.Opts(cull_mode=S.CULL_NONE), W, H, want_rgba=False)
        assert ((vis == want) == (nv != O.VIS_NONE)).all()


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("cull", [S.CULL_BACK, S.CULL_FRONT, S.CULL_NONE])
def test_sequential_raster_equals_per_pixel_brute_force(seed, cull):
    # brute force: per pixel, every triangle, GPU edge form; sequential: per
    # triangle, bbox pixels, CPU edge form (CCR/ERenderer.cpp:497-542)
    scene = random_triangle_scene(seed, 120)
    v1, d1, _, _ = O.render(scene, screen_camera(), S.Opts(cull_mode=cull), 96, 64, want_rgba=False)
    vb, db = O.render_bruteforce(scene, screen_camera(), cull, 96, 64)
    np.testing.assert_array_equal(v1, vb)
    np.testing.assert_array_equal(d1.view(np.uint32), db.view(np.uint32))


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_row_band_threads_equal_single_thread(threads):
    w = S.config1()
    a = O.render(w.scene, w.camera, w.opts, w.width, w.height, threads=1)
    b = O.render(w.scene, w.camera, w.opts, w.width, w.height, threads=threads)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))


def test_winner_depth_is_recomputed_depth():
    # invariant: stored depth == the winning triangle's interpolated z at the pixel
    scene = random_triangle_scene(3, 80)
    vis, depth, _, _ = O.render(scene, screen_camera(), S.Opts(cull_mode=S.CULL_NONE), 96, 64, want_rgba=False)
    flags, rec = O.setup(scene, screen_camera(), S.CULL_NONE, 96, 64)
    f = rec[:, :10].view(np.float32)
    ys, xs = np.nonzero(vis != O.VIS_NONE)
    for y, x in list(zip(ys, xs))[::7]:
        t = f[vis[y, x]]
        sx0, sy0, sx1, sy1, sx2, sy2, z0, z1, z2, area = [np.float32(v) for v in t]
        px, py = np.float32(x), np.float32(y)
        w2 = np.float32(np.float32((px - sx0) * (sy1 - sy0)) - np.float32((py - sy0) * (sx1 - sx0)))
        w0 = np.float32(np.float32((px - sx1) * (sy2 - sy1)) - np.float32((py - sy1) * (sx2 - sx1)))
        w1 = np.float32(np.float32((px - sx2) * (sy0 - sy2)) - np.float32((py - sy2) * (sx0 - sx2)))
        a0, a1, a2 = w0 / area, w1 / area, w2 / area
        z = np.float32(np.float32(np.float32(a0 * z0) + np.float32(a1 * z1)) + np.float32(a2 * z2))
        assert np.float32(z) == depth[y, x]


def test_bin_counts_match_counts_from_setup_records():
    w = S.config1()
    for tile, bt in ((8, 10_000), (16, 10_000), (32, 4)):
        counts, big = O.bin_counts(w.scene, w.camera, w.opts.cull_mode, w.width, w.height, tile, tile, bt)
        flags, rec = O.setup(w.scene, w.camera, w.opts.cull_mode, w.width, w.height)
        want = np.zeros_like(counts)
        want_big = []
        for g in np.nonzero(flags)[0]:
            xmin, ymin = int(rec[g, 10] & 0xFFFF), int(rec[g, 10] >> 16)
            xmax, ymax = int(rec[g, 11] & 0xFFFF), int(rec[g, 11] >> 16)
            if xmin >= xmax or ymin >= ymax:
                continue
            tx0, tx1, ty0, ty1 = xmin // tile, (xmax - 1) // tile, ymin // tile, (ymax - 1) // tile
            if (tx1 - tx0 + 1) * (ty1 - ty0 + 1) > bt:
                want_big.append(g)
                continue
            want[ty0:ty1 + 1, tx0:tx1 + 1] += 1
        np.testing.assert_array_equal(counts, want)
        assert big.tolist() == want_big


def test_backface_cull_on_screen_facing_sphere_keeps_front_half():
    w = S.config1()
    flags, rec = O.setup(w.scene, w.camera, S.CULL_BACK, w.width, w.height)
    nflags, _ = O.setup(w.scene, w.camera, S.CULL_NONE, w.width, w.height)
    ffl, _ = O.setup(w.scene, w.camera, S.CULL_FRONT, w.width, w.height)
    assert 0 < flags.sum() < nflags.sum()
    assert not (flags & ffl).any()  # a triangle is never both front and back facing
    # 64 zero-area pole triangles are culled even without face culling (D5)
    assert nflags.sum() <= 4032 - 64


def test_camera_matrices_sanity_bound_against_float64():
    cam = S.Camera(position=(1.5, 2.0, 7.0), forward=(0.3, 0.4, 0.86), fov_deg=60.0, near_z=0.5, far_z=50.0)
    fr = O.camera_frame(cam, 320, 200)
    f = np.array(cam.forward, np.float64)
    f /= np.linalg.norm(f)
    r = np.cross([0, 1, 0], f)
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    look = np.eye(4)
    look[:3, 0], look[:3, 1], look[:3, 2], look[:3, 3] = r, u, f, cam.position
    view = np.linalg.inv(look)
    t = math.tan(math.radians(cam.fov_deg) / 2)
    proj = np.eye(4)
    proj[0, 0] = 1 / ((320 / 200) * t)
    proj[1, 1] = 1 / t
    proj[2, 2] = cam.far_z / (cam.near_z - cam.far_z)
    proj[3, 2] = -1.0
    proj[2, 3] = cam.far_z * cam.near_z / (cam.near_z - cam.far_z)
    # data[col][row] storage -> transpose into row-major
    got_view = np.array(fr.view, np.float64).reshape(4, 4).T
    got_proj = np.array(fr.proj, np.float64).reshape(4, 4).T
    got_vp = np.array(fr.viewproj, np.float64).reshape(4, 4).T
    assert np.allclose(got_view, view, atol=2e-5)
    assert np.allclose(got_proj, proj, atol=1e-6)
    assert np.allclose(got_vp, proj @ view, atol=5e-5)


def test_mat4_inverse_and_product_bounds():
    rng = np.random.default_rng(0)
    for _ in range(20):
        m = rng.normal(size=(4, 4)).astype(np.float32) + 4 * np.eye(4, dtype=np.float32)
        data = np.ascontiguousarray(m.T).ravel()  # row-major -> data[col][row]
        inv = O.mat4_inverse(data).reshape(4, 4).T
        assert np.allclose(inv @ m, np.eye(4), atol=1e-4)
        prod = O.mat4_mul(data, data).reshape(4, 4).T
        assert np.allclose(prod, m.astype(np.float64) @ m, rtol=1e-5, atol=1e-4)


def test_point_sampling_known_texels_and_out_of_range_magenta():
    tex = (np.arange(16 * 8, dtype=np.uint32).reshape(8, 16) * 0x010203) | 0xFF000000
    # u * w + 0.5 truncated: u = 3/16 -> x = 3; v = 5/8 -> y = 5
    rgb = O.sample(tex, (3 / 16, 5 / 8), S.SAMPLE_POINT)
    t = int(tex[5, 3])
    assert list(rgb) == [np.float32(t & 255) / np.float32(255), np.float32((t >> 8) & 255) / np.float32(255),
                         np.float32((t >> 16) & 255) / np.float32(255)]
    for uv in ((1.0, 0.5), (0.5, 1.0), (-0.1, 0.5), (0.5, -0.2), (float("nan"), 0.5)):
        assert list(O.sample(tex, uv, S.SAMPLE_POINT)) == [1.0, 0.0, 1.0]
    # int() truncates toward zero: v * 8 + 0.5 = -0.06 still samples row 0
    assert list(O.sample(tex, (0.5, -0.07), S.SAMPLE_POINT)) == list(O.sample(tex, (0.5, 0.0), S.SAMPLE_POINT))
    # just inside: u * 16 + 0.5 = 15.9 -> x = 15; and -0.5/16 rounds to x = 0
    assert list(O.sample(tex, (15.4 / 16, 0.0), S.SAMPLE_POINT)) == list(O.sample(tex, (15 / 16, 0.0), 0))
    assert list(O.sample(tex, (-0.5 / 16, 0.0), S.SAMPLE_POINT)) == list(O.sample(tex, (0.0, 0.0), 0))


def test_linear_sampling_is_a_five_tap_average():
    const = np.full((8, 8), 0xFF3C7896, np.uint32)
    rgb = O.sample(const, (0.4, 0.6), S.SAMPLE_LINEAR)
    want = np.array([0x96, 0x78, 0x3C], np.float32) / np.float32(255)
    assert np.allclose(rgb, want, rtol=0, atol=2e-7)
    # a single bright texel: its 4 neighbours see 1/8 of it, itself 1/2
    spot = np.zeros((8, 8), np.uint32)
    spot[4, 4] = 0xFF0000FF  # red 255
    centre = O.sample(spot, (4 / 8, 4 / 8), S.SAMPLE_LINEAR)
    beside = O.sample(spot, (5 / 8, 4 / 8), S.SAMPLE_LINEAR)
    far = O.sample(spot, (6 / 8, 4 / 8), S.SAMPLE_LINEAR)
    assert abs(centre[0] - 0.5) < 1e-6 and abs(beside[0] - 0.125) < 1e-6 and far[0] == 0.0


def test_pack_rgba_saturates_like_the_gpu_conversion():
    assert O.pack_rgba([1.0, 0.5, 0.0]) == 0xFF007FFF
    assert O.pack_rgba([-0.5, float("nan"), 2.0]) == 0xFFFF0000
    assert O.pack_rgba([0.999, 0.0, 0.0]) & 0xFF == 254


def test_mlp_gradient_matches_finite_differences():
    in_dim, hidden = 8, 8
    rng = np.random.default_rng(1)
    n = O.mlp_num_params(in_dim, hidden)
    p = rng.normal(0, 0.5, n).astype(np.float32)
    x = rng.normal(0, 1, (32, in_dim)).astype(np.float32)
    y = rng.normal(1, 1, 32).astype(np.float32)
    tau = 0.8
    loss, g = O.mlp_loss_grad(p, in_dim, hidden, x, y, tau)
    h = 1e-2
    idx = rng.choice(n, 40, replace=False)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.mlp_loss_grad(pp, in_dim, hidden, x, y, tau)[0] - O.mlp_loss_grad(pm, in_dim, hidden, x, y, tau)[0]) / (
            pp[i] - pm[i])
        assert abs(fd - g[i]) <= 2e-3 + 2e-2 * abs(fd), (i, fd, g[i])
    # the forward output agrees with the loss's own forward
    pred = O.mlp_forward(p, in_dim, hidden, x)
    r = y.astype(np.float64) - pred
    want = np.mean(np.where(r > 0, tau * r, (tau - 1) * r))
    assert abs(want - loss) < 1e-5
