"""GPU parity: librast.so (HIP, gfx950) against the CPU oracle on identical
seeded inputs.  Visibility and depth must be bit-exact; RGBA8 within 1/255
per channel (the only non-bitwise step is powf in the Phong term)."""
import numpy as np
import pytest

import oracle as O
import rast as R
import scenes as S
from scene_util import mesh, random_triangle_scene, screen_camera

pytestmark = pytest.mark.gpu

RGBA_TOL = 1  # per channel, in units of 1/255


def gpu_render(scene, cam, opts, W, H, tile=16, policy=R.BIN_EXACT, frames=1, big_threshold=256):
    with R.Renderer(W, H, tile=tile, big_threshold=big_threshold) as r:
        r.add_scene(scene)
        for _ in range(frames):
            r.render(cam, opts, policy)
        vis, depth, rgba = r.readback()
        st = r.stats()
    return vis, depth, rgba, st


def assert_bitwise(name, got, want):
    g = got.view(np.uint32)
    w = want.view(np.uint32)
    bad = np.argwhere(g != w)
    if bad.size:
        y, x = bad[0]
        raise AssertionError(f"{name}: {len(bad)} pixels differ; first at (x={x}, y={y}): "
                             f"got {g[y, x]:#010x} want {w[y, x]:#010x}")


def assert_rgba_close(got, want, tol=RGBA_TOL):
    g = got.view(np.uint8).astype(np.int16)
    w = want.view(np.uint8).astype(np.int16)
    d = np.abs(g - w)
    if d.max() > tol:
        y, xc = np.argwhere(d > tol)[0]
        raise AssertionError(f"rgba: {(d > tol).sum()} channels off by > {tol}; max {d.max()}; "
                             f"first at pixel ({xc // 4}, {y})")


def check_frame(scene, cam, opts, W, H, tile=16, policy=R.BIN_EXACT, threads=8):
    ovis, odepth, orgba, _ = O.render(scene, cam, opts, W, H, threads=threads)
    vis, depth, rgba, st = gpu_render(scene, cam, opts, W, H, tile=tile, policy=policy)
    assert_bitwise("vis", vis, ovis)
    assert_bitwise("depth", depth, odepth)
    assert_rgba_close(rgba, orgba)
    return st


def test_device_is_gfx950():
    info = R.device_info(0)
    print(info)
    assert info["arch"].startswith("gfx950")


def test_config1_uvsphere_parity():
    w = S.config1()
    st = check_frame(w.scene, w.camera, w.opts, w.width, w.height)
    assert st["num_triangles"] == 4032


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("cull", [S.CULL_BACK, S.CULL_NONE, S.CULL_FRONT])
def test_random_triangles_parity(seed, cull):
    scene = random_triangle_scene(seed, 300)
    opts = S.Opts(cull_mode=cull)
    check_frame(scene, screen_camera(), opts, 160, 96)


def test_vertex_outputs_bitwise():
    w = S.config1()
    with R.Renderer(w.width, w.height) as r:
        r.add_scene(w.scene)
        r.render(w.camera, w.opts)
        r.sync()
        pos = r.debug_read(R.DEBUG_VERTEX_POS, np.float32).reshape(-1, 4)
        attr = r.debug_read(R.DEBUG_VERTEX_ATTR, np.float32).reshape(-1, 12)
    opos, oattr = O.vertex_transform(w.scene, w.camera, w.width, w.height)
    assert_bitwise("pos4", pos, opos)
    # GPU record: every attribute times 1/w: u v 1/w n.x | n.y n.z t.x t.y | t.z vd.xyz (the colour
    # times 1/w is formed at shading: test_vertex_colour_path_without_diffuse_map);
    # oracle: uv n tan vd c 1/w (the product is one binary32 multiply on both sides)
    gpu_in_oracle_order = attr[:, [0, 1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 2]]
    inv_w = oattr[:, 14:15].astype(np.float32)
    want = np.concatenate([(oattr[:, :11].astype(np.float32) * inv_w).astype(np.float32), inv_w], axis=1)
    assert_bitwise("attr", np.ascontiguousarray(gpu_in_oracle_order), np.ascontiguousarray(want))


def test_setup_records_bitwise():
    w = S.config1()
    with R.Renderer(w.width, w.height) as r:
        r.add_scene(w.scene)
        r.render(w.camera, w.opts)
        r.sync()
        alive = r.debug_read(R.DEBUG_TRI_ALIVE, np.uint8)
        rec = r.debug_read(R.DEBUG_TRI_RECORDS, np.uint32).reshape(-1, 16)
    oflags, orec = O.setup(w.scene, w.camera, w.opts.cull_mode, w.width, w.height)
    np.testing.assert_array_equal(alive, oflags)
    live = oflags.astype(bool)
    # GPU words 0-11: sx0 sy0 sx1 sy1 sx2 sy2 z0 area z2 z1 lo hi; oracle: ... z0 z1 z2 area lo hi
    gpu_in_oracle_order = rec[live][:, [0, 1, 2, 3, 4, 5, 6, 9, 8, 7, 10, 11]]
    assert_bitwise("trirec", np.ascontiguousarray(gpu_in_oracle_order), orec[live])


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_tile_counts_exact(tile):
    w = S.config1()
    with R.Renderer(w.width, w.height, tile=tile, big_threshold=64) as r:
        r.add_scene(w.scene)
        r.render(w.camera, w.opts)
        r.sync()
        counts = r.debug_read(R.DEBUG_TILE_COUNTS, np.uint32)
        big = r.debug_read(R.DEBUG_BIG_LIST, np.uint32)
    ocounts, obig = O.bin_counts(w.scene, w.camera, w.opts.cull_mode, w.width, w.height, tile, tile, 64)
    np.testing.assert_array_equal(counts, ocounts.ravel())
    assert sorted(big[1:1 + big[0]].tolist()) == obig.tolist()


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_tile_size_invariance(tile):
    w = S.config1()
    check_frame(w.scene, w.camera, w.opts, w.width, w.height, tile=tile)


def test_repeated_frames_bit_identical():
    w = S.config1()
    a = gpu_render(w.scene, w.camera, w.opts, w.width, w.height)
    b = gpu_render(w.scene, w.camera, w.opts, w.width, w.height, frames=3)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("ntri", [50, 400])
@pytest.mark.parametrize("tile", [8, 16, 32])
def test_big_triangles_screen_list(ntri, tile):
    """Two screen-covering triangles plus small ones; big_threshold 1 puts
    all that span more than one tile on the screen-wide list: 12-22 with 50
    triangles, which fits one wave (every wave bounds it on its own), and
    107-186 with 400, longer than a wave: one staged chunk for the 16 and 32
    tiles (256 threads), sorted front to back first for the 8 tiles (64
    threads)."""
    scene = random_triangle_scene(7, ntri)
    quad = mesh([(-50, -50, -20), (50, -50, -20), (50, 50, -20), (-50, 50, -20)], [(0, 1, 2), (0, 2, 3)])
    scene = S.Scene(scene.meshes + [quad])
    ovis, odepth, _, _ = O.render(scene, screen_camera(), S.Opts(), 128, 96, threads=4, want_rgba=False)
    for bt in (1, 4, 256):
        vis, depth, _, st = gpu_render(scene, screen_camera(), S.Opts(), 128, 96, tile=tile, big_threshold=bt)
        if bt == 1:
            assert (st["num_big"] <= 64) == (ntri == 50), st["num_big"]
            assert st["num_big"] <= 256, st["num_big"]
        assert_bitwise(f"vis bt={bt}", vis, ovis)
        assert_bitwise(f"depth bt={bt}", depth, odepth)


@pytest.mark.parametrize("micro", [0, 4, 16, 64, 1024])
def test_micro_limit_invariance(micro):
    w = S.config1()
    ovis, odepth, _, _ = O.render(w.scene, w.camera, w.opts, w.width, w.height, threads=4, want_rgba=False)
    scene = random_triangle_scene(11, 400)
    rvis, rdepth, _, _ = O.render(scene, screen_camera(), S.Opts(), 160, 96, threads=4, want_rgba=False)
    for sc, cam, opts, W, H, ov, od in ((w.scene, w.camera, w.opts, w.width, w.height, ovis, odepth),
                                         (scene, screen_camera(), S.Opts(), 160, 96, rvis, rdepth)):
        with R.Renderer(W, H) as r:
            r.set_tuning(micro, 16)
            r.add_scene(sc)
            r.render(cam, opts)
            vis, depth, _ = r.readback()
        assert_bitwise(f"vis micro={micro}", vis, ov)
        assert_bitwise(f"depth micro={micro}", depth, od)


@pytest.mark.parametrize("small", [0, 16, 256])
def test_small_limit_invariance(small):
    w = S.config1()
    ovis, odepth, _, _ = O.render(w.scene, w.camera, w.opts, w.width, w.height, threads=4, want_rgba=False)
    with R.Renderer(w.width, w.height) as r:
        r.set_tuning(0, small)
        r.add_scene(w.scene)
        r.render(w.camera, w.opts)
        vis, depth, _ = r.readback()
    assert_bitwise("vis", vis, ovis)
    assert_bitwise("depth", depth, odepth)


@pytest.mark.parametrize("tile,micro,small", [(32, 0, 1024), (16, 0, 1024), (32, 1024, 1024), (8, 0, 64)])
def test_crowded_tiles_and_largest_limits(tile, micro, small):
    """600 overlapping random triangles at 160x96: tens of thousands of queued
    items per tile (several 4096-item owner windows and several 256-triangle
    staging chunks), items and micro-triangle ranges up to the 1024-pixel
    limits; exact and previous-frame binning against the oracle."""
    scene = random_triangle_scene(23, 600)
    W, H = 160, 96
    ovis, odepth, _, _ = O.render(scene, screen_camera(), S.Opts(), W, H, threads=4, want_rgba=False)
    for policy in (R.BIN_EXACT, R.BIN_PREVFRAME):
        with R.Renderer(W, H, tile=tile) as r:
            r.set_tuning(micro, small)
            r.add_scene(scene)
            for _ in range(2):  # the heuristic sizes its queues from the frame before
                r.render(screen_camera(), S.Opts(), policy)
            vis, depth, _ = r.readback()
        assert_bitwise(f"vis tile={tile} micro={micro} small={small} policy={policy}", vis, ovis)
        assert_bitwise(f"depth tile={tile} micro={micro} small={small} policy={policy}", depth, odepth)


def test_config2_icospheres_full_size():
    w = S.config2()
    st = check_frame(w.scene, w.camera, w.opts, w.width, w.height, threads=16)
    assert st["num_triangles"] == 5_242_880


def test_config3_terrain_4k_full_size():
    w = S.config3()
    st = check_frame(w.scene, w.camera, w.opts, w.width, w.height, threads=16)
    assert st["num_triangles"] == 2_097_152 + 8


def small_terrain_workload():
    w = S.config3(n=192, tex_size=256, width=640, height=360)
    return w


@pytest.mark.parametrize("policy", [R.BIN_PREVFRAME, R.BIN_LEARNED])
@pytest.mark.parametrize("margin,slack", [(1.0, 0), (1.25, 4), (4.0, 64)])
def test_single_pass_policies_match_exact(policy, margin, slack):
    """Single-pass binning with any capacities (including ones that overflow
    on most tiles) renders exactly the frames of the exact two-pass binner."""
    w = small_terrain_workload()
    _, path = S.train_test_paths(4, 6)
    with R.Renderer(w.width, w.height) as ex, R.Renderer(w.width, w.height) as sp:
        ex.add_scene(w.scene)
        sp.add_scene(w.scene)
        sp.set_tuning(0, 16)  # bin every triangle so the queues carry the whole frame
        model = R.BinModel(sp, hidden=16, seed=3)
        model.attach(sp, margin, slack)
        overflowed = 0
        for cam in path:
            ex.render(cam, w.opts, R.BIN_EXACT)
            sp.render(cam, w.opts, policy)
            a = ex.readback()
            b = sp.readback()
            overflowed += sp.stats()["overflow_tiles"]
            assert_bitwise("vis", b[0], a[0])
            assert_bitwise("depth", b[1], a[1])
            np.testing.assert_array_equal(b[2], a[2])
        model.close()
    if margin == 1.0 and slack == 0:
        assert overflowed > 0  # the fallback path really ran


@pytest.mark.parametrize("vis_mode,sample", [(S.VIS_PBR, S.SAMPLE_LINEAR), (S.VIS_DEPTH, S.SAMPLE_POINT),
                                             (S.VIS_NORMAL, S.SAMPLE_POINT)])
def test_sampler_and_visualisation_modes(vis_mode, sample):
    w = S.config1()
    opts = S.Opts(visualisation=vis_mode, sample_state=sample)
    check_frame(w.scene, w.camera, opts, w.width, w.height)


@pytest.mark.parametrize("sample", [S.SAMPLE_POINT, S.SAMPLE_LINEAR])
def test_texture_paths_for_map_sizes_and_mixed_meshes(sample):
    """Three ways the shading pass finds texels: four maps of one size (the
    indices are computed once per pixel), four maps of different sizes, and
    waves whose pixels belong to two meshes (per-lane texture tables).  The
    second sphere's texture coordinates run past 1 (magenta texels)."""
    a = S.uv_sphere(40, 20)
    a.textures = [S.noise_texture(96, 64, 11 + k, k) for k in range(4)]
    a.world = S.world_matrix((-0.45, 0.0, 0.0), (0.3, 0.5, 0.0), 0.8)
    b = S.uv_sphere(40, 20)
    b.verts = b.verts.copy()
    b.verts[:, 3:5] *= 1.2
    b.textures = [S.noise_texture(64, 48, 21, 0), S.noise_texture(32, 32, 22, 1), None,
                  S.noise_texture(16, 64, 23, 3)]
    b.world = S.world_matrix((0.45, 0.1, -0.2), (0.0, -0.4, 0.2), 0.8)
    check_frame(S.Scene([a, b]), S.Camera(), S.Opts(sample_state=sample), 320, 240)


def test_vertex_colour_path_without_diffuse_map():
    """Random vertex colours on two meshes without maps; the second sphere
    (in front of the first) has its vertices after the first's, so its
    colours are found through its own vertex base."""
    rng = np.random.default_rng(2)
    a = S.uv_sphere(48, 24)
    a.verts[:, 11:14] = rng.uniform(0, 1, (a.verts.shape[0], 3)).astype(np.float32)
    b = S.uv_sphere(24, 12)
    b.verts[:, 11:14] = rng.uniform(0, 1, (b.verts.shape[0], 3)).astype(np.float32)
    b.world = S.world_matrix((0.6, 0.4, 1.3), (0.2, 0.1, 0.0), 0.35)
    check_frame(S.Scene([a, b]), S.Camera(), S.Opts(), 320, 240)


def test_triangle_strip_topology():
    # a zig-zag strip across the screen: odd triangles swap their last two indices
    pts = []
    for i in range(40):
        x = -1.4 + 2.8 * (i // 2) / 19.0
        pts.append((x, 0.6 if i % 2 == 0 else -0.6, -2.0 - 0.01 * i))
    strip = mesh(pts, list(range(40)), topology=S.TOPO_STRIP)
    for cull in (S.CULL_BACK, S.CULL_NONE):
        check_frame(S.Scene([strip]), screen_camera(), S.Opts(cull_mode=cull), 200, 120)


def test_multi_mesh_ids_and_world_matrices():
    a = random_triangle_scene(21, 60).meshes[0]
    b = random_triangle_scene(22, 60).meshes[0]
    b.world = S.world_matrix((0.2, -0.1, -0.3), (0.1, 0.2, 0.05), 0.9)
    c = random_triangle_scene(23, 60).meshes[0]
    scene = S.Scene([a, b, c])
    check_frame(scene, screen_camera(), S.Opts(cull_mode=S.CULL_NONE), 160, 96)


def test_frame_sequence_matches_oracle_every_frame():
    """Eleven frames of one renderer with a moving camera, each equal to the
    oracle's: no depth key a frame leaves behind may show in the next one.
    The sequence mixes eager frames, graph replays, a batch and the
    previous-frame single-pass policy, and includes a frame that sees
    nothing, whose buffers must come out empty."""
    scene = random_triangle_scene(31, 400)
    W, H = 160, 96
    steps = []
    for k in range(11):
        if k == 6:
            cam = S.Camera(position=(0.0, 0.0, 0.0), forward=(0.0, 0.0, -1.0), fov_deg=90.0,
                           near_z=0.1, far_z=100.0)  # looking away: no fragment at all
        else:
            cam = S.Camera(position=(0.04 * k, -0.03 * k, 0.08 * k), forward=(0.0, 0.0, 1.0), fov_deg=90.0,
                           near_z=0.1, far_z=100.0)
        steps.append(cam)
    opts = S.Opts(cull_mode=S.CULL_NONE)
    with R.Renderer(W, H) as r:
        r.add_scene(scene)
        for k, cam in enumerate(steps):
            if k in (3, 4):
                r.render(cam, opts, R.BIN_EXACT, R.FLAG_GRAPH)
            elif k == 8:
                R.render_batch([(r, cam, opts, R.BIN_EXACT, 0)])
            elif k == 9:
                r.render(cam, opts, R.BIN_PREVFRAME)
            else:
                r.render(cam, opts)
            vis, depth, rgba = r.readback()
            ovis, odepth, orgba, _ = O.render(scene, cam, opts, W, H, threads=4)
            assert_bitwise(f"vis, frame {k}", vis, ovis)
            assert_bitwise(f"depth, frame {k}", depth, odepth)
            assert_rgba_close(rgba, orgba)


@pytest.mark.parametrize("sample", [S.SAMPLE_POINT, S.SAMPLE_LINEAR])
def test_interleaved_maps_follow_texture_changes(sample):
    """Four maps of one size are sampled from an interleaved copy built on
    first use.  The frames stay the oracle's when the mesh is bound to another
    set of maps, when one map of the set is destroyed (its stand-in takes
    over and the copy is dropped) and when the first set is bound again."""
    m = S.uv_sphere(40, 20)
    first = [S.noise_texture(96, 64, 31 + k, k) for k in range(4)]
    second = [S.noise_texture(96, 64, 41 + k, k) for k in range(4)]
    m.textures = first
    scene = S.Scene([m])
    cam, opts, W, H = S.Camera(), S.Opts(sample_state=sample), 320, 240
    lib = R.lib()

    def same_as_oracle(r, textures, what):
        m.textures = textures
        r.render(cam, opts)
        vis, depth, rgba = r.readback()
        ovis, odepth, orgba, _ = O.render(scene, cam, opts, W, H, threads=4)
        assert_bitwise(f"vis, {what}", vis, ovis)
        assert_bitwise(f"depth, {what}", depth, odepth)
        assert_rgba_close(rgba, orgba)

    with R.Renderer(W, H) as r:
        r.add_scene(scene)
        mh = r.meshes[0]
        same_as_oracle(r, first, "first set")
        hs = [r.texture(t) for t in second]
        R.check(lib.rast_mesh_set_textures(mh, *hs), "rast_mesh_set_textures")
        same_as_oracle(r, second, "second set")
        R.check(lib.rast_texture_destroy(hs[1]), "rast_texture_destroy")
        del r.textures[id(second[1])]
        same_as_oracle(r, [second[0], None, second[2], second[3]], "normal map destroyed")
        R.check(lib.rast_mesh_set_textures(mh, *[r.texture(t) for t in first]), "rast_mesh_set_textures")
        same_as_oracle(r, first, "first set again")


@pytest.mark.parametrize("sample", [S.SAMPLE_POINT, S.SAMPLE_LINEAR])
def test_phong_highlight_at_full_gloss(sample):
    """The steepest specular term: gloss 255 (exponent 25) and a camera on
    the mirror direction of the light over a flat floor, so the angle sweeps
    from 0 up to its maximum |L| = 0.99938 across the frame and the sum
    saturates (MaxToOne) at the centre.  Both sides call powf (the GPU as
    exp2(e log2 x)); RGBA8 stays within 1/255."""
    n = 24
    pts, uvs = [], []
    for j in range(n + 1):
        for i in range(n + 1):
            pts.append((-3.0 + 6.0 * i / n, 0.0, -3.0 + 6.0 * j / n))
            uvs.append((i / n, j / n))
    tris = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = j * (n + 1) + i, j * (n + 1) + i + 1, (j + 1) * (n + 1) + i, (j + 1) * (n + 1) + i + 1
            tris += [(a, c, b), (b, c, d)]
    floor = mesh(pts, tris, uvs=uvs)
    floor.verts[:, 5:8] = (0.0, 1.0, 0.0)  # normal +y
    floor.verts[:, 8:11] = (1.0, 0.0, 0.0)  # tangent +x
    const = lambda r, g, b: np.full((4, 4), S.rgba(r, g, b), np.uint32)
    floor.textures = [const(40, 40, 40), const(128, 128, 255), const(255, 255, 255), const(255, 255, 255)]
    d = np.array([1.0, 1.0, -1.0]) / np.sqrt(3.0)  # the light's mirror direction off +y
    cam = S.Camera(position=tuple(5.0 * d), forward=tuple(d), fov_deg=60.0)
    st = check_frame(S.Scene([floor]), cam, S.Opts(cull_mode=S.CULL_NONE, sample_state=sample), 320, 240)
    assert st["num_visible"] > 0
