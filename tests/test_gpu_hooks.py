"""GPU: the reference's test and timing hooks over the C ABI -- DrawTexture /
DrawTextureGlobal (CCR/CUDARenderer.cu:1565-1617,1969-2012) against the
oracle's sampler, the event timer of CUDABenchMarker
(CCR/CUDABenchMarker.cpp:24-35), rast_frame_buffers' wait for a checked
frame, and handle lifetimes when a renderer is destroyed before the objects
holding its mesh and texture handles (the order of CCR/main.cpp:418,529)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import rast as R
import scenes as S
from scene_util import random_triangle_scene, screen_camera

pytestmark = pytest.mark.gpu


def texture():
    return S.noise_texture(48, 40, 77, 0)


@pytest.mark.parametrize("stretch", [True, False])
@pytest.mark.parametrize("sample", [S.SAMPLE_POINT, S.SAMPLE_LINEAR])
def test_draw_texture_global_matches_the_oracle_sampler(stretch, sample):
    """uv = (x, y) / window size (stretched) or / texture size (texels past
    the texture's edge sample magenta), sampled and packed as shading does."""
    W, H = 96, 64
    tex = texture()
    th, tw = tex.shape
    with R.Renderer(W, H) as r:
        r.draw_texture(tex, R.DRAW_TEXTURE_GLOBAL, stretch, sample)
        vis, depth, rgba = r.readback()
    assert np.all(vis == R.VIS_NONE) and np.all(depth == 1.0)
    dw, dh = (W, H) if stretch else (tw, th)
    want = np.zeros((H, W), np.uint32)
    for y in range(H):
        for x in range(W):
            uv = (np.float32(x) / np.float32(dw), np.float32(y) / np.float32(dh))
            want[y, x] = O.pack_rgba(O.sample(tex, uv, sample))
    bad = np.argwhere(rgba != want)
    assert bad.size == 0, f"{len(bad)} pixels differ, first {bad[0]}"
    if not stretch:
        assert np.any(rgba == O.pack_rgba((1.0, 0.0, 1.0)))  # magenta past the edge


def test_draw_texture_point_fetch_with_wrap():
    """DrawTexture: the texel nearest to (x / W, y / H) in normalized
    coordinates, (floor(u w) mod w, floor(v h) mod h), copied unchanged."""
    W, H = 100, 70
    tex = texture()
    th, tw = tex.shape
    with R.Renderer(W, H) as r:
        r.draw_texture(tex, R.DRAW_TEXTURE)
        _, _, rgba = r.readback()
    xs = np.arange(W, dtype=np.float32) / np.float32(W)
    ys = np.arange(H, dtype=np.float32) / np.float32(H)
    ix = np.floor(xs * np.float32(tw)).astype(np.int64) % tw
    iy = np.floor(ys * np.float32(th)).astype(np.int64) % th
    np.testing.assert_array_equal(rgba, tex[iy[:, None], ix[None, :]])


def test_event_timer_brackets_frames_and_batches():
    scene = random_triangle_scene(5, 200)
    with R.Renderer(160, 96) as r:
        r.add_scene(scene)
        with pytest.raises(R.RastError):
            r.stop_timer()  # not started
        r.start_timer()
        for _ in range(3):
            r.render(screen_camera(), S.Opts())
        ms = r.stop_timer()
        assert 0.0 < ms < 10_000.0
        r.start_timer()
        R.render_batch([(r, screen_camera(), S.Opts(), R.BIN_EXACT, 0)])
        ms_batch = r.stop_timer()
        assert 0.0 < ms_batch < 10_000.0


def test_frame_buffers_wait_for_a_checked_frame():
    lib = R.lib()
    vis, depth, rgba = C.c_void_p(), C.c_void_p(), C.c_void_p()
    with R.Renderer(64, 48) as r:
        r.add_scene(random_triangle_scene(6, 20))
        assert lib.rast_frame_buffers(r.h, C.byref(vis), C.byref(depth), C.byref(rgba)) != R.RAST_OK  # no frame yet
        r.render(screen_camera(), S.Opts())
        R.check(lib.rast_frame_buffers(r.h, C.byref(vis), C.byref(depth), C.byref(rgba)), "rast_frame_buffers")
        assert vis.value and depth.value and rgba.value


def test_handles_outlive_their_renderer():
    """Mesh and texture handles of a destroyed renderer stay valid as empty
    handles: destroying them afterwards succeeds, binding them fails."""
    lib = R.lib()
    r = R.Renderer(64, 48)
    r.add_scene(random_triangle_scene(8, 10))
    mh = r.meshes[0]
    th = r.texture(texture())
    r.render(screen_camera(), S.Opts())
    r.sync()
    r.close()
    assert lib.rast_mesh_set_textures(mh, th, None, None, None) != R.RAST_OK
    R.check(lib.rast_texture_destroy(th), "rast_texture_destroy after the renderer")
    R.check(lib.rast_mesh_destroy(mh), "rast_mesh_destroy after the renderer")
