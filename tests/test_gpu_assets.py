"""GPU: a mesh read by the OBJ reader renders like the oracle (vis and depth
bit for bit, RGBA8 within 1/255); so do a mesh textured with the PNG decoder's
output (the reference's uv_grid.png) and the reference's vehicle.obj with its
four PNG maps (when tests/golden/assets holds them); a texture uploaded from
rows with a pitch renders exactly like the same texels uploaded packed; the
framebuffer writes to BMP and reads back unchanged."""
import ctypes as C
import os

import numpy as np
import pytest

import assets as A
import oracle as O
import rast as R
import scenes as S
from test_assets import cube_obj

pytestmark = pytest.mark.gpu


def cube_scene():
    mesh, _ = A.parse_obj(cube_obj())
    mesh.world = S.world_matrix((0.0, 0.0, 0.0), (0.5, 0.7, 0.1), 1.0)
    mesh.textures = [S.checker(64, 64, 8), None, None, None]
    return S.Scene([mesh])


def test_obj_mesh_renders_like_the_oracle():
    scene = cube_scene()
    cam = S.Camera(position=(0.0, 0.0, 5.0))
    W, H = 320, 240
    ovis, odepth, orgba, _ = O.render(scene, cam, S.Opts(), W, H, threads=4)
    with R.Renderer(W, H) as r:
        r.add_scene(scene)
        r.render(cam, S.Opts())
        vis, depth, rgba = r.readback()
    np.testing.assert_array_equal(vis, ovis)
    np.testing.assert_array_equal(depth.view(np.uint32), odepth.view(np.uint32))
    d = np.abs(rgba.view(np.uint8).astype(np.int16) - orgba.view(np.uint8).astype(np.int16))
    assert d.max() <= 1
    assert (vis != R.VIS_NONE).sum() > 5000


def test_pitched_texture_upload_matches_packed(tmp_path):
    scene = cube_scene()
    tex = S.noise_texture(100, 60, 7, 0)  # width not a multiple of any alignment
    pitch = 100 * 4 + 64
    padded = np.zeros((60, pitch // 4), np.uint32)
    padded[:, :100] = tex
    padded[:, 100:] = 0xDEADBEEF
    cam = S.Camera(position=(0.0, 0.0, 5.0))
    frames = []
    for pitched in (False, True):
        with R.Renderer(256, 192) as r:
            if pitched:
                h = C.c_void_p()
                R.check(R.lib().rast_texture_create_pitched(r.h, padded.ctypes.data, 100, 60, pitch, C.byref(h)),
                        "rast_texture_create_pitched")
            else:
                h = r.texture(tex)
            m = r.add_mesh(S.Mesh(verts=scene.meshes[0].verts, indices=scene.meshes[0].indices,
                                  world=scene.meshes[0].world))
            R.check(R.lib().rast_mesh_set_textures(m, h, None, None, None), "rast_mesh_set_textures")
            r.render(cam, S.Opts())
            frames.append(r.readback())
    for a, b in zip(*frames):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    p = str(tmp_path / "frame.bmp")
    A.write_bmp(p, frames[0][2])
    np.testing.assert_array_equal(A.read_bmp(p), frames[0][2])


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def assert_like_oracle(scene, cam, opts, W, H, min_covered):
    ovis, odepth, orgba, _ = O.render(scene, cam, opts, W, H, threads=8)
    with R.Renderer(W, H) as r:
        r.add_scene(scene)
        r.render(cam, opts)
        vis, depth, rgba = r.readback()
    np.testing.assert_array_equal(vis, ovis)
    np.testing.assert_array_equal(depth.view(np.uint32), odepth.view(np.uint32))
    d = np.abs(rgba.view(np.uint8).astype(np.int16) - orgba.view(np.uint8).astype(np.int16))
    assert d.max() <= 1
    assert (vis != R.VIS_NONE).sum() > min_covered
    return rgba


def test_png_texture_renders_like_the_oracle():
    scene = cube_scene()
    with open(os.path.join(GOLDEN, "uv_grid.png"), "rb") as f:
        grid = A.decode_png(f.read())
    scene.meshes[0].textures = [grid, None, None, None]
    for sample in (S.SAMPLE_POINT, S.SAMPLE_LINEAR):
        rgba = assert_like_oracle(scene, S.Camera(position=(0.0, 0.0, 5.0)), S.Opts(sample_state=sample), 320, 240,
                                  5000)
        red = rgba & 0xFF
        assert (red < 40).sum() > 1000 and (red > 215).sum() > 1000  # the grid's black and white cells


@pytest.mark.parametrize("cam", [S.Camera(position=(0.0, 5.0, 65.0)),
                                 S.Camera(position=(30.0, 15.0, 45.0), forward=(0.5345, 0.2673, 0.8018))])
def test_vehicle_with_its_png_maps_renders_like_the_oracle(cam):
    """The reference's vehicle scene (CCR/main.cpp:164-178: vehicle.obj read
    with the y axis inverted, diffuse, normal, specular and gloss maps) at
    the reference app's camera distance and from a raised three-quarter view."""
    d = os.path.join(GOLDEN, "assets")
    names = ["vehicle_diffuse.png", "vehicle_normal.png", "vehicle_specular.png", "vehicle_gloss.png"]
    if not all(os.path.exists(os.path.join(d, n)) for n in names + ["vehicle.obj"]):
        pytest.skip("reference vehicle not copied (tests/golden/copy_reference_assets.py)")
    mesh, vtype = A.load_obj(os.path.join(d, "vehicle.obj"), invert_y=True)
    assert vtype & A.VERTEX_UV and vtype & A.VERTEX_NORMAL
    mesh.textures = [A.load_png(os.path.join(d, n)) for n in names]
    assert_like_oracle(S.Scene([mesh]), cam, S.Opts(), 640, 480, 20000)
