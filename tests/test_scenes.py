"""Scene generators: sizes from the workload definitions, winding, determinism."""
import math

import numpy as np

import scenes as S


def test_uv_sphere_has_4032_triangles():
    m = S.uv_sphere(64, 32)
    assert m.num_triangles == 4032
    assert m.verts.shape == (33 * 65, 14)
    assert np.allclose(np.linalg.norm(m.verts[:, :3], axis=1), 1.0, atol=1e-6)


def test_icosphere_subdivision_5_counts_and_outward_winding():
    m = S.icosphere(5)
    assert m.num_triangles == 20_480 and m.verts.shape[0] == 10_242
    p = m.verts[:, :3].astype(np.float64)
    t = m.indices.reshape(-1, 3)
    n = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    assert (np.einsum("ij,ij->i", n, p[t].sum(axis=1)) > 0).all()


def test_terrain_counts_and_upward_winding():
    m = S.terrain(32)
    assert m.num_triangles == 2 * 32 * 32 and m.verts.shape[0] == 33 * 33
    p = m.verts[:, :3].astype(np.float64)
    t = m.indices.reshape(-1, 3)
    n = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    assert (n[:, 1] > 0).all()
    uv = m.verts[:, 3:5]
    assert (uv > 0).all() and (uv < 1).all()


def test_workload_sizes():
    w = S.config1()
    assert (w.width, w.height) == (640, 480) and w.scene.num_triangles == 4032
    assert S.icosphere(5).num_triangles * 256 == 5_242_880
    assert S.terrain(1024).num_triangles == 2_097_152


def test_deterministic_sine_cosine():
    for x in np.linspace(-20, 20, 2001):
        assert abs(S.lib().rsc_sin(float(x)) - math.sin(x)) < 1e-12
        assert abs(S.lib().rsc_cos(float(x)) - math.cos(x)) < 1e-12


def test_camera_paths_stay_outside_the_terrain_and_look_down():
    for seed in (S.SEED_BASE + 0x300, S.SEED_BASE + 0x301):
        for cam in S.camera_path(seed, 50):
            assert cam.position[2] > 40.0  # beyond the +z edge of the 80-unit terrain
            assert cam.forward[1] > 0.0     # forward points away from the view: view looks down


def test_backdrop_faces_the_camera():
    cam = S.terrain_camera()
    m = S.backdrop(cam, 16 / 9)
    p = m.verts[:, :3].astype(np.float64)
    t = m.indices.reshape(-1, 3)
    n = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    to_cam = np.array(cam.position) - p[t[:, 0]]
    assert (np.einsum("ij,ij->i", n, to_cam) > 0).all()
