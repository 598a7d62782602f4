"""OBJ meshes in, BMP framebuffers out (librast_io): the C++ reader against
the oracle's Python restatement of CCR/ObjParser.cpp (bitwise), closed
forms (de-duplicated counts, axis-aligned tangents, inverted v), the
reference's input rules, and a BMP round trip."""
import os

import numpy as np
import pytest

import assets as A
import obj_oracle as OO

REF_RES = "/root/reference/CustomCUDARasterizer/Resources"


def cube_obj(crlf=True):
    """A unit cube in the reference's 3ds Max layout ("v" + two spaces),
    uvs and normals per face, comment / group / material lines, trailing
    spaces on faces."""
    L = ["# cube", "", "o Cube", "g Box001"]
    P = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
    L += [f"v  {x:.4f} {y:.4f} {z:.4f}" for x, y, z in P]
    L += ["vt 0.0000 0.0000 0.0000", "vt 1.0000 0.0000 0.0000", "vt 1.0000 1.0000 0.0000", "vt 0.0000 1.0000 0.0000"]
    N = [(0, 0, 1), (0, 0, -1), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)]
    L += [f"vn {x:.4f} {y:.4f} {z:.4f}" for x, y, z in N]
    L += ["usemtl mat", "s 1"]
    # faces counter-clockwise seen from outside; uv (0,0),(1,0),(1,1),(0,1) over each quad
    quads = [((5, 6, 7, 8), 1), ((2, 1, 4, 3), 2), ((6, 2, 3, 7), 3), ((1, 5, 8, 4), 4), ((8, 7, 3, 4), 5),
             ((1, 2, 6, 5), 6)]
    for (a, b, c, d), n in quads:
        L.append(f"f {a}/1/{n} {b}/2/{n} {c}/3/{n} ")
        L.append(f"f {a}/1/{n} {c}/3/{n} {d}/4/{n} ")
    return ("\r\n" if crlf else "\n").join(L).encode() + b"\r\n"


def check_same(text, invert_y=False):
    mesh, vtype = A.parse_obj(text, invert_y)
    ov, oi, ot = OO.parse(text, invert_y)
    assert vtype == ot
    np.testing.assert_array_equal(mesh.indices, oi)
    assert mesh.verts.shape == ov.shape
    np.testing.assert_array_equal(mesh.verts.view(np.uint32), ov.view(np.uint32))
    return mesh, vtype


def test_cube_deduplicates_to_24_vertices_with_axis_tangents():
    mesh, vtype = check_same(cube_obj())
    assert vtype == A.VERTEX_UV | A.VERTEX_NORMAL | A.VERTEX_TANGENT
    assert mesh.verts.shape[0] == 24 and mesh.indices.size == 36
    n, t = mesh.verts[:, 5:8], mesh.verts[:, 8:11]
    # u runs along one face edge: the tangent is an exact unit axis, orthogonal to the normal
    assert np.all(np.sum(np.abs(t), axis=1) == 1.0)
    assert np.all(np.sum(t * n, axis=1) == 0.0)
    # +z face: u along +x
    front = mesh.verts[:, 7] == 1.0
    np.testing.assert_array_equal(t[front], np.tile([1, 0, 0], (int(front.sum()), 1)))


def test_line_endings_do_not_matter():
    a, _ = A.parse_obj(cube_obj(crlf=True))
    b, _ = A.parse_obj(cube_obj(crlf=False))
    np.testing.assert_array_equal(a.verts, b.verts)
    np.testing.assert_array_equal(a.indices, b.indices)


def test_invert_y_flips_v_exactly():
    a, _ = check_same(cube_obj(), invert_y=True)
    b, _ = A.parse_obj(cube_obj(), invert_y=False)
    np.testing.assert_array_equal(a.verts[:, 4], np.float32(1) - b.verts[:, 4])


def test_positions_only_mesh():
    text = b"v  0 0 0\nv  1 0 0\nv  0 1 0\nv  1 1 0\nf 1 2 3\nf 3 2 4\nf 1 2 3\n"
    mesh, vtype = check_same(text)
    assert vtype == 0
    assert mesh.verts.shape[0] == 4
    np.testing.assert_array_equal(mesh.indices, [0, 1, 2, 2, 1, 3, 0, 1, 2])
    np.testing.assert_array_equal(mesh.verts[:, 11:14], np.ones((4, 3), np.float32))  # colour (1, 1, 1)
    np.testing.assert_array_equal(mesh.verts[:, 3:5], np.zeros((4, 2), np.float32))


def test_uv_only_mesh_splits_vertices_by_uv():
    text = b"v  0 0 0\nv  1 0 0\nv  0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nvt 0.5 0.5\nf 1/1 2/2 3/3\nf 1/4 2/2 3/3\n"
    mesh, vtype = check_same(text)
    assert vtype == A.VERTEX_UV
    assert mesh.verts.shape[0] == 4  # vertex 1 appears with two uvs
    np.testing.assert_array_equal(mesh.indices, [0, 1, 2, 3, 1, 2])


def test_only_the_first_three_vertices_of_a_face_are_used():
    text = b"v  0 0 0\nv  1 0 0\nv  1 1 0\nv  0 1 0\nf 1 2 3 4\n"
    mesh, _ = check_same(text)
    np.testing.assert_array_equal(mesh.indices, [0, 1, 2])


def test_normals_equal_within_two_ulp_share_a_vertex():
    n1 = np.float32(0.6)
    n2 = np.nextafter(n1, np.float32(1))  # one ulp apart: the same normal for AreEqual
    text = (b"v  0 0 0\nv  1 0 0\nv  0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\n"
            + f"vn 0 {float(n1):.9g} 0.8\nvn 0 {float(n2):.9g} 0.8\n".encode()
            + b"f 1/1/1 2/2/1 3/3/1\nf 1/1/2 2/2/2 3/3/2\n")
    mesh, _ = check_same(text)
    assert mesh.verts.shape[0] == 3


@pytest.mark.parametrize("bad", [b"v  0 0 0\nv  1 0 0\nv  0 1 0\nf 1//1 2//1 3//1\n",
                                 b"v  0 0 0\nv  1 0 0\nv  0 1 0\nf -1 -2 -3\n",
                                 b"v  0 0 0\nv  1 0 0\nf 1 2 3\n",
                                 b"v  0 0 0\nv  1 0 0\nv  0 1 0\nf 1 2\n",
                                 b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",
                                 b"# nothing\n"])
def test_input_the_reference_cannot_parse_is_rejected(bad):
    with pytest.raises(A.AssetError):
        A.parse_obj(bad)


def test_load_requires_the_obj_extension(tmp_path):
    p = tmp_path / "mesh.txt"
    p.write_bytes(cube_obj())
    with pytest.raises(A.AssetError):
        A.load_obj(str(p))
    q = tmp_path / "mesh.obj"
    q.write_bytes(cube_obj())
    mesh, _ = A.load_obj(str(q))
    assert mesh.verts.shape[0] == 24


@pytest.mark.parametrize("name", ["lowpoly_bunny.obj", "tuktuk.obj", "vehicle.obj"])
def test_reference_assets_parse_like_the_oracle(name):
    path = os.path.join(REF_RES, name)
    if not os.path.exists(path):
        pytest.skip("the reference's assets are not present here")
    with open(path, "rb") as f:
        text = f.read()
    mesh, vtype = check_same(text)
    faces = sum(1 for line in text.split(b"\n") if line.startswith(b"f "))
    assert mesh.indices.size == 3 * faces
    if name == "lowpoly_bunny.obj":
        assert vtype == 0
        positions = sum(1 for line in text.split(b"\n") if line.startswith(b"v "))
        assert mesh.verts.shape[0] <= positions


def test_bmp_round_trip_and_header(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 2 ** 32, (37, 53), dtype=np.uint64).astype(np.uint32)
    p = str(tmp_path / "frame.bmp")
    A.write_bmp(p, img)
    with open(p, "rb") as f:
        b = f.read()
    assert b[:2] == b"BM"
    assert int.from_bytes(b[2:6], "little") == len(b) == 54 + 37 * 53 * 4
    assert int.from_bytes(b[18:22], "little") == 53 and int.from_bytes(b[22:26], "little") == 37
    # bottom-up rows, bytes B, G, R, A
    last_row_first_px = img[36, 0]
    assert b[54:58] == bytes([(last_row_first_px >> 16) & 255, (last_row_first_px >> 8) & 255,
                              last_row_first_px & 255, last_row_first_px >> 24])
    np.testing.assert_array_equal(A.read_bmp(p), img)
