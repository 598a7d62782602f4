"""Small hand-made scenes for closed-form and brute-force checks."""
import numpy as np

import scenes as S


def vertex(p, uv=(0.0, 0.0), n=(0.0, 0.0, 1.0), t=(1.0, 0.0, 0.0), c=(1.0, 1.0, 1.0)):
    return np.array(list(p) + list(uv) + list(n) + list(t) + list(c), np.float32)


def mesh(points, tris, world=None, textures=None, topology=S.TOPO_LIST, uvs=None, colours=None):
    vs = []
    for i, p in enumerate(points):
        vs.append(vertex(p, uv=uvs[i] if uvs is not None else (0.5, 0.5),
                         c=colours[i] if colours is not None else (1.0, 1.0, 1.0)))
    m = S.Mesh(verts=np.stack(vs), indices=np.asarray(tris, np.uint32).ravel(), topology=topology)
    if world is not None:
        m.world = np.asarray(world, np.float32)
    m.textures = textures
    return m


def screen_camera():
    """Camera at the origin looking down -z (forward = +z)."""
    return S.Camera(position=(0.0, 0.0, 0.0), forward=(0.0, 0.0, 1.0), fov_deg=90.0, near_z=0.1, far_z=100.0)


def random_triangle_scene(seed, n_tris, spread=1.2, zmin=-6.0, zmax=-1.5):
    """Random triangles in front of screen_camera(), both windings."""
    rng = np.random.default_rng(seed)
    pts, tris = [], []
    for t in range(n_tris):
        c = rng.uniform(-spread, spread, 2)
        z = rng.uniform(zmin, zmax)
        size = rng.uniform(0.02, 0.6)
        for _ in range(3):
            d = rng.normal(0, size, 2)
            pts.append((float(c[0] + d[0]) * -z / 2, float(c[1] + d[1]) * -z / 2, float(z + rng.normal(0, 0.3))))
        tris.append((3 * t, 3 * t + 1, 3 * t + 2))
    return S.Scene([mesh(pts, tris)])
