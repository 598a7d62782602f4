"""The conservative depth plane used by the resolve pass to skip screen-list
triangles (cull_plane in rast_kernels.hip) must never exceed a depth the
exact fragment test computes.  Checked on the CPU with the oracle: single
large triangles at random orientations and depths, every covered pixel's
stored depth (the computed z of that triangle) against the bound, for every
tile of the frame."""
import numpy as np
import pytest

import oracle as O
import scenes as S
from scene_util import mesh, screen_camera

U = 2.0 ** -24


def cull_plane(sx, sy, z, X0, Y0, T):
    """The kernel's bound, restated: (c, dzdx, dzdy) in float32, or None."""
    f = np.float32
    sx = [f(v) for v in sx]; sy = [f(v) for v in sy]; z = [f(v) for v in z]
    zlo, zhi = min(z), max(z)
    if not (zlo > 0 and zhi < 2):
        return None
    ax, ay = f(sx[1] - sx[0]), f(sy[1] - sy[0])
    bx, by = f(sx[2] - sx[1]), f(sy[2] - sy[1])
    cx, cy = f(sx[0] - sx[2]), f(sy[0] - sy[2])
    emax2 = max(f(ax * ax + ay * ay), f(bx * bx + by * by), f(cx * cx + cy * cy))
    ex, ey, fx2, fy2 = f(sx[0] - sx[1]), f(sy[0] - sy[1]), f(sx[0] - sx[2]), f(sy[0] - sy[2])
    area = abs(f(f(ex * fy2) - f(ey * fx2)))
    if not (area >= f(0.1) * emax2 and emax2 >= 256 and emax2 < 1e12):
        return None
    e1x, e1y, e1z = float(sx[1]) - float(sx[0]), float(sy[1]) - float(sy[0]), float(z[1]) - float(z[0])
    e2x, e2y, e2z = float(sx[2]) - float(sx[0]), float(sy[2]) - float(sy[0]), float(z[2]) - float(z[0])
    nz = e1x * e2y - e1y * e2x
    if nz == 0:
        return None
    dzdx = -(e1y * e2z - e1z * e2y) / nz
    dzdy = -(e1z * e2x - e1x * e2z) / nz
    zt = float(z[0]) + dzdx * (X0 - float(sx[0])) + dzdy * (Y0 - float(sy[0]))
    margin = 512 * U * float(zhi) + 8 * U * (abs(zt) + (abs(dzdx) + abs(dzdy)) * T)
    return f(zt - margin), f(dzdx), f(dzdy)


def random_big_triangle(rng, W, H):
    # vertices around the screen centre at varied depths: large, well-conditioned, tilted in depth
    cx, cy = rng.uniform(-0.3, 0.3, 2)
    r = rng.uniform(0.6, 3.0)
    a0 = rng.uniform(0, 2 * np.pi)
    pts = []
    for k in range(3):
        a = a0 + k * 2 * np.pi / 3 + rng.uniform(-0.4, 0.4)
        d = rng.uniform(-9.0, -1.5)
        pts.append(((cx + r * np.cos(a)) * -d, (cy + r * np.sin(a)) * -d, d))
    return mesh(pts, [(0, 1, 2)])


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("T", [8, 16, 32])
def test_depth_plane_never_exceeds_a_computed_depth(seed, T):
    rng = np.random.default_rng(1000 + seed)
    W, H = 192, 128
    m = random_big_triangle(rng, W, H)
    scene = S.Scene([m])
    cam = screen_camera()
    opts = S.Opts(cull_mode=S.CULL_NONE)
    flags, rec = O.setup(scene, cam, opts.cull_mode, W, H)
    if not flags[0]:
        pytest.skip("triangle culled")
    vis, depth, _, _ = O.render(scene, cam, opts, W, H, threads=2, want_rgba=False)
    r = rec[0].view(np.float32)
    sx, sy, z = (r[0], r[2], r[4]), (r[1], r[3], r[5]), (r[6], r[7], r[8])
    covered = vis != 0xFFFFFFFF
    assert covered.sum() > 100
    checked = 0
    worst = np.inf
    for Y0 in range(0, H, T):
        for X0 in range(0, W, T):
            pl = cull_plane(sx, sy, z, X0, Y0, T)
            if pl is None:
                continue
            c, dx, dy = pl
            for py in range(Y0, min(Y0 + T, H)):
                xs = np.arange(X0, min(X0 + T, W))
                cov = covered[py, xs]
                if not cov.any():
                    continue
                zc = (c + dx * (xs - X0).astype(np.float32)).astype(np.float32)
                zc = (zc + dy * np.float32(py - Y0)).astype(np.float32)
                d = depth[py, xs]
                assert np.all(zc[cov] <= d[cov]), f"bound above a computed depth at row {py}"
                worst = min(worst, float(np.min(d[cov] - zc[cov])))
                checked += int(cov.sum())
    assert checked > 100
    # the bound is useful: within 1e-3 of the depths it bounds
    assert worst < 1e-3
