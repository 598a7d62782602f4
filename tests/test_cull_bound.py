"""The conservative depth bounds the resolve pass uses to skip screen-list
triangles (cull_plane in rast_kernels.hip) must hold for every depth the
exact fragment test computes.  Checked on the CPU with the oracle: single
triangles (large, thin, and straddling the camera plane, which the
reference's projection leaves unclipped), every covered pixel's stored depth
(that triangle's computed z) against lo(p) <= z <= lo(p) + w, for every tile;
a tile the bound skips (outside an edge, or out of depth range) must hold
no covered pixel, and the skips must occur."""
import numpy as np
import pytest

import oracle as O
import scenes as S
from scene_util import mesh, screen_camera

U = 2.0 ** -24


def cull_plane(rec, X0, Y0, T):
    """The kernel's bound, restated in float64 / float32: (x, y, z, w) or None."""
    f = np.float32
    sx = [float(rec[0]), float(rec[2]), float(rec[4])]
    sy = [float(rec[1]), float(rec[3]), float(rec[5])]
    sz = [float(rec[6]), float(rec[7]), float(rec[8])]
    area = float(rec[9])
    ax, ay = float(f(rec[2] - rec[0])), float(f(rec[3] - rec[1]))
    bx, by = float(f(rec[4] - rec[2])), float(f(rec[5] - rec[3]))
    cx, cy = float(f(rec[0] - rec[4])), float(f(rec[1] - rec[5]))
    R = max(max(abs(sx[i] - X0), abs(sx[i] - (X0 + T)), abs(sy[i] - Y0), abs(sy[i] - (Y0 + T))) for i in range(3))
    zmax = max(abs(z) for z in sz)
    E = max(abs(ax), abs(ay), abs(bx), abs(by), abs(cx), abs(cy))
    if not R < 1e9:
        return None
    lim = -7 * U * R * E
    for a, b in ((0, 1), (1, 2), (2, 0)):
        ex, ey = sx[b] - sx[a], sy[b] - sy[a]
        if all((cxx - sx[a]) * ey - (cyy - sy[a]) * ex < lim * (1 + 1e-12) - 1e-300
               for cxx in (X0, X0 + T - 1) for cyy in (Y0, Y0 + T - 1)):
            return "skip"
    if not (area > 0 and zmax < 1e6):
        return None
    k = 18.06 * R * E / area + 6.08 * E * E / area + 6.06
    if not (k * U < 1e-3):
        return None
    e1x, e1y, e1z = sx[1] - sx[0], sy[1] - sy[0], sz[1] - sz[0]
    e2x, e2y, e2z = sx[2] - sx[0], sy[2] - sy[0], sz[2] - sz[0]
    nz = e1x * e2y - e1y * e2x
    if nz == 0:
        return None
    dzdx = -(e1y * e2z - e1z * e2y) / nz
    dzdy = -(e1z * e2x - e1x * e2z) / nz
    zt = sz[0] + dzdx * (X0 - sx[0]) + dzdy * (Y0 - sy[0])
    span = (abs(dzdx) + abs(dzdy)) * T
    M = 2 * U * zmax * k + 8 * U * (abs(zt) + span) + 1e-12 * (abs(zt) + (abs(dzdx) + abs(dzdy)) * R)
    if not (M < 0.25):
        return None
    if zt - span - M > 1.0 or zt + span + M < 0.0:
        return "skip"
    return f(zt - M), f(dzdx), f(dzdy), f(2 * M + 8 * U * (abs(zt) + span))


def triangle(rng, kind):
    if kind == "big":
        cx, cy = rng.uniform(-0.3, 0.3, 2)
        r = rng.uniform(0.6, 3.0)
        a0 = rng.uniform(0, 2 * np.pi)
        pts = []
        for k in range(3):
            a = a0 + k * 2 * np.pi / 3 + rng.uniform(-0.4, 0.4)
            d = rng.uniform(-9.0, -1.5)
            pts.append(((cx + r * np.cos(a)) * -d, (cy + r * np.sin(a)) * -d, d))
    elif kind == "thin":
        p0 = np.array([rng.uniform(-2, 2), rng.uniform(-1.5, 1.5), rng.uniform(-8, -2)])
        d = rng.normal(0, 1, 3)
        d[2] *= 0.3
        pts = [tuple(p0), tuple(p0 + 3 * d), tuple(p0 + 3 * d + rng.normal(0, 0.12, 3))]
    else:  # straddling the camera plane: one vertex behind the camera
        pts = [(rng.uniform(-3, 3), rng.uniform(-2, 2), rng.uniform(-6, -2)),
               (rng.uniform(-3, 3), rng.uniform(-2, 2), rng.uniform(-6, -2)),
               (rng.uniform(-3, 3), rng.uniform(-2, 2), rng.uniform(0.3, 3.0))]
    return mesh(pts, [(0, 1, 2)])


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("kind", ["big", "thin", "straddle"])
@pytest.mark.parametrize("T", [8, 32])
def test_depth_bounds_hold_for_every_computed_depth(seed, kind, T):
    rng = np.random.default_rng(2000 + 17 * seed + len(kind))
    W, H = 160, 112
    scene = S.Scene([triangle(rng, kind)])
    cam, opts = screen_camera(), S.Opts(cull_mode=S.CULL_NONE)
    flags, rec = O.setup(scene, cam, opts.cull_mode, W, H)
    if not flags[0]:
        pytest.skip("triangle culled")
    vis, depth, _, _ = O.render(scene, cam, opts, W, H, threads=2, want_rgba=False)
    r = rec[0].view(np.float32)
    covered = vis != 0xFFFFFFFF
    checked = skipped = 0
    for Y0 in range(0, H, T):
        for X0 in range(0, W, T):
            pl = cull_plane(r, X0, Y0, T)
            if pl is None:
                continue
            ys, xs = slice(Y0, min(Y0 + T, H)), slice(X0, min(X0 + T, W))
            if isinstance(pl, str):
                assert not covered[ys, xs].any(), f"tile ({X0},{Y0}) skipped but covered"
                skipped += 1
                continue
            c, dx, dy, w = pl
            for py in range(Y0, min(Y0 + T, H)):
                px = np.arange(X0, min(X0 + T, W))
                cov = covered[py, px]
                if not cov.any():
                    continue
                lo = (c + dx * (px - X0).astype(np.float32)).astype(np.float32)
                lo = (lo + dy * np.float32(py - Y0)).astype(np.float32)
                hi = (lo + w).astype(np.float32)
                d = depth[py, px]
                assert np.all(lo[cov] <= d[cov]), f"lower bound above a computed depth, row {py}"
                assert np.all(hi[cov] >= d[cov]), f"upper bound below a computed depth, row {py}"
                checked += int(cov.sum())
    assert checked + skipped > 0 or not covered.any()


def test_tiles_away_from_a_triangle_are_skipped():
    # a triangle in the middle of the screen: tiles a few tiles away from it are outside an edge
    scene = S.Scene([mesh([(-0.3, -0.2, -3.0), (0.3, -0.2, -3.0), (0.0, 0.3, -3.0)], [(0, 1, 2)])])
    W, H, T = 256, 192, 16
    flags, rec = O.setup(scene, screen_camera(), S.CULL_NONE, W, H)
    r = rec[0].view(np.float32)
    marks = [[cull_plane(r, X0, Y0, T) for X0 in range(0, W, T)] for Y0 in range(0, H, T)]
    assert marks[0][0] == "skip" and marks[-1][-1] == "skip"
    centre = marks[H // T // 2][W // T // 2]
    assert centre is not None and centre != "skip"
