"""The conservative bounds the resolve pass uses before the exact fragment
test (tile_bound and fragment_closer in rast_kernels.hip), restated in
float32 / float64, must hold for every depth the oracle's fragment test
computes.  Single triangles (large, thin, and straddling the camera plane,
which the reference's projection leaves unclipped): every covered pixel's
stored depth (that triangle's computed z) against lo(p) <= z <= lo(p) + w
for every tile; a tile skipped as outside an edge or out of depth range
holds no covered pixel; with the tile's largest depth at the nearest
fragment the tile is never skipped; the division-free reject never drops a
covered pixel at its own depth or as out of range, and does reject
fragments behind the current depth."""
import numpy as np
import pytest

import oracle as O
import scenes as S
from scene_util import mesh, screen_camera

U = 2.0 ** -24
f = np.float32


def _tri(rec):
    sx0, sy0, sx1, sy1, sx2, sy2, z0, z1, z2, A = [f(v) for v in rec[:10]]
    e = (f(sx1 - sx0), f(sy1 - sy0), f(sx2 - sx1), f(sy2 - sy1), f(sx0 - sx2), f(sy0 - sy2))
    return (sx0, sy0, sx1, sy1, sx2, sy2), e, (z0, z1, z2), A


def edges(rec, px, py):
    """edges() in rast_kernels.hip: the three computed edge functions (w0, w1, w2)."""
    (sx0, sy0, sx1, sy1, sx2, sy2), (ax, ay, bx, by, cx, cy), _, _ = _tri(rec)
    px, py = f(px), f(py)
    w2 = f(f(f(px - sx0) * ay) - f(f(py - sy0) * ax))
    w0 = f(f(f(px - sx1) * by) - f(f(py - sy1) * bx))
    w1 = f(f(f(px - sx2) * cy) - f(f(py - sy2) * cx))
    return w0, w1, w2


def depth_sum(a, zs):
    return f(f(f(a[0] * zs[0]) + f(a[1] * zs[1])) + f(a[2] * zs[2]))


def tile_bound(rec, X0, Y0, T, tmax=1.0):
    """The kernel's tile_bound, restated in float32 / float64: (x, y, z, w) or "skip"."""
    s, e, zs, A = _tri(rec)
    none = (f(-1e30), f(0), f(0), f(1e30))
    if not max(abs(float(v)) for v in s + e) < 1e18:
        return none
    ws = [edges(rec, X0 + (c & 1) * (T - 1), Y0 + (c >> 1) * (T - 1)) for c in range(4)]
    lo = [min(w[i] for w in ws) for i in range(3)]
    hi = [max(w[i] for w in ws) for i in range(3)]
    if min(hi) < 0:
        return "skip"
    Ad = float(A)
    zmax = max(abs(float(z)) for z in zs)
    if Ad > 0 and zmax < 1e6:
        ox, oy = [float(s[0]), float(s[2]), float(s[4])], [float(s[1]), float(s[3]), float(s[5])]
        ex, ey = [float(e[0]), float(e[2]), float(e[4])], [float(e[1]), float(e[3]), float(e[5])]
        zz = [float(zs[2]), float(zs[0]), float(zs[1])]
        xa, xb, ya, yb = float(X0), float(X0 + T - 1), float(Y0), float(Y0 + T - 1)
        L0 = Lx = Ly = B = Lmag = 0.0
        for j in range(3):
            dxa, dxb, dya, dyb = xa - ox[j], xb - ox[j], ya - oy[j], yb - oy[j]
            Dx, Dy = max(abs(dxa), abs(dxb)), max(abs(dya), abs(dyb))
            wc = [dx * ey[j] - dy * ex[j] for dx in (dxa, dxb) for dy in (dya, dyb)]
            Wabs = max(abs(w) for w in wc)
            eps = 3.0001 * U * (Dx * abs(ey[j]) + Dy * abs(ex[j]))
            B += abs(zz[j]) * (eps + 4.0001 * U * (Wabs + eps))
            Lmag += abs(zz[j]) * Wabs
            L0 += zz[j] * wc[0]
            Lx += zz[j] * ey[j]
            Ly -= zz[j] * ex[j]
        rA = 1.0 / Ad
        B, Lmag, L0, Lx, Ly = B * rA, Lmag * rA, L0 * rA, Lx * rA, Ly * rA
        span = (abs(Lx) + abs(Ly)) * (T - 1)
        M = 1.01 * B + 8 * U * (abs(L0) + span + 1.0) + 1e-12 * Lmag + 1e-30
        if M < 0.25:
            Lmin = L0 + min(Lx, 0.0) * (T - 1) + min(Ly, 0.0) * (T - 1)
            Lmax = L0 + max(Lx, 0.0) * (T - 1) + max(Ly, 0.0) * (T - 1)
            if Lmin - M > 1.0 or Lmax + M < 0.0 or Lmin - M > tmax:
                return "skip"
            return f(L0 - M), f(Lx), f(Ly), f(2 * M + 8 * U * (abs(L0) + span + 1.0))
    with np.errstate(all="ignore"):
        lo_a = [f(max(lo[i], f(0)) / A) for i in range(3)]
        hi_a = [f(hi[i] / A) for i in range(3)]
        zlo = depth_sum([lo_a[i] if zs[i] >= 0 else hi_a[i] for i in range(3)], zs)
        zhi = depth_sum([hi_a[i] if zs[i] >= 0 else lo_a[i] for i in range(3)], zs)
    if zlo > 1 or zhi < 0 or zlo > tmax:
        return "skip"
    return (max(zlo, f(-1e30)), f(0), f(0), f(1e30)) if zlo == zlo else none


def closer_rejects(rec, px, py, dbest):
    """The division-free reject of fragment_closer: True when the fragment at
    (px, py) is claimed to lose against depth dbest or to be out of range."""
    _, _, zs, A = _tri(rec)
    w = edges(rec, px, py)
    if not all(x >= 0 for x in w):
        return False
    with np.errstate(all="ignore"):
        r = f(f(1) / A)
        p = [f(f(w[i] * r) * zs[i]) for i in range(3)]
        zq = f(f(p[0] + p[1]) + p[2])
        m = f(f(f(f(abs(p[0]) + abs(p[1])) + abs(p[2])) * f(10.0 / 16777216.0)) + f(1e-37))
        return bool(f(zq - m) > dbest or f(zq + m) < 0)


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
    elif kind == "straddle":  # one vertex behind the camera
        pts = [(rng.uniform(-3, 3), rng.uniform(-2, 2), rng.uniform(-6, -2)),
               (rng.uniform(-3, 3), rng.uniform(-2, 2), rng.uniform(-6, -2)),
               (rng.uniform(-3, 3), rng.uniform(-2, 2), rng.uniform(0.3, 3.0))]
    else:  # "graze": a ground triangle across the plane w = 0 (z = 1 for this camera), all
        # three vertices near it: screen coordinates of 1e3..1e4, depths of -200..150
        pts = [(rng.uniform(-1, 1), -0.2, 1.0 - rng.uniform(0.005, 0.05)),
               (rng.uniform(-1, 1), -0.2, 1.0 - rng.uniform(0.005, 0.05)),
               (rng.uniform(-1, 1), -0.2, 1.0 + rng.uniform(0.005, 0.05))]
    return mesh(pts, [(0, 1, 2)])


def drawn(rng, kind, cam, opts, W, H):
    """A triangle of the kind with its setup record and oracle frame; grazing
    triangles are redrawn until one covers pixels (about one in four does)."""
    for _ in range(40):
        scene = S.Scene([triangle(rng, kind)])
        flags, rec = O.setup(scene, cam, opts.cull_mode, W, H)
        vis, depth, _, _ = O.render(scene, cam, opts, W, H, threads=2, want_rgba=False)
        if kind != "graze" or (vis != 0xFFFFFFFF).any():
            break
    return scene, flags, rec, vis, depth


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("kind", ["big", "thin", "straddle", "graze"])
@pytest.mark.parametrize("T", [8, 32])
def test_depth_bounds_hold_for_every_computed_depth(seed, kind, T):
    rng = np.random.default_rng(2000 + 17 * seed + len(kind))
    W, H = 160, 112
    cam, opts = screen_camera(), S.Opts(cull_mode=S.CULL_NONE)
    scene, flags, rec, vis, depth = drawn(rng, kind, cam, opts, W, H)
    if not flags[0]:
        pytest.skip("triangle culled")
    r = rec[0].view(np.float32)
    covered = vis != 0xFFFFFFFF
    checked = skipped = occluded = planes = 0
    for Y0 in range(0, H, T):
        for X0 in range(0, W, T):
            pl = tile_bound(r, X0, Y0, T)
            ys, xs = slice(Y0, min(Y0 + T, H)), slice(X0, min(X0 + T, W))
            if isinstance(pl, str):
                assert not covered[ys, xs].any(), f"tile ({X0},{Y0}) skipped but covered"
                skipped += 1
                continue
            if covered[ys, xs].any():
                # behind every pixel: with tmax just below the tile's nearest
                # fragment the triangle may be skipped, with tmax at it never
                dmin = float(depth[ys, xs][covered[ys, xs]].min())
                assert tile_bound(r, X0, Y0, T, tmax=dmin) != "skip", f"tile ({X0},{Y0}) skipped at its own depth"
                below = np.nextafter(f(dmin), f(-1))
                pl2 = tile_bound(r, X0, Y0, T, tmax=float(below))
                occluded += pl2 == "skip"
            c, dx, dy, w = pl
            planes += bool(w < 1e29)
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
    if kind == "graze":
        assert covered.any() and planes > 0  # the per-pixel plane applies to them


def test_tiles_away_from_a_triangle_are_skipped():
    # a triangle in the middle of the screen: tiles a few tiles away from it are outside an edge
    scene = S.Scene([mesh([(-0.3, -0.2, -3.0), (0.3, -0.2, -3.0), (0.0, 0.3, -3.0)], [(0, 1, 2)])])
    W, H, T = 256, 192, 16
    flags, rec = O.setup(scene, screen_camera(), S.CULL_NONE, W, H)
    r = rec[0].view(np.float32)
    marks = [[tile_bound(r, X0, Y0, T) for X0 in range(0, W, T)] for Y0 in range(0, H, T)]
    assert marks[0][0] == "skip" and marks[-1][-1] == "skip"
    centre = marks[H // T // 2][W // T // 2]
    assert centre != "skip"
    # an occluder at depth 0.1 everywhere: every tile is skipped
    assert all(tile_bound(r, X0, Y0, T, tmax=0.1) == "skip" for X0 in range(0, W, T) for Y0 in range(0, H, T))


def closer_rejects_grid(rec, px, py, dbest):
    """closer_rejects over arrays of pixels (elementwise float32, same rounding)."""
    (sx0, sy0, sx1, sy1, sx2, sy2), (ax, ay, bx, by, cx, cy), zs, A = _tri(rec)
    px, py = px.astype(f), py.astype(f)
    with np.errstate(all="ignore"):
        w2 = (px - sx0) * ay - (py - sy0) * ax
        w0 = (px - sx1) * by - (py - sy1) * bx
        w1 = (px - sx2) * cy - (py - sy2) * cx
        inside = (w0 >= 0) & (w1 >= 0) & (w2 >= 0)
        r = f(f(1) / A)
        p0, p1, p2 = (w0 * r) * zs[0], (w1 * r) * zs[1], (w2 * r) * zs[2]
        zq = (p0 + p1) + p2
        m = ((np.abs(p0) + np.abs(p1)) + np.abs(p2)) * f(10.0 / 16777216.0) + f(1e-37)
        return inside & (((zq - m) > dbest) | ((zq + m) < 0)), inside


@pytest.mark.parametrize("kind", ["big", "thin", "straddle", "graze"])
def test_division_free_reject_keeps_every_fragment_that_can_win(kind):
    rng = np.random.default_rng(7000 + len(kind))
    W, H = 160, 112
    rejected_range = rejected_depth = 0
    for _ in range(6):
        cam, opts = screen_camera(), S.Opts(cull_mode=S.CULL_NONE)
        scene, flags, rec, vis, depth = drawn(rng, kind, cam, opts, W, H)
        if not flags[0]:
            continue
        r = rec[0].view(np.float32)
        covered = vis != 0xFFFFFFFF
        py, px = np.mgrid[0:H, 0:W]
        # out of range: never a pixel the oracle covers
        rej, inside = closer_rejects_grid(r, px, py, f(1))
        assert not (rej & covered).any()
        rejected_range += int((rej & inside).sum())
        # at its own depth (ties go to the ID) never rejected; well in front it is
        d = np.where(covered, depth, f(1)).astype(f)
        rej_at, _ = closer_rejects_grid(r, px, py, d)
        assert not (rej_at & covered).any()
        rej_front, _ = closer_rejects_grid(r, px, py, (d - f(1e-3)).astype(f))
        rejected_depth += int((rej_front & covered).sum())
    assert rejected_depth > 0
    if kind in ("straddle", "graze"):
        assert rejected_range > 0


def test_skip_marker_is_above_every_plane():
    """The kernel reads x > 1e37 as "skip the tile".  Planes of grazing
    triangles start well above 1 at the tile corner and fall steeply across
    it, with covered pixels in the tile: a marker test as low as 2 drops them."""
    W, H = 160, 112
    cam, opts = screen_camera(), S.Opts(cull_mode=S.CULL_NONE)
    high_with_fragments = 0
    for seed in range(5):
        rng = np.random.default_rng(2000 + 17 * seed + 5)
        scene, flags, rec, vis, depth = drawn(rng, "graze", cam, opts, W, H)
        r = rec[0].view(np.float32)
        covered = vis != 0xFFFFFFFF
        for T in (8, 16, 32):
            for Y0 in range(0, H, T):
                for X0 in range(0, W, T):
                    pl = tile_bound(r, X0, Y0, T)
                    if isinstance(pl, str):
                        continue
                    assert abs(float(pl[0])) < 5e5
                    if pl[0] > 2 and covered[Y0:Y0 + T, X0:X0 + T].any():
                        high_with_fragments += 1
    assert high_with_fragments > 0
