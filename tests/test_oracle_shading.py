"""Closed-form pins of the oracle's shading (ShadePixelUnsafe,
CCR/CUDAROPs.cu:472-514) and of its per-pixel attribute interpolation
(RasterizePixelAtomic, CCR/CUDAROPs.cu:347-380), from geometry computed
independently in float64.  Each check is chosen so that a plausible slip in
the oracle fails it: a flipped binormal, a transposed tangent frame, a
dropped / pi or light intensity, a negated light direction, a clamped
back-lit term (the reference's ClampFast(x, 0, x) returns x unchanged,
CCR/CUDAROPs.cu:36-40), the textbook reflection instead of the reference's
-d + 2 (d.n) n (CCR/Vector.h:76-78), affine instead of perspective-correct
interpolation, or a tangent multiplied by w like the normal.  CPU only."""
import math

import numpy as np
import pytest

import oracle as O
import scenes as S
from scene_util import vertex

L = np.array([0.577, -0.577, -0.577])   # light direction of ShadePixelUnsafe
K_DIFFUSE = 7.0 / math.pi                # light intensity / pi


def max_to_one(c):
    m = max(c[0], max(c[1], c[2]))
    return c / m if m > 1.0 else c


def close(got, want, rel=2e-5, abs_=2e-6):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    assert np.all(np.abs(got - want) <= abs_ + rel * np.abs(want)), (got, want)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def frames(seed, n):
    """n orthonormal (tangent, normal) pairs, the normals spread over the sphere."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nrm = unit(rng.normal(size=3))
        t = unit(np.cross(nrm, unit(rng.normal(size=3))))
        out.append((t.astype(np.float32), nrm.astype(np.float32)))
    return out


FLAT = (0.5, 0.5, 1.0)  # a normal-map sample whose tangent-space normal is exactly (0, 0, 1)


@pytest.mark.parametrize("seed", range(3))
def test_lambert_term_lit_and_back_lit(seed):
    """Flat normal map, no specular: colour = 0.05 + d (7 / pi) (-N.L), with N
    the interpolated normal and no clamp of the back-lit (negative) term."""
    diff = np.array([0.9, 0.55, 0.2], np.float32)
    vd = unit([0.1, -0.3, -1.0]).astype(np.float32)
    seen_back = seen_lit = False
    for t, n in frames(seed, 24):
        out = O.shade_pbr(diff, FLAT, (0, 0, 0), (0.5, 0.5, 0.5), n, t, vd)
        lam = -float(np.dot(n.astype(np.float64), L))
        want = max_to_one(0.05 + diff.astype(np.float64) * K_DIFFUSE * lam)
        close(out, want)
        seen_back |= lam < -0.1
        seen_lit |= lam > 0.1
    assert seen_back and seen_lit


def test_normal_map_axes_follow_tangent_binormal_normal():
    """Tangent frame {tan, -Cross(tan, n), n} as columns: with tan = +x and
    n = +z a red-only map bends N to +x, a green-only map to +y (the binormal
    is -(x cross z) = +y), and N = (0, 0, 1) stays n."""
    t = np.array([1, 0, 0], np.float32)
    n = np.array([0, 0, 1], np.float32)
    vd = unit([0.0, 0.2, -1.0]).astype(np.float32)
    diff = np.array([0.3, 0.3, 0.3], np.float32)
    for nrm, N in (((1.0, 0.5, 0.5), (1, 0, 0)), ((0.0, 0.5, 0.5), (-1, 0, 0)),
                   ((0.5, 1.0, 0.5), (0, 1, 0)), ((0.5, 0.0, 0.5), (0, -1, 0)), (FLAT, (0, 0, 1))):
        out = O.shade_pbr(diff, nrm, (0, 0, 0), (0.5, 0.5, 0.5), n, t, vd)
        want = 0.05 + 0.3 * K_DIFFUSE * -float(np.dot(N, L))
        close(out, [want] * 3)


@pytest.mark.parametrize("gloss", [0.0, 0.2, 0.5, 1.0])
def test_phong_peak_falloff_and_max_to_one(gloss):
    """N = +z: the reference's Reflect(L, N) = -L + 2 (L.N) N is
    (-0.577, 0.577, -0.577).  A view direction at angle theta from it gives
    spec * clamp01(|R| cos theta) ^ (25 gloss), with pow(x, 0) = 1; the
    result is divided by its largest channel when that exceeds 1."""
    t = np.array([1, 0, 0], np.float32)
    n = np.array([0, 0, 1], np.float32)
    R = np.array([-0.577, 0.577, -0.577])
    r_len = np.linalg.norm(R)
    axis = unit(np.cross(R, [0.0, 0.0, 1.0]))
    spec = np.array([1.0, 0.5, 0.25], np.float32)
    e = 25.0 * np.float32(gloss)
    for deg in (0.0, 2.0, 10.0, 35.0, 80.0, 120.0):
        th = math.radians(deg)
        # rotate R / |R| by theta about an axis perpendicular to it (Rodrigues)
        rh = R / r_len
        v = rh * math.cos(th) + np.cross(axis, rh) * math.sin(th) + axis * np.dot(axis, rh) * (1 - math.cos(th))
        out = O.shade_pbr((0, 0, 0), FLAT, spec, (gloss, 0, 0), n, t, v.astype(np.float32))
        a = min(max(r_len * math.cos(th), 0.0), 1.0)
        p = 1.0 if e == 0 else a ** e
        want = max_to_one(0.05 + spec.astype(np.float64) * p)
        close(out, want, rel=5e-5)
        if deg == 0.0 and gloss in (0.0, 0.2):
            assert out.max() == pytest.approx(1.0, abs=1e-6)  # saturated and rescaled


# ------------------------------------------------ interpolated attributes

def view_basis(cam):
    """The reference camera frame (CCR/Camera.cpp:26-31,97-117) in float64."""
    fwd = unit(cam.forward)
    right = unit(np.cross([0.0, 1.0, 0.0], fwd))
    up = unit(np.cross(fwd, right))
    return np.asarray(cam.position, np.float64), right, up, fwd


def pixel_ray_hit(cam, W, H, px, py, Q, N):
    """Where pixel (px, py)'s sample point meets the plane through Q with
    normal N.  Clip w = 1 - z_view (the projection keeps [3][3] = 1), so the
    pixel's points are x_view = a (1 - z), y_view = b (1 - z); returns the
    world point and its view-space z."""
    pos, right, up, fwd = view_basis(cam)
    t = math.tan(math.radians(cam.fov_deg) / 2)
    a = (2.0 * px / W - 1.0) * (W / H) * t
    b = (1.0 - 2.0 * py / H) * t
    c0 = np.dot(pos - Q, N)
    k = np.dot(a * right + b * up, N)
    f = np.dot(fwd, N)
    z = -(c0 + k) / (f - k)
    X = pos + right * a * (1 - z) + up * b * (1 - z) + fwd * z
    return X, z


def oblique_quad():
    rng = np.random.default_rng(5)
    e1 = unit([1.0, 0.25, 0.35])
    e2 = unit(np.cross(unit([0.2, 1.0, 0.3]), e1))
    nrm = unit(np.cross(e1, e2))
    Q = np.array([0.3, -0.2, -4.0])
    h = 1.6
    corners = [Q - h * e1 - h * e2, Q + h * e1 - h * e2, Q + h * e1 + h * e2, Q - h * e1 + h * e2]
    uvs = [(0, 0), (1, 0), (1, 1), (0, 1)]
    tan = e1.astype(np.float32)
    vs = [np.array(list(p) + list(uv) + list(nrm) + list(tan) + [1, 1, 1], np.float32) for p, uv in zip(corners, uvs)]
    m = S.Mesh(verts=np.stack(vs), indices=np.array([0, 1, 2, 0, 2, 3], np.uint32))
    cam = S.Camera(position=(1.2, 1.5, 0.5), forward=tuple(unit(np.array([1.2, 1.5, 0.5]) - Q)), fov_deg=60.0)
    del rng
    return m, cam, Q, e1, e2, nrm, h, np.asarray(corners)


def test_interpolated_uv_normal_tangent_and_view_direction():
    """Oblique textured quad: at every covered pixel, uv is the texture
    coordinate of the ray-plane intersection (perspective-correct), n the
    vertex normal, tan the vertex tangent divided by the pixel's clip w
    (tan is not multiplied back by w), and vd the normalized world-space
    barycentric blend of the vertices' normalized view directions."""
    m, cam, Q, e1, e2, nrm, h, P = oblique_quad()
    W, H = 160, 120
    vis, inp = O.shade_inputs(S.Scene([m]), cam, S.CULL_NONE, W, H)
    ys, xs = np.nonzero(vis != O.VIS_NONE)
    assert len(xs) > 1000
    pos = np.asarray(cam.position, np.float64)
    vdk = [unit(p - pos) for p in P]
    tris = {0: (0, 1, 2), 1: (0, 2, 3)}
    worst = np.zeros(4)
    for x, y in zip(xs, ys):
        X, z = pixel_ray_hit(cam, W, H, float(x), float(y), Q, nrm)
        d = X - Q
        uv = np.array([(np.dot(d, e1) / h + 1) / 2, (np.dot(d, e2) / h + 1) / 2])
        got = inp[y, x].astype(np.float64)
        worst[0] = max(worst[0], np.abs(got[0:2] - uv).max())
        worst[1] = max(worst[1], np.abs(got[2:5] - nrm).max())
        tan_w = e1 / (1.0 - z)
        worst[2] = max(worst[2], np.abs(got[5:8] - tan_w).max() / np.abs(tan_w).max())
        # world barycentrics of X in the winning triangle
        a, b, c = (P[i] for i in tris[int(vis[y, x])])
        T = np.stack([b - a, c - a], axis=1)
        l12 = np.linalg.lstsq(T, X - a, rcond=None)[0]
        lam = (1 - l12.sum(), l12[0], l12[1])
        ia, ib, ic = tris[int(vis[y, x])]
        vd = unit(lam[0] * vdk[ia] + lam[1] * vdk[ib] + lam[2] * vdk[ic])
        worst[3] = max(worst[3], np.abs(got[8:11] - vd).max())
    assert worst[0] < 5e-5, f"uv off by {worst[0]}"
    assert worst[1] < 2e-6, f"normal off by {worst[1]}"
    assert worst[2] < 2e-5, f"tangent / w off by {worst[2]} (relative)"
    assert worst[3] < 2e-5, f"view direction off by {worst[3]}"


def test_fetched_texel_is_the_ray_plane_intersections_texel():
    """Away from texel edges, the texel the oracle samples (a texture whose
    texels encode their own index) is the one under the exact intersection."""
    m, cam, Q, e1, e2, nrm, h, _ = oblique_quad()
    W, H, n = 160, 120, 64
    idx = np.arange(n * n, dtype=np.uint32)
    tex = (idx & 0xFF) | ((idx >> 8) << 8) | np.uint32(0xFF000000)
    tex = tex.reshape(n, n)
    vis, inp = O.shade_inputs(S.Scene([m]), cam, S.CULL_NONE, W, H)
    ys, xs = np.nonzero(vis != O.VIS_NONE)
    checked = 0
    for x, y in zip(xs, ys):
        X, _ = pixel_ray_hit(cam, W, H, float(x), float(y), Q, nrm)
        d = X - Q
        fu = ((np.dot(d, e1) / h + 1) / 2) * n + 0.5
        fv = ((np.dot(d, e2) / h + 1) / 2) * n + 0.5
        if min(abs(fu - round(fu)), abs(fv - round(fv))) < 0.02 or not (0 <= fu < n and 0 <= fv < n):
            continue
        rgb = O.sample(tex, inp[y, x, 0:2], S.SAMPLE_POINT)
        got = int(round(rgb[0] * 255)) + 256 * int(round(rgb[1] * 255))
        assert got == int(fu) + int(fv) * n, (x, y)
        checked += 1
    assert checked > 1000


def test_lambert_frame_of_a_flat_quad_end_to_end():
    """orc_render of a flat quad with constant maps (normal map 128,128,255,
    no specular) at several orientations, lit and back-lit: every covered
    pixel is pack(0.05 + d (7 / pi) (-N.L)), unclamped, with
    N = (tan_p + B_p) / 255 + n, where 2 * 128 / 255 - 1 = 1 / 255, the
    pixel's tangent tan_p is the vertex tangent over the pixel's clip w (the
    interpolation leaves it divided by w) and B_p = -(tan_p x n)."""
    d = np.array([0.05, 0.4, 0.9])
    dtex = np.full((4, 4), S.rgba(*(int(round(c * 255)) for c in d)), np.uint32)
    dd = np.array([int(round(c * 255)) / 255.0 for c in d])
    maps = [dtex, np.full((4, 4), S.rgba(128, 128, 255), np.uint32), np.full((4, 4), S.rgba(0, 0, 0), np.uint32),
            np.full((4, 4), S.rgba(128, 128, 128), np.uint32)]
    W, H = 96, 72
    checked_back = 0
    for k, (t, n) in enumerate(frames(11, 10)):
        t64, n64 = t.astype(np.float64), n.astype(np.float64)
        e2 = unit(np.cross(n64, t64))
        Q = np.array([0.0, 0.0, -3.0])
        corners = [Q - t64 - e2, Q + t64 - e2, Q + t64 + e2, Q - t64 + e2]
        vs = [vertex(p, uv=(0.5, 0.5), n=n, t=t) for p in corners]
        mesh = S.Mesh(verts=np.stack(vs), indices=np.array([0, 1, 2, 0, 2, 3], np.uint32), textures=maps)
        cam = S.Camera(position=tuple(Q + 3.5 * n64 + 0.01), forward=tuple(n64), fov_deg=50.0)
        vis, _, rgba, _ = O.render(S.Scene([mesh]), cam, S.Opts(cull_mode=S.CULL_NONE), W, H)
        ys, xs = np.nonzero(vis != O.VIS_NONE)
        assert len(xs) > 500
        lam_seen = []
        for x, y in zip(xs[::7], ys[::7]):
            _, z = pixel_ray_hit(cam, W, H, float(x), float(y), Q, n64)
            tp = t64 / (1.0 - z)
            N = (tp - np.cross(tp, n64)) / 255.0 + n64
            lam = -np.dot(N, L)
            want = max_to_one(0.05 + dd * K_DIFFUSE * lam)
            want_u8 = np.array([0 if not (v * 255 > 0) else min(int(v * 255), 255) for v in want])
            got = rgba[y, x:x + 1].view(np.uint8)[:3].astype(int)
            assert np.abs(got - want_u8).max() <= 1, (k, x, y, got, want_u8)
            lam_seen.append(lam)
        checked_back += max(lam_seen) < 0
    assert checked_back >= 2


# ----------------------------------------------------- vertex attributes

ROT90 = {
    "identity": np.eye(3),
    "x90": np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], float),
    "y90": np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], float),
    "z90": np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], float),
}


@pytest.mark.parametrize("name", list(ROT90))
def test_vertex_normals_and_tangents_under_exact_rotations(name):
    """World matrices with entries 0 / +-1 (and a translation): n' and tan'
    are the rotated input vectors exactly, and each component of vd is
    within 2 ulp of 1 (the vector's length) of normalize(world * p - camera
    position) in float64."""
    R = ROT90[name]
    world = np.eye(4)
    world[:3, :3] = R
    world[:3, 3] = (0.25, -0.5, 0.75)
    m = S.uv_sphere(24, 12)
    m.world = world.T.astype(np.float32).ravel()  # data[col][row]
    cam = S.Camera(position=(0.4, 0.7, 4.0))
    _, attr = O.vertex_transform(S.Scene([m]), cam, 320, 240)
    n_in = m.verts[:, 5:8].astype(np.float64)
    t_in = m.verts[:, 8:11].astype(np.float64)
    np.testing.assert_array_equal(attr[:, 2:5], (n_in @ R.T).astype(np.float32))
    np.testing.assert_array_equal(attr[:, 5:8], (t_in @ R.T).astype(np.float32))
    p = m.verts[:, 0:3].astype(np.float64) @ R.T + world[:3, 3]
    vd = p - np.asarray(cam.position, np.float64)
    vd /= np.linalg.norm(vd, axis=1, keepdims=True)
    err = np.abs(attr[:, 8:11] - vd).max()
    assert err <= 2 * np.spacing(np.float32(1.0)), err
