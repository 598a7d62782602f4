"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never by the product path.  The C sources restate the reference's
arithmetic (see rast_oracle.h for the file:line list).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

VIS_NONE = 0xFFFFFFFF


class OrcTexture(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("w", C.c_uint32), ("h", C.c_uint32)]


class OrcMesh(C.Structure):
    _fields_ = [
        ("verts", C.c_void_p),
        ("num_verts", C.c_uint32),
        ("indices", C.c_void_p),
        ("num_indices", C.c_uint32),
        ("topology", C.c_int),
        ("world", C.c_float * 16),
        ("tex", OrcTexture * 4),
    ]


class OrcCamera(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("forward", C.c_float * 3),
        ("fov_deg", C.c_float),
        ("near_z", C.c_float),
        ("far_z", C.c_float),
    ]


class OrcOpts(C.Structure):
    _fields_ = [
        ("cull_mode", C.c_int),
        ("sample_state", C.c_int),
        ("visualisation", C.c_int),
        ("clear_rgba", C.c_uint32),
        ("default_normal", C.c_uint32),
        ("default_specular", C.c_uint32),
        ("default_gloss", C.c_uint32),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("num_triangles", C.c_uint64),
        ("num_visible", C.c_uint64),
        ("num_covered_pixels", C.c_uint64),
        ("distinct_texels", C.c_uint64),
        ("distinct_ref_verts", C.c_uint64),
    ]


class OrcFrame(C.Structure):
    _fields_ = [
        ("view", C.c_float * 16),
        ("proj", C.c_float * 16),
        ("viewproj", C.c_float * 16),
        ("cam_pos", C.c_float * 3),
        ("cam_fwd", C.c_float * 3),
        ("cam_right", C.c_float * 3),
        ("cam_up", C.c_float * 3),
    ]


_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    args = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, capture_output=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_camera_frame.argtypes = [C.POINTER(OrcCamera), C.c_uint32, C.c_uint32, C.POINTER(OrcFrame)]
        L.orc_render.argtypes = [C.POINTER(OrcMesh), C.c_int, C.POINTER(OrcCamera), C.POINTER(OrcOpts),
                                 C.c_uint32, C.c_uint32, C.c_int, vp, vp, vp, C.POINTER(OrcStats)]
        L.orc_render.restype = C.c_int
        L.orc_setup.argtypes = [C.POINTER(OrcMesh), C.c_int, C.POINTER(OrcFrame), C.c_int, C.c_uint32,
                                C.c_uint32, vp, vp]
        L.orc_vertex_transform.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcFrame), vp, vp]
        L.orc_bin_counts.argtypes = [C.POINTER(OrcMesh), C.c_int, C.POINTER(OrcCamera), C.c_int, C.c_uint32,
                                     C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, C.c_uint32,
                                     C.POINTER(C.c_uint32)]
        L.orc_render_bruteforce.argtypes = [C.POINTER(OrcMesh), C.c_int, C.POINTER(OrcCamera), C.c_int,
                                            C.c_uint32, C.c_uint32, vp, vp]
        L.orc_sample.argtypes = [C.POINTER(OrcTexture), vp, C.c_int, vp]
        L.orc_shade_pbr.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_shade_inputs.argtypes = [C.POINTER(OrcMesh), C.c_int, C.POINTER(OrcCamera), C.c_int, C.c_uint32,
                                       C.c_uint32, vp, vp]
        L.orc_shade_inputs.restype = C.c_int
        L.orc_pack_rgba.argtypes = [vp]
        L.orc_pack_rgba.restype = C.c_uint32
        L.orc_mat4_mul.argtypes = [vp, vp, vp]
        L.orc_mat4_inverse.argtypes = [vp, vp]
        L.orc_mlp_num_params.argtypes = [C.c_int, C.c_int]
        L.orc_mlp_num_params.restype = C.c_size_t
        L.orc_mlp_forward.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, vp]
        L.orc_mlp_loss_grad.argtypes = [vp, C.c_int, C.c_int, vp, vp, C.c_int, C.c_float, vp]
        L.orc_mlp_loss_grad.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class _MeshPack:
    """Keeps numpy buffers alive while the oracle reads them."""

    def __init__(self, scene):
        self.keep = []
        n = len(scene.meshes)
        self.arr = (OrcMesh * n)()
        for i, m in enumerate(scene.meshes):
            v = np.ascontiguousarray(m.verts, dtype=np.float32)
            ix = np.ascontiguousarray(m.indices, dtype=np.uint32)
            self.keep += [v, ix]
            om = self.arr[i]
            om.verts = _ptr(v)
            om.num_verts = v.shape[0]
            om.indices = _ptr(ix)
            om.num_indices = ix.size
            om.topology = m.topology
            om.world[:] = [float(x) for x in np.asarray(m.world, dtype=np.float32).ravel()]
            for s in range(4):
                t = m.textures[s] if m.textures is not None else None
                if t is None:
                    om.tex[s].texels = None
                    om.tex[s].w = om.tex[s].h = 0
                else:
                    tt = np.ascontiguousarray(t, dtype=np.uint32)
                    self.keep.append(tt)
                    om.tex[s].texels = _ptr(tt)
                    om.tex[s].h, om.tex[s].w = tt.shape


def camera_struct(cam) -> OrcCamera:
    c = OrcCamera()
    c.position[:] = [float(x) for x in cam.position]
    c.forward[:] = [float(x) for x in cam.forward]
    c.fov_deg, c.near_z, c.far_z = cam.fov_deg, cam.near_z, cam.far_z
    return c


def opts_struct(opts) -> OrcOpts:
    o = OrcOpts()
    o.cull_mode = opts.cull_mode
    o.sample_state = opts.sample_state
    o.visualisation = opts.visualisation
    o.clear_rgba = opts.clear_rgba
    o.default_normal = opts.default_normal
    o.default_specular = opts.default_specular
    o.default_gloss = opts.default_gloss
    return o


def camera_frame(cam, width: int, height: int) -> OrcFrame:
    fr = OrcFrame()
    lib().orc_camera_frame(C.byref(camera_struct(cam)), width, height, C.byref(fr))
    return fr


def render(scene, cam, opts, width: int, height: int, threads: int = 1, want_rgba: bool = True,
           want_stats: bool = False):
    """Full frame on the CPU: returns (vis u32[H,W], depth f32[H,W], rgba u32[H,W] | None, stats | None)."""
    pack = _MeshPack(scene)
    vis = np.empty((height, width), np.uint32)
    depth = np.empty((height, width), np.float32)
    rgba = np.empty((height, width), np.uint32) if want_rgba else None
    st = OrcStats() if want_stats else None
    rc = lib().orc_render(pack.arr, len(scene.meshes), C.byref(camera_struct(cam)), C.byref(opts_struct(opts)),
                          width, height, threads, _ptr(vis), _ptr(depth),
                          _ptr(rgba) if rgba is not None else None, C.byref(st) if st is not None else None)
    if rc != 0:
        raise RuntimeError(f"orc_render failed: {rc}")
    stats = None
    if st is not None:
        stats = {k: int(getattr(st, k)) for k, _ in OrcStats._fields_}
    return vis, depth, rgba, stats


def render_bruteforce(scene, cam, cull_mode: int, width: int, height: int):
    pack = _MeshPack(scene)
    vis = np.empty((height, width), np.uint32)
    depth = np.empty((height, width), np.float32)
    rc = lib().orc_render_bruteforce(pack.arr, len(scene.meshes), C.byref(camera_struct(cam)), cull_mode,
                                     width, height, _ptr(vis), _ptr(depth))
    if rc != 0:
        raise RuntimeError(f"orc_render_bruteforce failed: {rc}")
    return vis, depth


def setup(scene, cam, cull_mode: int, width: int, height: int):
    """Per-triangle setup records (all triangles, global ID order): flags u8[T], rec u32[T,12]."""
    pack = _MeshPack(scene)
    fr = camera_frame(cam, width, height)
    T = scene.num_triangles
    flags = np.zeros(T, np.uint8)
    rec = np.zeros((T, 12), np.uint32)
    rc = lib().orc_setup(pack.arr, len(scene.meshes), C.byref(fr), cull_mode, width, height, _ptr(flags), _ptr(rec))
    if rc != 0:
        raise RuntimeError("orc_setup failed")
    return flags, rec


def vertex_transform(scene, cam, width: int, height: int, mesh_index: int = 0):
    pack = _MeshPack(scene)
    fr = camera_frame(cam, width, height)
    nv = scene.meshes[mesh_index].verts.shape[0]
    pos = np.zeros((nv, 4), np.float32)
    attr = np.zeros((nv, 16), np.float32)
    lib().orc_vertex_transform(C.byref(pack.arr[mesh_index]), C.byref(fr), _ptr(pos), _ptr(attr))
    return pos, attr


def bin_counts(scene, cam, cull_mode: int, width: int, height: int, tile_w: int, tile_h: int,
               big_threshold: int):
    pack = _MeshPack(scene)
    ntx, nty = -(-width // tile_w), -(-height // tile_h)
    counts = np.zeros((nty, ntx), np.uint32)
    cap = 1 << 16
    big = np.zeros(cap, np.uint32)
    nb = C.c_uint32(0)
    rc = lib().orc_bin_counts(pack.arr, len(scene.meshes), C.byref(camera_struct(cam)), cull_mode, width, height,
                              tile_w, tile_h, big_threshold, _ptr(counts), _ptr(big), cap, C.byref(nb))
    if rc != 0:
        raise RuntimeError("orc_bin_counts failed")
    return counts, big[: min(nb.value, cap)].copy()


def sample(texels: np.ndarray, uv, mode: int):
    t = OrcTexture()
    tt = np.ascontiguousarray(texels, dtype=np.uint32)
    t.texels = _ptr(tt)
    t.h, t.w = tt.shape
    uvb = np.asarray(uv, np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_sample(C.byref(t), _ptr(uvb), mode, _ptr(out))
    return out


def shade_pbr(diff, nrm, spc, gls, n, tan, vd) -> np.ndarray:
    """The oracle's ShadePixelUnsafe on given map samples and attributes."""
    args = [np.ascontiguousarray(a, np.float32) for a in (diff, nrm, spc, gls, n, tan, vd)]
    out = np.zeros(3, np.float32)
    lib().orc_shade_pbr(*[_ptr(a) for a in args], _ptr(out))
    return out


def shade_inputs(scene, cam, cull_mode: int, width: int, height: int):
    """Per-pixel shading inputs of the oracle's frame: (vis u32[H,W],
    inputs f32[H,W,11] = u v n.xyz tan.xyz vd.xyz, NaN where nothing is visible)."""
    pack = _MeshPack(scene)
    vis = np.empty((height, width), np.uint32)
    out = np.empty((height, width, 11), np.float32)
    rc = lib().orc_shade_inputs(pack.arr, len(scene.meshes), C.byref(camera_struct(cam)), cull_mode, width, height,
                                _ptr(vis), _ptr(out))
    if rc != 0:
        raise RuntimeError(f"orc_shade_inputs failed: {rc}")
    return vis, out


def pack_rgba(rgb) -> int:
    a = np.asarray(rgb, np.float32)
    return int(lib().orc_pack_rgba(_ptr(a)))


def mat4_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32).ravel()
    b = np.ascontiguousarray(b, np.float32).ravel()
    o = np.zeros(16, np.float32)
    lib().orc_mat4_mul(_ptr(a), _ptr(b), _ptr(o))
    return o


def mat4_inverse(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32).ravel()
    o = np.zeros(16, np.float32)
    lib().orc_mat4_inverse(_ptr(a), _ptr(o))
    return o


def mlp_num_params(in_dim: int, hidden: int) -> int:
    return int(lib().orc_mlp_num_params(in_dim, hidden))


def mlp_forward(params: np.ndarray, in_dim: int, hidden: int, x: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32)
    out = np.zeros(xx.shape[0], np.float32)
    lib().orc_mlp_forward(_ptr(p), in_dim, hidden, _ptr(xx), xx.shape[0], _ptr(out))
    return out


def mlp_loss_grad(params: np.ndarray, in_dim: int, hidden: int, x: np.ndarray, target: np.ndarray, tau: float):
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32)
    tt = np.ascontiguousarray(target, np.float32)
    g = np.zeros_like(p)
    loss = lib().orc_mlp_loss_grad(_ptr(p), in_dim, hidden, _ptr(xx), _ptr(tt), xx.shape[0], tau, _ptr(g))
    return float(loss), g
