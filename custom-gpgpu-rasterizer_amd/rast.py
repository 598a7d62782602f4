"""ctypes binding of librast.so (the C ABI of include/rast.h).

Used by the tests and bench.py to drive the HIP rasterizer exactly the way a
foreign caller would: plain pointers and sizes through the C ABI.  The
Renderer class follows the reference host API (CCR/CUDARenderer.h:31-74):
LoadScene -> add_scene, SetupBins -> setup_bins, Render -> render,
Present -> readback.  It raises RastError with the library's message on
any failure; there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "librast.so")

RAST_OK = 0
VIS_NONE = 0xFFFFFFFF
BIN_EXACT, BIN_LEARNED, BIN_PREVFRAME = 0, 1, 2
DRAW_TEXTURE, DRAW_TEXTURE_GLOBAL = 0, 1
FLAG_TIMING, FLAG_NO_SHADE, FLAG_GRAPH = 1, 2, 4
RESOLVE_PER_TILE, RESOLVE_PERSISTENT, RESOLVE_PERSISTENT_SHADE = 0, 1, 2
DEBUG_VERTEX_POS, DEBUG_VERTEX_ATTR, DEBUG_TRI_RECORDS, DEBUG_TRI_ALIVE, DEBUG_TILE_COUNTS, DEBUG_BIG_LIST, DEBUG_TILE_CAPS = range(7)

EXPORTS = [
    "rast_last_error", "rast_device_query", "rast_renderer_create", "rast_renderer_destroy", "rast_setup_bins",
    "rast_resize", "rast_set_tuning", "rast_set_resolve_mode", "rast_mesh_create", "rast_mesh_destroy", "rast_mesh_set_world", "rast_mesh_set_textures",
    "rast_mesh_num_triangles", "rast_mesh_triangle_base", "rast_texture_create", "rast_texture_destroy",
    "rast_camera_frame", "rast_render", "rast_sync", "rast_readback", "rast_get_stats", "rast_frame_buffers",
    "rast_debug_read", "rast_binmodel_create", "rast_binmodel_destroy", "rast_binmodel_num_features",
    "rast_renderer_set_binmodel", "rast_binmodel_record", "rast_binmodel_clear_data", "rast_binmodel_train",
    "rast_binmodel_get_params", "rast_binmodel_set_params", "rast_binmodel_num_params", "rast_binmodel_forward",
    "rast_binmodel_loss_grad", "rast_render_batch", "rast_renderer_create_view", "rast_texture_create_pitched",
    "rast_timer_start", "rast_timer_stop", "rast_draw_texture", "rast_binmodel_forward_noalloc",
]


class RastError(RuntimeError):
    pass


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("forward", C.c_float * 3), ("fov_deg", C.c_float),
                ("near_z", C.c_float), ("far_z", C.c_float)]


class FrameOpts(C.Structure):
    _fields_ = [("cull_mode", C.c_int), ("sample_state", C.c_int), ("visualisation", C.c_int),
                ("bin_policy", C.c_int), ("clear_rgba", C.c_uint32), ("default_normal", C.c_uint32),
                ("default_specular", C.c_uint32), ("default_gloss", C.c_uint32), ("flags", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("num_triangles", C.c_uint64), ("num_visible", C.c_uint64), ("num_bin_pairs", C.c_uint64),
                ("num_big", C.c_uint32), ("num_tiles", C.c_uint32), ("overflow_tiles", C.c_uint32),
                ("capacity_total", C.c_uint32), ("rerenders", C.c_uint32), ("overflow_entries", C.c_uint32)] + [
        (n, C.c_float) for n in ("ms_vertex", "ms_setup", "ms_count", "ms_scan", "ms_fill", "ms_fallback",
                                 "ms_raster", "ms_shade", "ms_total", "ms_model")]


class FrameConsts(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("viewproj", C.c_float * 16),
                ("cam_pos", C.c_float * 3), ("cam_fwd", C.c_float * 3), ("cam_right", C.c_float * 3),
                ("cam_up", C.c_float * 3)]


class Job(C.Structure):
    _fields_ = [("target", C.c_void_p), ("camera", Camera), ("opts", FrameOpts)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_int),
                ("clock_khz", C.c_int), ("global_mem_bytes", C.c_uint64), ("lds_bytes_per_block", C.c_int),
                ("wavefront_size", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RastError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name).restype = C.c_int
        vp = C.c_void_p
        L.rast_last_error.argtypes = [C.c_char_p, C.c_size_t]
        L.rast_device_query.argtypes = [C.c_int, C.POINTER(DeviceInfo)]
        L.rast_renderer_create.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.POINTER(vp)]
        L.rast_renderer_destroy.argtypes = [vp]
        L.rast_renderer_create_view.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(vp)]
        L.rast_setup_bins.argtypes = [vp, C.c_uint32, C.c_uint32]
        L.rast_resize.argtypes = [vp, C.c_uint32, C.c_uint32]
        L.rast_set_tuning.argtypes = [vp, C.c_uint32, C.c_uint32]
        L.rast_set_resolve_mode.argtypes = [vp, C.c_int]
        L.rast_mesh_create.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, C.c_int, vp, C.POINTER(vp)]
        L.rast_mesh_destroy.argtypes = [vp]
        L.rast_mesh_set_world.argtypes = [vp, vp]
        L.rast_mesh_set_textures.argtypes = [vp, vp, vp, vp, vp]
        L.rast_mesh_num_triangles.argtypes = [vp, C.POINTER(C.c_uint32)]
        L.rast_mesh_triangle_base.argtypes = [vp, C.POINTER(C.c_uint64)]
        L.rast_texture_create.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.POINTER(vp)]
        L.rast_texture_destroy.argtypes = [vp]
        L.rast_texture_create_pitched.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_size_t, C.POINTER(vp)]
        L.rast_camera_frame.argtypes = [C.POINTER(Camera), C.c_uint32, C.c_uint32, C.POINTER(FrameConsts)]
        L.rast_render.argtypes = [vp, C.POINTER(Camera), C.POINTER(FrameOpts)]
        L.rast_sync.argtypes = [vp]
        L.rast_readback.argtypes = [vp, vp, vp, vp]
        L.rast_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.rast_frame_buffers.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
        L.rast_debug_read.argtypes = [vp, C.c_int, vp, C.c_size_t, C.POINTER(C.c_size_t)]
        L.rast_binmodel_create.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(vp)]
        L.rast_binmodel_destroy.argtypes = [vp]
        L.rast_binmodel_num_features.argtypes = []
        L.rast_renderer_set_binmodel.argtypes = [vp, vp, C.c_float, C.c_uint32]
        L.rast_binmodel_record.argtypes = [vp, vp]
        L.rast_binmodel_clear_data.argtypes = [vp]
        L.rast_binmodel_train.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.POINTER(C.c_float)]
        L.rast_binmodel_get_params.argtypes = [vp, vp, C.c_size_t]
        L.rast_binmodel_set_params.argtypes = [vp, vp, C.c_size_t]
        L.rast_binmodel_num_params.argtypes = [vp, C.POINTER(C.c_size_t)]
        L.rast_binmodel_forward.argtypes = [vp, vp, C.c_uint32, vp]
        L.rast_binmodel_forward_noalloc.argtypes = [vp, vp, C.c_uint32, vp]
        L.rast_binmodel_loss_grad.argtypes = [vp, vp, vp, C.c_uint32, C.c_float, vp, vp]
        L.rast_render_batch.argtypes = [C.POINTER(Job), C.c_uint32]
        L.rast_timer_start.argtypes = [vp]
        L.rast_timer_stop.argtypes = [vp, C.POINTER(C.c_float)]
        L.rast_draw_texture.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().rast_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    if rc != RAST_OK:
        raise RastError(f"{what} failed ({rc}): {last_error()}")


def camera_struct(cam: S.Camera) -> Camera:
    c = Camera()
    c.position[:] = [float(x) for x in cam.position]
    c.forward[:] = [float(x) for x in cam.forward]
    c.fov_deg, c.near_z, c.far_z = cam.fov_deg, cam.near_z, cam.far_z
    return c


def opts_struct(o: S.Opts, policy: int = BIN_EXACT, flags: int = 0) -> FrameOpts:
    f = FrameOpts()
    f.cull_mode, f.sample_state, f.visualisation = o.cull_mode, o.sample_state, o.visualisation
    f.bin_policy = policy
    f.clear_rgba, f.default_normal = o.clear_rgba, o.default_normal
    f.default_specular, f.default_gloss = o.default_specular, o.default_gloss
    f.flags = flags
    return f


def camera_frame(cam: S.Camera, width: int, height: int) -> FrameConsts:
    fc = FrameConsts()
    check(lib().rast_camera_frame(C.byref(camera_struct(cam)), width, height, C.byref(fc)), "rast_camera_frame")
    return fc


def device_info(device: int = 0) -> dict:
    d = DeviceInfo()
    check(lib().rast_device_query(device, C.byref(d)), "rast_device_query")
    return {"name": d.name.decode(), "arch": d.arch.decode(), "compute_units": d.compute_units,
            "clock_khz": d.clock_khz, "global_mem_bytes": d.global_mem_bytes}


class Renderer:
    """CUDARenderer equivalent over the C ABI."""

    def __init__(self, width: int, height: int, device: int = 0, tile: int = 16, big_threshold: int = 256,
                 owner: Optional["Renderer"] = None):
        self.h = C.c_void_p()
        if owner is None:
            check(lib().rast_renderer_create(device, width, height, C.byref(self.h)), "rast_renderer_create")
        else:
            check(lib().rast_renderer_create_view(owner.h, width, height, C.byref(self.h)),
                  "rast_renderer_create_view")
        self.width, self.height = width, height
        self.meshes, self.textures = [], {}
        self.owner = owner
        self.setup_bins(tile, big_threshold)

    def view(self, width: int, height: int, tile: int = 16, big_threshold: int = 256) -> "Renderer":
        """Another framebuffer drawing this renderer's meshes."""
        return Renderer(width, height, tile=tile, big_threshold=big_threshold, owner=self)

    def setup_bins(self, tile: int, big_threshold: int = 256) -> None:
        check(lib().rast_setup_bins(self.h, tile, big_threshold), "rast_setup_bins")
        self.tile = tile

    def set_tuning(self, micro_limit: int = 16, small_limit: int = 16) -> None:
        check(lib().rast_set_tuning(self.h, micro_limit, small_limit), "rast_set_tuning")

    def set_resolve_mode(self, mode: int) -> None:
        """RESOLVE_PER_TILE, RESOLVE_PERSISTENT or RESOLVE_PERSISTENT_SHADE."""
        check(lib().rast_set_resolve_mode(self.h, mode), "rast_set_resolve_mode")

    def set_queue_depth(self, margin: float = 1.0, slack: int = 0) -> None:
        """Queue depth of the previous-frame heuristic, ceil(prev * margin) +
        slack per tile (detaches any capacity model)."""
        check(lib().rast_renderer_set_binmodel(self.h, None, margin, slack), "rast_renderer_set_binmodel")

    def texture(self, texels: np.ndarray) -> C.c_void_p:
        key = id(texels)
        if key not in self.textures:
            t = np.ascontiguousarray(texels, np.uint32)
            h = C.c_void_p()
            check(lib().rast_texture_create(self.h, t.ctypes.data, t.shape[1], t.shape[0], C.byref(h)),
                  "rast_texture_create")
            self.textures[key] = (h, texels)
        return self.textures[key][0]

    def add_mesh(self, m: S.Mesh) -> C.c_void_p:
        v = np.ascontiguousarray(m.verts, np.float32)
        ix = np.ascontiguousarray(m.indices, np.uint32)
        w = np.ascontiguousarray(m.world, np.float32)
        h = C.c_void_p()
        check(lib().rast_mesh_create(self.h, v.ctypes.data, v.shape[0], ix.ctypes.data, ix.size, m.topology,
                                     w.ctypes.data, C.byref(h)), "rast_mesh_create")
        if m.textures is not None:
            ts = [self.texture(t) if t is not None else None for t in m.textures]
            check(lib().rast_mesh_set_textures(h, *ts), "rast_mesh_set_textures")
        self.meshes.append(h)
        return h

    def add_scene(self, scene: S.Scene) -> None:
        for m in scene.meshes:
            self.add_mesh(m)

    def render(self, cam: S.Camera, opts: S.Opts, policy: int = BIN_EXACT, flags: int = 0) -> None:
        check(lib().rast_render(self.h, C.byref(camera_struct(cam)), C.byref(opts_struct(opts, policy, flags))),
              "rast_render")

    def sync(self) -> None:
        check(lib().rast_sync(self.h), "rast_sync")

    def start_timer(self) -> None:
        """StartTimer (CCR/CUDARenderer.cu:398-402): a HIP event on the renderer's stream."""
        check(lib().rast_timer_start(self.h), "rast_timer_start")

    def stop_timer(self) -> float:
        """StopTimer: milliseconds of device time since start_timer."""
        ms = C.c_float(0)
        check(lib().rast_timer_stop(self.h, C.byref(ms)), "rast_timer_stop")
        return ms.value

    def draw_texture(self, texels: np.ndarray, mode: int = DRAW_TEXTURE_GLOBAL, stretch: bool = True,
                     sample_state: int = S.SAMPLE_POINT) -> None:
        """DrawTexture / DrawTextureGlobal (CCR/CUDARenderer.cu:1969-2012): the
        frame becomes the texture; read it with readback()."""
        check(lib().rast_draw_texture(self.h, self.texture(texels), mode, int(stretch), sample_state),
              "rast_draw_texture")

    def readback(self, vis=True, depth=True, rgba=True):
        out = []
        arrs = []
        for want, dt in ((vis, np.uint32), (depth, np.float32), (rgba, np.uint32)):
            a = np.empty((self.height, self.width), dt) if want else None
            arrs.append(a)
        check(lib().rast_readback(self.h, *[a.ctypes.data if a is not None else None for a in arrs]),
              "rast_readback")
        out = arrs
        return tuple(out)

    def stats(self) -> dict:
        s = Stats()
        check(lib().rast_get_stats(self.h, C.byref(s)), "rast_get_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def debug_read(self, which: int, dtype=np.uint8) -> np.ndarray:
        size = C.c_size_t(0)
        check(lib().rast_debug_read(self.h, which, None, 0, C.byref(size)), "rast_debug_read")
        buf = np.empty(size.value, np.uint8)
        check(lib().rast_debug_read(self.h, which, buf.ctypes.data, size.value, C.byref(size)), "rast_debug_read")
        return buf.view(dtype)

    def close(self) -> None:
        if self.h:
            check(lib().rast_renderer_destroy(self.h), "rast_renderer_destroy")
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class BinModel:
    """Learned bin-capacity MLP bound to a renderer."""

    def __init__(self, renderer: Renderer, hidden: int = 64, seed: int = 1):
        self.h = C.c_void_p()
        check(lib().rast_binmodel_create(renderer.h, hidden, seed, C.byref(self.h)), "rast_binmodel_create")
        self.hidden = hidden
        n = C.c_size_t(0)
        check(lib().rast_binmodel_num_params(self.h, C.byref(n)), "rast_binmodel_num_params")
        self.num_params = n.value

    def attach(self, renderer: Renderer, margin: float = 1.25, slack: int = 4) -> None:
        check(lib().rast_renderer_set_binmodel(renderer.h, self.h, margin, slack), "rast_renderer_set_binmodel")

    def detach(self, renderer: Renderer) -> None:
        check(lib().rast_renderer_set_binmodel(renderer.h, None, 1.0, 0), "rast_renderer_set_binmodel")

    def record(self, renderer: Renderer) -> None:
        check(lib().rast_binmodel_record(self.h, renderer.h), "rast_binmodel_record")

    def clear(self) -> None:
        check(lib().rast_binmodel_clear_data(self.h), "rast_binmodel_clear_data")

    def train(self, epochs: int, batch: int = 8192, lr: float = 2e-3, tau: float = 0.9) -> float:
        loss = C.c_float(0)
        check(lib().rast_binmodel_train(self.h, epochs, batch, lr, tau, C.byref(loss)), "rast_binmodel_train")
        return loss.value

    def params(self) -> np.ndarray:
        p = np.empty(self.num_params, np.float32)
        check(lib().rast_binmodel_get_params(self.h, p.ctypes.data, p.size), "rast_binmodel_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(lib().rast_binmodel_set_params(self.h, p.ctypes.data, p.size), "rast_binmodel_set_params")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty(x.shape[0], np.float32)
        check(lib().rast_binmodel_forward(self.h, x.ctypes.data, x.shape[0], out.ctypes.data), "rast_binmodel_forward")
        return out

    def loss_grad(self, x: np.ndarray, y: np.ndarray, tau: float):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(self.num_params, np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().rast_binmodel_loss_grad(self.h, x.ctypes.data, y.ctypes.data, x.shape[0], tau,
                                            loss.ctypes.data, g.ctypes.data), "rast_binmodel_loss_grad")
        return float(loss[0]), g

    def close(self) -> None:
        if self.h:
            check(lib().rast_binmodel_destroy(self.h), "rast_binmodel_destroy")
            self.h = C.c_void_p()


def render_batch(jobs) -> None:
    """Enqueue [(renderer, camera, opts, policy, flags), ...] as one graph launch."""
    arr = (Job * len(jobs))()
    for j, (r, cam, opts, policy, flags) in enumerate(jobs):
        arr[j].target = r.h
        arr[j].camera = camera_struct(cam)
        arr[j].opts = opts_struct(opts, policy, flags)
    check(lib().rast_render_batch(arr, len(jobs)), "rast_render_batch")


def render_workload(w: S.Workload, policy: int = BIN_EXACT, tile: Optional[int] = None):
    """Render one frame of a workload; returns (vis, depth, rgba, stats)."""
    with Renderer(w.width, w.height, tile=tile or w.tile) as r:
        r.add_scene(w.scene)
        r.render(w.camera, w.opts, policy)
        vis, depth, rgba = r.readback()
        st = r.stats()
    return vis, depth, rgba, st
