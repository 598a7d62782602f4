"""Synthetic benchmark scenes, built by the deterministic generators in
csrc/scenes.cpp (librast_scenes.so).

The reference renders OBJ assets loaded by its app shell (CCR/main.cpp:48-228,
CCR/ObjParser.cpp); the benchmark workloads here are procedural and seeded so
that the GPU path and the CPU oracle consume byte-identical inputs.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
SCENES_LIB = os.path.join(LIB_DIR, "librast_scenes.so")

TOPO_LIST, TOPO_STRIP = 0, 1
CULL_BACK, CULL_FRONT, CULL_NONE = 0, 1, 2
SAMPLE_POINT, SAMPLE_LINEAR = 0, 1
VIS_PBR, VIS_DEPTH, VIS_NORMAL = 0, 1, 2

SEED_BASE = 0x5EED0000


def rgba(r: int, g: int, b: int, a: int = 255) -> int:
    return (r & 255) | ((g & 255) << 8) | ((b & 255) << 16) | ((a & 255) << 24)


# Reference clear colour RGBColor(0.25, 0.25, 0.25) packed by (u8)(c * 255)
# (CCR/CUDARenderer.h:135, CCR/RGBRaw.h:23-32).
CLEAR_RGBA = rgba(63, 63, 63)
DEFAULT_NORMAL = rgba(128, 128, 255)
DEFAULT_SPECULAR = rgba(128, 128, 128)
DEFAULT_GLOSS = rgba(128, 128, 128)


@dataclass
class Mesh:
    verts: np.ndarray              # float32 [nv, 14]: p uv n tan c (IVertex)
    indices: np.ndarray            # uint32 [ni]
    topology: int = TOPO_LIST
    world: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32).ravel())
    textures: Optional[list] = None  # [diffuse, normal, specular, gloss], uint32 [h, w] or None

    @property
    def num_triangles(self) -> int:
        n = int(self.indices.size)
        return n // 3 if self.topology == TOPO_LIST else max(n - 2, 0)


@dataclass
class Scene:
    meshes: List[Mesh]

    @property
    def num_triangles(self) -> int:
        return sum(m.num_triangles for m in self.meshes)

    @property
    def num_vertices(self) -> int:
        return sum(int(m.verts.shape[0]) for m in self.meshes)


@dataclass
class Camera:
    position: tuple = (0.0, 0.0, 3.0)
    forward: tuple = (0.0, 0.0, 1.0)   # reference Camera::m_Forward (away from the view)
    fov_deg: float = 45.0
    near_z: float = 0.1
    far_z: float = 100.0


@dataclass
class Opts:
    cull_mode: int = CULL_BACK
    sample_state: int = SAMPLE_POINT
    visualisation: int = VIS_PBR
    clear_rgba: int = CLEAR_RGBA
    default_normal: int = DEFAULT_NORMAL
    default_specular: int = DEFAULT_SPECULAR
    default_gloss: int = DEFAULT_GLOSS


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SCENES_LIB):
            raise RuntimeError(f"{SCENES_LIB} is not built; run __graft_entry__.build()")
        L = C.CDLL(SCENES_LIB)
        vp, u32p = C.c_void_p, C.POINTER(C.c_uint32)
        L.rsc_uv_sphere.argtypes = [C.c_uint32, C.c_uint32, C.c_float, vp, u32p, vp, u32p]
        L.rsc_icosphere.argtypes = [C.c_uint32, C.c_float, vp, u32p, vp, u32p]
        L.rsc_terrain.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_float, vp, u32p,
                                  vp, u32p]
        L.rsc_texture_checker.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, vp]
        L.rsc_texture_noise.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, vp]
        L.rsc_world_matrix.argtypes = [vp, vp, C.c_float, vp]
        L.rsc_uniform.argtypes = [C.c_uint32, C.c_uint32, C.c_float, C.c_float]
        L.rsc_uniform.restype = C.c_float
        L.rsc_sin.argtypes = [C.c_double]
        L.rsc_sin.restype = C.c_double
        L.rsc_cos.argtypes = [C.c_double]
        L.rsc_cos.restype = C.c_double
        L.rsc_camera_path.argtypes = [C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float, vp]
        L.rsc_backdrop.argtypes = [vp, vp, C.c_float, C.c_float, C.c_float, C.c_float, vp, vp]
        _lib = L
    return _lib


def _mesh_from(fn, *args) -> Mesh:
    nv, ni = C.c_uint32(0), C.c_uint32(0)
    rc = fn(*args, None, C.byref(nv), None, C.byref(ni))
    if rc != 0:
        raise ValueError("generator rejected its arguments")
    v = np.zeros((nv.value, 14), np.float32)
    ix = np.zeros(ni.value, np.uint32)
    rc = fn(*args, v.ctypes.data, C.byref(nv), ix.ctypes.data, C.byref(ni))
    if rc != 0:
        raise RuntimeError("generator failed")
    return Mesh(verts=v, indices=ix)


def uv_sphere(slices=64, stacks=32, radius=1.0) -> Mesh:
    return _mesh_from(lib().rsc_uv_sphere, slices, stacks, radius)


def icosphere(subdiv=5, radius=1.0) -> Mesh:
    return _mesh_from(lib().rsc_icosphere, subdiv, radius)


def terrain(n=1024, size=80.0, amplitude=6.0, octaves=6, seed=SEED_BASE + 3, uv_tiling=1.0) -> Mesh:
    return _mesh_from(lib().rsc_terrain, n, size, amplitude, octaves, seed, uv_tiling)


def checker(w=256, h=256, square=32, c0=rgba(235, 235, 235), c1=rgba(200, 60, 40)) -> np.ndarray:
    t = np.zeros((h, w), np.uint32)
    if lib().rsc_texture_checker(w, h, square, c0, c1, t.ctypes.data) != 0:
        raise RuntimeError("checker failed")
    return t


def noise_texture(w: int, h: int, seed: int, kind: int) -> np.ndarray:
    t = np.zeros((h, w), np.uint32)
    if lib().rsc_texture_noise(w, h, seed, kind, t.ctypes.data) != 0:
        raise RuntimeError("noise texture failed")
    return t


def noise_maps(size: int, seed: int) -> list:
    # the generator releases the GIL inside its ctypes call: the four maps in parallel
    with ThreadPoolExecutor(4) as ex:
        return list(ex.map(lambda k: noise_texture(size, size, seed + k, k), range(4)))


def world_matrix(translate, euler, scale) -> np.ndarray:
    t = np.asarray(translate, np.float32)
    e = np.asarray(euler, np.float32)
    out = np.zeros(16, np.float32)
    lib().rsc_world_matrix(t.ctypes.data, e.ctypes.data, float(scale), out.ctypes.data)
    return out


def uniform(seed: int, index: int, lo: float, hi: float) -> float:
    return float(lib().rsc_uniform(seed, index, lo, hi))


def camera_path(seed: int, frames: int, terrain_size=80.0, height=9.0, pitch_deg=10.0) -> List[Camera]:
    out = np.zeros((frames, 6), np.float32)
    lib().rsc_camera_path(seed, frames, terrain_size, height, pitch_deg, out.ctypes.data)
    return [Camera(position=tuple(map(float, r[:3])), forward=tuple(map(float, r[3:]))) for r in out]


def backdrop(cam: Camera, aspect: float, d0=120.0, step=2.0) -> Mesh:
    v = np.zeros((24, 14), np.float32)
    ix = np.zeros(24, np.uint32)
    p = np.asarray(cam.position, np.float32)
    f = np.asarray(cam.forward, np.float32)
    lib().rsc_backdrop(p.ctypes.data, f.ctypes.data, cam.fov_deg, aspect, d0, step, v.ctypes.data, ix.ctypes.data)
    return Mesh(verts=v, indices=ix)


# --------------------------------------------------------------- workloads

@dataclass
class Workload:
    name: str
    scene: Scene
    camera: Camera
    width: int
    height: int
    opts: Opts
    tile: int = 16


def config1() -> Workload:
    """UV sphere 64x32 (4,032 tris), 256^2 checker, 640x480, FOV 45 at distance 3."""
    m = uv_sphere(64, 32, 1.0)
    m.textures = [checker(), None, None, None]
    return Workload("uvsphere64x32_640x480", Scene([m]), Camera(), 640, 480, Opts())


def config2(grid=16, subdiv=5, tex_size=2048) -> Workload:
    """16x16 icospheres (subdiv 5, 20,480 tris each, 5,242,880 total), seeded
    rotation and scale in [0.5, 1.5], four 2048^2 noise maps, 1920x1080."""
    seed = SEED_BASE + 2
    base = icosphere(subdiv, 1.0)
    maps = noise_maps(tex_size, seed)
    meshes = []
    sx, sy = 4.4, 2.5
    for j in range(grid):
        for i in range(grid):
            k = j * grid + i
            t = ((i - (grid - 1) / 2) * sx, ((grid - 1) / 2 - j) * sy, 0.0)
            e = tuple(uniform(seed, 3 * k + q, 0.0, 6.2831853) for q in range(3))
            s = uniform(seed, 1000 + k, 0.5, 1.5)
            meshes.append(Mesh(verts=base.verts, indices=base.indices, world=world_matrix(t, e, s), textures=maps))
    cam = Camera(position=(0.0, 0.0, 45.0))
    return Workload("icospheres16x16_s5_1920x1080", Scene(meshes), cam, 1920, 1080, Opts())


def terrain_camera() -> Camera:
    return camera_path(SEED_BASE + 3, 1)[0]


def config3(n=1024, tex_size=2048, width=3840, height=2160, camera: Optional[Camera] = None) -> Workload:
    """1024^2-quad fBm terrain (2,097,152 tris) at a 10 degree grazing angle
    plus 8 screen-covering backdrop triangles, 3840x2160."""
    seed = SEED_BASE + 3
    cam = camera or terrain_camera()
    ter = terrain(n, 80.0, 6.0, 6, seed, 1.0)
    maps = noise_maps(tex_size, seed)
    ter.textures = maps
    bd = backdrop(terrain_camera(), width / height)
    bd.textures = maps
    return Workload(f"terrain{n}_grazing10_{width}x{height}", Scene([ter, bd]), cam, width, height, Opts())


def train_test_paths(train=200, test=100):
    """Disjoint seeded camera paths over the terrain (different seeds)."""
    return (camera_path(SEED_BASE + 0x300, train), camera_path(SEED_BASE + 0x301, test))


@dataclass
class Job:
    """One render request of a batched workload."""
    name: str
    scene: Scene
    camera: Camera
    width: int
    height: int
    opts: Opts = field(default_factory=Opts)

    @property
    def pixels(self) -> int:
        return self.width * self.height


def orbit_camera(yaw: float, elevation: float, distance: float, fov_deg=45.0) -> Camera:
    """Camera on a sphere around the origin, looking at it (the reference's
    forward points from the target to the eye)."""
    ce = lib().rsc_cos(elevation)
    d = (ce * lib().rsc_sin(yaw), lib().rsc_sin(elevation), ce * lib().rsc_cos(yaw))
    return Camera(position=tuple(float(distance * q) for q in d), forward=tuple(float(q) for q in d),
                  fov_deg=fov_deg, near_z=0.1, far_z=100.0)


def config4(jobs=256, tex_size=512):
    """256 uneven render requests: resolution per axis log-uniform in
    [64, 1024]; each job one seeded icosphere (subdivision 3-6, 1,280-81,920
    tris) or terrain (2n^2 tris log-uniform in [1k, 200k]) with its own
    camera and a tex_size^2 RGBA8 noise texture."""
    seed = SEED_BASE + 4
    # the generators release the GIL inside ctypes calls: build textures in parallel
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        texs = list(ex.map(lambda j: noise_texture(tex_size, tex_size, seed + 1000 + j, 0), range(jobs)))
    ico = {}
    out = []
    for j in range(jobs):
        def u(q, lo, hi):
            return uniform(seed, 16 * j + q, lo, hi)
        W = int(round(math.exp(u(0, math.log(64), math.log(1024)))))
        H = int(round(math.exp(u(1, math.log(64), math.log(1024)))))
        tex = texs[j]
        if u(2, 0.0, 1.0) < 0.5:
            sub = min(3 + int(u(3, 0.0, 4.0)), 6)
            if sub not in ico:
                ico[sub] = icosphere(sub, 1.0)
            base = ico[sub]
            m = Mesh(verts=base.verts, indices=base.indices, textures=[tex, None, None, None])
            cam = orbit_camera(u(4, 0.0, 6.2831853), u(5, -0.9, 0.9), u(6, 2.2, 6.0))
            name = f"ico{sub}"
        else:
            tris = math.exp(u(3, math.log(1000), math.log(200000)))
            n = max(int(round(math.sqrt(tris / 2))), 23)
            m = terrain(n, 80.0, 6.0, 6, seed + 2000 + j, 1.0)
            m.textures = [tex, None, None, None]
            cam = camera_path(seed + 3000 + j, 1)[0]
            name = f"terrain{n}"
        out.append(Job(f"{name}_{W}x{H}", Scene([m]), cam, W, H))
    return out


def cube_directions():
    """Face directions of a cube map whose body diagonal is vertical, so no
    face looks straight up or down (the reference camera builds its frame from
    the world up axis, which a face direction must not be parallel to)."""
    a = math.atan(1.0 / math.sqrt(2.0))  # 35.26 deg: elevation of each face normal
    dirs = []
    for k in range(3):
        yaw = 2.0 * math.pi * k / 3.0
        ce, se = math.cos(a), math.sin(a)
        d = (ce * math.sin(yaw), se, ce * math.cos(yaw))
        dirs.append(d)
        dirs.append(tuple(-q for q in d))
    return dirs


@dataclass
class Flythrough:
    scene: Scene
    path: List[Camera]
    main: tuple        # (width, height) of the main view
    face: int          # cube face size
    opts: Opts = field(default_factory=Opts)

    def views(self, frame: int):
        """(camera, width, height) of the main view and the six cube faces."""
        cam = self.path[frame]
        out = [(cam, self.main[0], self.main[1])]
        for d in cube_directions():
            out.append((Camera(position=cam.position, forward=d, fov_deg=90.0, near_z=0.1, far_z=100.0),
                        self.face, self.face))
        return out


def config5(frames=300, n=1160, spheres=64, subdiv=5, tex_size=2048, main=(3840, 2160), face=1024) -> Flythrough:
    """Terrain of n^2 quads (1160^2: 2,691,200 tris) plus 64 instanced
    subdivision-5 icospheres (1,310,720 tris), 4,001,920 tris in all, with
    four tex_size^2 maps; a seeded camera path; per frame one main view and
    six cube-map faces."""
    seed = SEED_BASE + 5
    maps = noise_maps(tex_size, seed)
    ter = terrain(n, 80.0, 6.0, 6, seed, 1.0)
    ter.textures = maps
    base = icosphere(subdiv, 1.0)
    meshes = [ter]
    for k in range(spheres):
        t = (uniform(seed, 4 * k, -34.0, 34.0), uniform(seed, 4 * k + 1, 7.5, 14.0),
             uniform(seed, 4 * k + 2, -34.0, 34.0))
        e = tuple(uniform(seed, 1000 + 3 * k + q, 0.0, 6.2831853) for q in range(3))
        sc = uniform(seed, 4 * k + 3, 1.0, 2.5)
        meshes.append(Mesh(verts=base.verts, indices=base.indices, world=world_matrix(t, e, sc), textures=maps))
    path = camera_path(seed + 0x100, frames)
    return Flythrough(Scene(meshes), path, main, face)
