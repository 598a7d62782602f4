"""Asset formats either side of the rasterizer (librast_io.so over ctypes):
Wavefront OBJ meshes in (the reference's ObjParser, CCR/ObjParser.cpp:23-584),
PNG textures in (IMG_Load in CUDATexture::Create, CCR/CUDATexture.cpp:24-45)
and 32-bit BMP framebuffers out (SaveBackbufferToImage,
CCR/ERenderer.cpp:174-177)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB = os.path.join(HERE, "lib", "librast_io.so")

VERTEX_UV, VERTEX_NORMAL, VERTEX_TANGENT = 1, 4, 8

_lib = None


class AssetError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(IO_LIB):
            raise AssetError(f"{IO_LIB} is not built; run __graft_entry__.build()")
        L = C.CDLL(IO_LIB)
        vp, u32p = C.c_void_p, C.POINTER(C.c_uint32)
        L.rio_obj_parse.argtypes = [C.c_char_p, C.c_size_t, C.c_int, C.POINTER(vp), u32p, C.POINTER(vp), u32p,
                                    C.POINTER(C.c_int)]
        L.rio_obj_load.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp), u32p, C.POINTER(vp), u32p, C.POINTER(C.c_int)]
        L.rio_free.argtypes = [vp]
        L.rio_free.restype = None
        L.rio_write_bmp.argtypes = [C.c_char_p, vp, C.c_uint32, C.c_uint32]
        L.rio_last_error.argtypes = [C.c_char_p, C.c_size_t]
        L.rio_png_decode.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(vp), u32p, u32p]
        L.rio_png_load.argtypes = [C.c_char_p, C.POINTER(vp), u32p, u32p]
        L.rio_png_encode.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(vp), C.POINTER(C.c_size_t)]
        for f in ("rio_obj_parse", "rio_obj_load", "rio_write_bmp", "rio_last_error", "rio_png_decode",
                  "rio_png_load", "rio_png_encode"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def _error(what: str) -> AssetError:
    buf = C.create_string_buffer(512)
    lib().rio_last_error(buf, 512)
    return AssetError(f"{what}: {buf.value.decode(errors='replace')}")


def _take(call, what):
    v, i = C.c_void_p(), C.c_void_p()
    nv, ni, vt = C.c_uint32(0), C.c_uint32(0), C.c_int(0)
    if call(C.byref(v), C.byref(nv), C.byref(i), C.byref(ni), C.byref(vt)) != 0:
        raise _error(what)
    try:
        verts = np.ctypeslib.as_array(C.cast(v, C.POINTER(C.c_float)), (nv.value, 14)).copy()
        idx = np.ctypeslib.as_array(C.cast(i, C.POINTER(C.c_uint32)), (ni.value,)).copy()
    finally:
        lib().rio_free(v)
        lib().rio_free(i)
    return S.Mesh(verts=verts, indices=idx), vt.value


def parse_obj(text: bytes, invert_y: bool = False):
    """(Mesh, vertex_type flags) from the bytes of an OBJ file."""
    return _take(lambda *a: lib().rio_obj_parse(text, len(text), int(invert_y), *a), "rio_obj_parse")


def load_obj(path: str, invert_y: bool = False):
    return _take(lambda *a: lib().rio_obj_load(path.encode(), int(invert_y), *a), "rio_obj_load")


def write_bmp(path: str, rgba: np.ndarray) -> None:
    a = np.ascontiguousarray(rgba, np.uint32)
    if lib().rio_write_bmp(path.encode(), a.ctypes.data, a.shape[1], a.shape[0]) != 0:
        raise _error("rio_write_bmp")


def read_bmp(path: str) -> np.ndarray:
    """RGBA8 texels (memory order R,G,B,A) of a 32-bit uncompressed BMP."""
    with open(path, "rb") as f:
        b = f.read()
    if b[:2] != b"BM":
        raise AssetError("not a BMP")
    off = int.from_bytes(b[10:14], "little")
    w = int.from_bytes(b[18:22], "little", signed=True)
    h = int.from_bytes(b[22:26], "little", signed=True)
    bpp = int.from_bytes(b[28:30], "little")
    if bpp != 32:
        raise AssetError(f"{bpp}-bit BMP")
    px = np.frombuffer(b, np.uint8, count=abs(w * h) * 4, offset=off).reshape(abs(h), w, 4)
    if h > 0:
        px = px[::-1]
    rgba = px[..., [2, 1, 0, 3]].astype(np.uint32)
    return rgba[..., 0] | rgba[..., 1] << 8 | rgba[..., 2] << 16 | rgba[..., 3] << 24


def _texels(call, what) -> np.ndarray:
    t = C.c_void_p()
    w, h = C.c_uint32(0), C.c_uint32(0)
    if call(C.byref(t), C.byref(w), C.byref(h)) != 0:
        raise _error(what)
    try:
        return np.ctypeslib.as_array(C.cast(t, C.POINTER(C.c_uint32)), (h.value, w.value)).copy()
    finally:
        lib().rio_free(t)


def decode_png(data: bytes) -> np.ndarray:
    """RGBA8 texels (memory order R,G,B,A; rows top to bottom), shape (h, w)."""
    return _texels(lambda *a: lib().rio_png_decode(data, len(data), *a), "rio_png_decode")


def load_png(path: str) -> np.ndarray:
    return _texels(lambda *a: lib().rio_png_load(path.encode(), *a), "rio_png_load")


def encode_png(rgba: np.ndarray, filter: int = -1) -> bytes:
    """An 8-bit RGBA PNG of RGBA8 texels; filter 0..4 for every row, -1 for
    row y filtered with type y % 5."""
    a = np.ascontiguousarray(rgba, np.uint32)
    p, n = C.c_void_p(), C.c_size_t(0)
    if lib().rio_png_encode(a.ctypes.data, a.shape[1], a.shape[0], filter, C.byref(p), C.byref(n)) != 0:
        raise _error("rio_png_encode")
    try:
        return C.string_at(p, n.value)
    finally:
        lib().rio_free(p)
