"""TEST INFRASTRUCTURE ONLY: a pure-Python restatement of the reference's OBJ
reader, the checker of librast_io's rio_obj_parse.  Nothing in the product
imports it.

Follows kdedeurw/Custom-GPGPU-Rasterizer CustomCUDARasterizer/ObjParser.cpp:
  ParseData             :23-87    line classification by the prefix before the first space
  StorePosition         :111-116  "v" + two spaces: the first three characters skipped
  StoreFace             :118-164  three vertices, '/'-separated 1-based indices, component = k % (count / 3)
  StoreNormal           :166-171  GetNormalized
  StoreUV               :173-179  v -> 1 - v when the y axis is inverted
  GetFirstSecondThird   :181-192  std::getline(.., ' ') three times
  AssignVertices        :381-553  de-duplication; tangents :446-473
and CCR/MathUtilities.h:30-38 (AreEqual, 2 ulp), CCR/Vector.h:41-74
(Normalize, Reject).  Arithmetic in numpy float32 scalars, one rounding per
operation, in the reference's order.  Where the reference reads memory it
never wrote, this restatement uses tangent 0, colour (1, 1, 1), normal 0.
"""
import ctypes

import numpy as np

F = np.float32
EPS = F(np.finfo(np.float32).eps)
TINY = F(np.finfo(np.float32).tiny)


class ObjError(ValueError):
    pass


_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]


def _stof(s):
    # std::stof is strtof on the characters (leading whitespace, the longest
    # number prefix, rounded once to binary32); no number -> the reference throws
    b = s.encode("latin-1")
    end = ctypes.c_char_p()
    buf = ctypes.create_string_buffer(b)
    v = _libc.strtof(buf, ctypes.byref(end))
    if ctypes.cast(end, ctypes.c_void_p).value == ctypes.cast(buf, ctypes.c_void_p).value:
        raise ObjError(f"no number in {s!r}")
    return F(v)


def _stoi(s):
    t = s.lstrip(" \t\r\n\f\v")
    end = 0
    if end < len(t) and t[end] in "+-":
        end += 1
    while end < len(t) and t[end].isdigit():
        end += 1
    if end == 0 or not t[:end].lstrip("+-"):
        raise ObjError(f"no integer in {s!r}")
    return int(t[:end])


def _three(s):
    out = s.split(" ")
    return (out + ["", "", ""])[:3]


def _normalized(v):
    m = F(np.sqrt(F(F(F(v[0] * v[0]) + F(v[1] * v[1])) + F(v[2] * v[2]))))
    if m == F(0):
        return [F(0), F(0), F(0)]
    inv = F(F(1) / m)
    return [F(v[0] * inv), F(v[1] * inv), F(v[2] * inv)]


def _dot(a, b):
    return F(F(F(a[0] * b[0]) + F(a[1] * b[1])) + F(a[2] * b[2]))


def _are_equal(a, b):
    d = F(abs(F(a - b)))
    return bool(d <= F(F(EPS * F(abs(F(a + b)))) * F(2)) or d < TINY)


def parse(text: bytes, invert_y=False):
    """Returns (verts float32 [n, 14] as p uv n tan c, indices uint32, vertex_type)."""
    pos, uvs, nrm = [], [], []
    pi, ti, ni = [], [], []
    for raw in text.decode("latin-1").split("\n"):
        line = raw
        if not line or line[0] == "#":
            continue
        prefix = line.split(" ", 1)[0]
        if prefix == "v":
            a, b, c = _three(line[3:])
            pos.append([_stof(a), _stof(b), _stof(c)])
        elif prefix == "f":
            toks = _three(line[2:])
            idx = []
            for tok in toks:
                if not tok:
                    raise ObjError("face with fewer than three vertices")
                for field in tok.split("/"):
                    v = _stoi(field)
                    if v < 1:
                        raise ObjError("relative or zero index")
                    idx.append(v - 1)
            k = len(idx) // 3
            for q, v in enumerate(idx):
                c = q % k
                (pi, ti, ni)[c].append(v)
        elif prefix == "vt":
            a, b, _ = _three(line[3:])
            y = _stof(b)
            if invert_y:
                y = F(F(1) - y)
            uvs.append([_stof(a), y])
        elif prefix == "vn":
            a, b, c = _three(line[3:])
            nrm.append(_normalized([_stof(a), _stof(b), _stof(c)]))
    has_uv, has_n = bool(uvs), bool(nrm)
    vtype = (1 if has_uv else 0) | (4 if has_n else 0)
    if has_n and not has_uv:
        raise ObjError("normals without uvs")
    for lst, name, have in ((ti, "uv", has_uv), (ni, "normal", has_n)):
        if have and len(lst) != len(pi):
            raise ObjError(f"faces without {name} indices")
    for i in range(len(pi)):
        if pi[i] >= len(pos) or (has_uv and ti[i] >= len(uvs)) or (has_n and ni[i] >= len(nrm)):
            raise ObjError("index out of range")
    # the reference scans every earlier unique vertex (its "black list") in
    # order; only entries with the same position index (and uv index, when the
    # file has uvs) can match, so the scan is kept per index pair
    verts, index, black = [], [], {}
    for i in range(len(pi)):
        key = (pi[i], ti[i] if has_uv else 0)
        match = None
        for (bn, vi) in black.get(key, []):
            if has_uv and has_n:
                if all(_are_equal(x, y) for x, y in zip(nrm[bn], nrm[ni[i]])):
                    match = vi  # no break: the last one wins
            else:
                match = vi
                break
        if match is None:
            match = len(verts)
            uv = uvs[ti[i]] if has_uv else [F(0), F(0)]
            n = nrm[ni[i]] if has_n else [F(0), F(0), F(0)]
            verts.append(list(pos[pi[i]]) + list(uv) + list(n) + [F(0)] * 3 + [F(1)] * 3)
            black.setdefault(key, []).append((ni[i] if has_n else 0, match))
        index.append(match)
    if has_uv and has_n:
        vtype |= 8
        for i in range(0, len(index) - 2, 3):
            a, b, c = verts[index[i]], verts[index[i + 1]], verts[index[i + 2]]
            e0 = [F(b[k] - a[k]) for k in range(3)]
            e1 = [F(c[k] - a[k]) for k in range(3)]
            dx = [F(b[3] - a[3]), F(c[3] - a[3])]
            dy = [F(b[4] - a[4]), F(c[4] - a[4])]
            with np.errstate(divide="ignore", invalid="ignore"):
                r = F(F(1) / F(F(dx[0] * dy[1]) - F(dx[1] * dy[0])))
                t = [F(F(F(e0[k] * dy[1]) - F(e1[k] * dy[0])) * r) for k in range(3)]
            for v in (a, b, c):
                for k in range(3):
                    v[8 + k] = F(v[8 + k] + t[k])
        for v in verts:
            tan, n = v[8:11], v[5:8]
            with np.errstate(divide="ignore", invalid="ignore"):
                s = F(_dot(tan, n) / _dot(n, n))
                rej = [F(tan[k] - F(n[k] * s)) for k in range(3)]
                v[8:11] = _normalized(rej)
    return np.array(verts, np.float32).reshape(-1, 14), np.array(index, np.uint32), vtype
