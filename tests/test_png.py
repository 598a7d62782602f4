"""PNG ingest (rio_png_decode, the IMG_Load step of CUDATexture::Create,
CCR/CUDATexture.cpp:24-45), checked without a GPU against:
  - the committed uv_grid.png of the reference's Resources, whose every pixel
    the 32-pixel black/white checker rule fixes (known cells);
  - a second decoder written here in numpy on Python's zlib, on files whose
    rows use each of the five filters, and on the reference's vehicle maps
    (filters 1-4) when tests/golden/assets holds them;
  - encode -> decode round trips for every filter type;
  - hand-built files of the other colour types (grey, RGB with a tRNS key,
    palette with tRNS, grey + alpha) with the texels the specification gives;
  - rejection of a bad signature, a bad CRC, an interlaced or 16-bit file.
"""
import os
import struct
import zlib

import numpy as np
import pytest

import assets as A

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIG = b"\x89PNG\r\n\x1a\n"


def chunk(kind: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body))


def png(width, height, ctype, rows, extra=b"", interlace=0, depth=8):
    """A PNG of already-filtered rows (each starting with its filter byte)."""
    ihdr = struct.pack(">IIBBBBB", width, height, depth, ctype, 0, 0, interlace)
    return SIG + chunk(b"IHDR", ihdr) + extra + chunk(b"IDAT", zlib.compress(b"".join(rows))) + chunk(b"IEND", b"")


def reference_decode(data: bytes, max_rows=None) -> np.ndarray:
    """Second decoder (8-bit RGBA only, per the specification): chunks,
    zlib, and the five filters undone byte by byte with numpy rows."""
    assert data[:8] == SIG
    at, z, w, h = 8, b"", 0, 0
    while at < len(data):
        n = struct.unpack(">I", data[at:at + 4])[0]
        kind = data[at + 4:at + 8]
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", data[at + 8:at + 18])
            assert depth == 8 and ctype == 6
        elif kind == b"IDAT":
            z += data[at + 8:at + 8 + n]
        at += 12 + n
    raw = np.frombuffer(zlib.decompress(z), np.uint8)
    stride = 4 * w
    rows = h if max_rows is None else min(h, max_rows)
    out = np.zeros((rows, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(rows):
        f = raw[y * (stride + 1)]
        cur = raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)].astype(np.int32)
        if f == 0:
            rec = cur
        elif f == 2:
            rec = (cur + prev) & 255
        else:
            rec = np.zeros(stride, np.int32)
            for x in range(0, stride, 4):
                a = rec[x - 4:x] if x else np.zeros(4, np.int32)
                b = prev[x:x + 4]
                c = prev[x - 4:x] if x else np.zeros(4, np.int32)
                if f == 1:
                    p = a
                elif f == 3:
                    p = (a + b) >> 1
                else:
                    pa, pb, pc = np.abs(b - c), np.abs(a - c), np.abs(a + b - 2 * c)
                    p = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
                rec[x:x + 4] = (cur[x:x + 4] + p) & 255
        out[y] = rec
        prev = rec
    px = out.reshape(rows, w, 4).astype(np.uint32)
    return px[..., 0] | px[..., 1] << 8 | px[..., 2] << 16 | px[..., 3] << 24


def random_texels(h, w, seed):
    rng = np.random.default_rng(seed)
    # smooth-ish content as well as noise, so every predictor gets exercised
    base = rng.integers(0, 256, (h, w, 4))
    base[: h // 2] = (np.arange(w)[None, :, None] * 3 + np.arange(h // 2)[:, None, None] * 5) % 256
    t = base.astype(np.uint32)
    return t[..., 0] | t[..., 1] << 8 | t[..., 2] << 16 | t[..., 3] << 24


def test_uv_grid_fixture_is_the_32_pixel_checker():
    with open(os.path.join(GOLDEN, "uv_grid.png"), "rb") as f:
        data = f.read()
    t = A.decode_png(data)
    assert t.shape == (512, 512)
    y, x = np.mgrid[0:512, 0:512]
    want = np.where((y // 32 + x // 32) % 2 == 0, 0xFFFFFFFF, 0xFF000000).astype(np.uint32)
    np.testing.assert_array_equal(t, want)
    np.testing.assert_array_equal(reference_decode(data, max_rows=40), want[:40])


@pytest.mark.parametrize("filt", [-1, 0, 1, 2, 3, 4])
def test_round_trip_and_second_decoder_for_every_filter(filt):
    t = random_texels(23, 37, 5 + filt)
    data = A.encode_png(t, filt)
    np.testing.assert_array_equal(A.decode_png(data), t)
    np.testing.assert_array_equal(reference_decode(data), t)
    # the encoder wrote the filter asked for on every row
    raw = zlib.decompress(data[data.index(b"IDAT") + 4:])
    types = {raw[y * (37 * 4 + 1)] for y in range(23)}
    assert types == ({0, 1, 2, 3, 4} if filt == -1 else {filt})


@pytest.mark.parametrize("name", ["vehicle_diffuse.png", "vehicle_gloss.png", "vehicle_normal.png",
                                  "vehicle_specular.png"])
def test_reference_maps_agree_with_the_second_decoder(name):
    path = os.path.join(GOLDEN, "assets", name)
    if not os.path.exists(path):
        pytest.skip("reference maps not copied (tests/golden/copy_reference_assets.py)")
    with open(path, "rb") as f:
        data = f.read()
    t = A.load_png(path)
    assert t.shape == (1024, 1024)
    rows = 48
    np.testing.assert_array_equal(t[:rows], reference_decode(data, max_rows=rows))


def test_other_colour_types_decode_as_the_specification_says():
    # grey, with a tRNS key
    g = png(3, 1, 0, [bytes([0, 10, 200, 10])], extra=chunk(b"tRNS", b"\x00\x0a"))
    np.testing.assert_array_equal(A.decode_png(g)[0], [0x000A0A0A, 0xFFC8C8C8, 0x000A0A0A])
    # RGB, with a tRNS key, row filtered with Sub
    rgb = png(2, 1, 2, [bytes([1, 1, 2, 3, 4, 4, 4])], extra=chunk(b"tRNS", b"\x00\x05\x00\x06\x00\x07"))
    np.testing.assert_array_equal(A.decode_png(rgb)[0], [0xFF030201, 0x00070605])
    # palette with partial tRNS, two rows, the second filtered with Up
    plte = chunk(b"PLTE", bytes([255, 0, 0, 0, 255, 0, 0, 0, 255]))
    trns = chunk(b"tRNS", bytes([128]))
    p = png(3, 2, 3, [bytes([0, 0, 1, 2]), bytes([2, 1, 1, 0])], extra=plte + trns)
    np.testing.assert_array_equal(A.decode_png(p), [[0x800000FF, 0xFF00FF00, 0xFFFF0000],
                                                    [0xFF00FF00, 0xFFFF0000, 0xFFFF0000]])
    # grey + alpha, Average then Paeth
    ga = png(2, 2, 4, [bytes([3, 100, 50, 20, 30]), bytes([4, 1, 2, 3, 4])])
    # row 0: (100, 50), (20 + 50, 30 + 25) = (70, 55); row 1 Paeth: a=0,b=row0 -> b
    np.testing.assert_array_equal(A.decode_png(ga), [[0x32646464, 0x37464646], [0x34656565, 0x3B494949]])


def test_malformed_files_are_rejected():
    good = A.encode_png(random_texels(4, 4, 1), 0)
    A.decode_png(good)
    with pytest.raises(A.AssetError, match="signature"):
        A.decode_png(b"\x89PNX" + good[4:])
    bad_crc = bytearray(good)
    bad_crc[good.index(b"IDAT") + 6] ^= 0x40
    with pytest.raises(A.AssetError, match="CRC"):
        A.decode_png(bytes(bad_crc))
    row = bytes([0] + [1, 2, 3, 4] * 2)
    with pytest.raises(A.AssetError, match="interlaced"):
        A.decode_png(png(2, 1, 6, [row], interlace=1))
    with pytest.raises(A.AssetError, match="bit depth"):
        A.decode_png(png(1, 1, 6, [bytes(9)], depth=16))
    with pytest.raises(A.AssetError, match="filter type"):
        A.decode_png(png(2, 1, 6, [bytes([7]) + row[1:]]))
    with pytest.raises(A.AssetError, match="expected"):
        A.decode_png(png(2, 2, 6, [row]))  # one row of two
    with pytest.raises(A.AssetError, match="truncated|past the end"):
        A.decode_png(good[:-20])
