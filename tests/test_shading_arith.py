"""Arithmetic identities the shading kernel relies on, checked on the CPU in
IEEE single precision.

* The unit table (c / 255 for a byte c) is built without a division: q = c x
  r, r = 1/255 rounded, then one residual correction fma(fma(-q, 255, c), r,
  q).  Checked for all 256 byte values against the correctly rounded
  quotient (the reference's `(float)c / 255.f`, CCR/CUDATextureSampler.cuh).
* SampleLinear's float accumulation (neighbours x 1/8 in order, then centre
  / 2, then x 1/255; CCR/CUDATextureSampler.cuh:112-190) equals the integer
  form (n0 + n1 + n2 + n3 + 4 c) x ((1/255) / 8) bit for bit, with red and
  blue summed side by side in 16-bit halves of one word.
"""
import ctypes

import numpy as np
import pytest

f32 = np.float32
REV = f32(1) / f32(255)


def _fmaf():
    libm = ctypes.CDLL("libm.so.6")
    fn = libm.fmaf
    fn.restype = ctypes.c_float
    fn.argtypes = [ctypes.c_float] * 3
    return lambda a, b, c: f32(fn(float(a), float(b), float(c)))


def test_unit_table_without_division_is_the_correctly_rounded_quotient():
    fma = _fmaf()
    # fused: the residual 1 - r x 255 is exact (a product of 8- and 24-bit
    # mantissas), which an unfused multiply-add would round to zero
    resid = fma(-REV, f32(255), f32(1))
    assert resid != 0 and float(resid) == 1.0 - float(REV) * 255.0
    plain_wrong = 0
    for c in range(256):
        q = f32(c) * REV
        got = fma(fma(-q, f32(255), f32(c)), REV, q)
        want = f32(c) / f32(255)
        assert got == want, c
        plain_wrong += q != want
    assert plain_wrong > 0  # the correction is needed


def linear_float(n, c):
    """The reference accumulation for one channel (values 0..255)."""
    w = f32(1) / f32(8)
    acc = f32(0)
    for t in n:
        acc = f32(acc + f32(t) * w)
    acc = f32(acc + f32(c) / f32(2))
    return f32(acc * REV)


def linear_packed(texels, centre):
    """The kernel's integer form: returns (r, g, b) for RGBA8 words."""
    rb = sum(int(t) & 0x00FF00FF for t in texels) + ((int(centre) & 0x00FF00FF) << 2)
    gs = sum((int(t) >> 8) & 0xFF for t in texels) + (((int(centre) >> 8) & 0xFF) << 2)
    k = REV * f32(0.125)
    return f32(f32(rb & 0xFFFF) * k), f32(f32(gs) * k), f32(f32(rb >> 16) * k)


@pytest.mark.parametrize("seed", range(3))
def test_linear_sampling_integer_form_is_bitwise_the_float_accumulation(seed):
    rng = np.random.default_rng(seed)
    words = rng.integers(0, 2**32, size=(4000, 5), dtype=np.uint64).astype(np.uint32)
    # the extremes: all zero, all 255, one channel saturated
    words[:3] = [[0] * 5, [0xFFFFFFFF] * 5, [0x00FF00FF] * 5]
    for row in words:
        n, c = row[:4], row[4]
        r, g, b = linear_packed(n, c)
        for sh, got in ((0, r), (8, g), (16, b)):
            want = linear_float([(int(t) >> sh) & 0xFF for t in n], (int(c) >> sh) & 0xFF)
            assert got.view(np.uint32) == want.view(np.uint32), (row, sh)


def test_linear_sampling_sums_cover_every_channel_total():
    # every reachable total S = n0 + n1 + n2 + n3 + 4c (0..2040) gives the same float
    k = REV * f32(0.125)
    for total in range(0, 2041):
        assert f32(f32(total) * k) == f32(f32(f32(total) * f32(0.125)) * REV), total
