"""Arithmetic identity the raster kernels rely on, checked exhaustively on
the CPU in IEEE single precision.

Item decoding (setup's micro-triangle items and resolve's queued-triangle
items): item q of a pixel range of row width w sits in row
trunc(fl(fl(q + 0.5) x fl(1 / w))), with no correction step, for every
q < 1024 and 1 <= w <= 1024 (both item counts are bounded by a tuning limit
of at most 1024).  A planted variant without the half offset fails: q = w
then lands in row 0 for some w.
"""
import numpy as np

f32 = np.float32


def _rows(offset: float):
    w = np.arange(1, 1025, dtype=np.int64)
    q = np.arange(0, 1024, dtype=np.int64)
    inv = (f32(1) / w.astype(f32)).astype(f32)                      # fl(1 / w), correctly rounded
    qq = (q.astype(f32) + f32(offset)).astype(f32)                  # exact
    prod = (qq[:, None] * inv[None, :]).astype(f32)                 # one rounding
    return prod.astype(np.int64), q[:, None] // w[None, :]


def test_item_row_is_exact_for_every_item_and_width():
    got, want = _rows(0.5)
    np.testing.assert_array_equal(got, want)


def test_item_row_without_the_half_offset_is_wrong_somewhere():
    got, want = _rows(0.0)
    assert (got != want).any()
