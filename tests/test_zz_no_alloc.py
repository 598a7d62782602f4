"""GPU: rast_binmodel_forward_noalloc, the capacity model's forward pass with
its device staging in the library's static per-GPU array (no allocation in
the call), against rast_binmodel_forward on the same model and samples:
bitwise equal (the same kernel), over two calls in a row, the second with a
smaller input, and a clear error above the array's maximum."""
import numpy as np
import pytest

import rast as R

pytestmark = pytest.mark.gpu

F = 8
MAX_SAMPLES = 1 << 16


def forward_noalloc(model, x):
    fn = R.lib().rast_binmodel_forward_noalloc
    x = np.ascontiguousarray(x, np.float32)
    out = np.full(x.shape[0], np.nan, np.float32)
    R.check(fn(model.h, x.ctypes.data, x.shape[0], out.ctypes.data), "rast_binmodel_forward_noalloc")
    return out


@pytest.mark.parametrize("hidden", [16, 64])
def test_forward_noalloc_matches_forward(hidden):
    with R.Renderer(64, 64) as r:
        m = R.BinModel(r, hidden=hidden, seed=3)
        rng = np.random.default_rng(40 + hidden)
        m.set_params(rng.normal(0, 0.3, m.num_params).astype(np.float32))
        try:
            for n in (MAX_SAMPLES, 1237):  # the second call smaller: stale tail must not matter
                x = rng.normal(0, 1, (n, F)).astype(np.float32)
                want = m.forward(x)
                got = forward_noalloc(m, x)
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), n
            with pytest.raises(R.RastError, match="above the maximum"):
                forward_noalloc(m, np.zeros((MAX_SAMPLES + 1, F), np.float32))
        finally:
            m.close()
