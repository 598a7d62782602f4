"""GPU: the C++ classes of include/rast.hpp (CUDARenderer, CUDAMesh,
CUDATexture, CUDATextureManager over the C ABI), driven with app-side types
shaped like the reference's; the frame must equal the oracle's (the program
checks vis/depth bitwise and RGBA8 within 1/255 and exits non-zero otherwise),
and a texture loaded from a PNG file through CUDATexture's path form must
draw the fixture's checker."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "build", "test_rast_cpp")


def test_cpp_interface_renders_the_oracle_frame():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built; run __graft_entry__.build()")
    r = subprocess.run([BIN, ROOT], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
