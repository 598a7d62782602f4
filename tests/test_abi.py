"""The C ABI library loads without a GPU, exports every entry point that
include/rast.h declares, computes the host camera math bitwise like the
oracle, and reports errors through status codes (no GPU calls here)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
import rast as R
import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|double|float|void)\s+\**(r(?:ast|sc|io)_\w+)\s*\(", text, re.M)))


def exported(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_librast_exports_every_declared_entry_point():
    names = declared("rast.h")
    assert len(names) >= 35
    missing = sorted(set(names) - exported(R.LIB_PATH))
    assert not missing, missing
    # and the Python binding knows all of them
    assert set(names) <= set(R.EXPORTS) | {"rast_set_tuning"}


def test_scene_library_exports_every_declared_entry_point():
    names = declared("rast_scene.h")
    assert not sorted(set(names) - exported(S.SCENES_LIB))


def test_io_library_exports_every_declared_entry_point():
    import assets as A

    names = declared("rast_io.h")
    assert len(names) >= 5
    assert not sorted(set(names) - exported(A.IO_LIB))


def test_library_loads_and_binds_without_a_gpu():
    L = R.lib()
    for name in R.EXPORTS:
        assert getattr(L, name) is not None


@pytest.mark.parametrize("cam", [
    S.Camera(),
    S.Camera(position=(1.5, -2.0, 7.0), forward=(0.3, 0.4, 0.86), fov_deg=60.0, near_z=0.5, far_z=50.0),
    S.terrain_camera(),
])
@pytest.mark.parametrize("size", [(640, 480), (3840, 2160), (1024, 1024)])
def test_host_camera_math_is_bitwise_the_oracle(cam, size):
    got = R.camera_frame(cam, *size)
    want = O.camera_frame(cam, *size)
    for field, _ in R.FrameConsts._fields_:
        g = np.ctypeslib.as_array(getattr(got, field)).view(np.uint32)
        w = np.ctypeslib.as_array(getattr(want, field)).view(np.uint32)
        np.testing.assert_array_equal(g, w, err_msg=field)


def test_errors_are_status_codes_with_messages():
    L = R.lib()
    h = C.c_void_p()
    assert L.rast_renderer_create(0, 0, 480, C.byref(h)) == -1
    assert "out of range" in R.last_error()
    assert L.rast_render(None, None, None) == -1
    assert L.rast_camera_frame(None, 1, 1, None) == -1
    assert L.rast_mesh_destroy(None) == -1
    assert L.rast_binmodel_num_features() == 8
    buf = C.create_string_buffer(4)
    n = L.rast_last_error(buf, 4)
    assert n > 3 and len(buf.value) == 3  # truncated and terminated


def test_renderer_create_without_a_device_fails_cleanly():
    h = C.c_void_p()
    rc = R.lib().rast_renderer_create(0, 64, 64, C.byref(h))
    if rc == 0:  # a GPU is present: nothing to check here
        R.lib().rast_renderer_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc in (-5, -2) and not h.value
    assert R.last_error()


def test_cpp_header_compiles_against_reference_shaped_types():
    """include/rast.hpp: the reference class names compile as C++17 with
    app-side types shaped like the reference's (syntax and template check)."""
    import shutil
    import subprocess

    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = os.path.join(ROOT, "tests", "cpp", "test_rast_cpp.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "oracle"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
