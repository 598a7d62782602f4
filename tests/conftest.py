"""Shared test setup: import paths for the package (custom-gpgpu-rasterizer_amd/)
and the CPU oracle (oracle/), and the `gpu` marker for tests that need an MI355X."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "custom-gpgpu-rasterizer_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); calls librast.so through its C ABI")


@pytest.fixture(scope="session", autouse=True)
def _built_libraries():
    """Build the oracle and the scene generator if they are missing (cheap, CPU only)."""
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    import build as pkg_build

    if not os.path.exists(os.path.join(PKG, "lib", "librast_scenes.so")):
        pkg_build.build_scenes()
    yield
