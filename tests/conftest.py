import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "edipic-2d_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libedgpu.so on the device)")


def gpu_available():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = ctypes.CDLL(name)
        except OSError:
            continue
        n = ctypes.c_int(0)
        try:
            rc = h.hipGetDeviceCount(ctypes.byref(n))
        except Exception:
            return False
        return rc == 0 and n.value > 0
    return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible to HIP")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def oracle():
    import oracle_ffi
    if not os.path.exists(oracle_ffi.LIB_PATH):
        oracle_ffi.build()
    return oracle_ffi


@pytest.fixture(scope="session")
def edgpu():
    import edgpu
    edgpu.lib()  # raises when the library is missing: GPU tests must run the HIP path
    return edgpu
