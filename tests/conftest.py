import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libmagreete_hip.so")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def host_kernels(tmp_path_factory):
    """The kernels' Bessel and Green's-function code (bessel.h, green.h) built
    for the host with g++, exposed through tests/host_shim.cpp."""
    import ctypes
    import subprocess
    out = tmp_path_factory.mktemp("host") / "libmg_host.so"
    src = os.path.join(ROOT, "tests", "host_shim.cpp")
    inc = os.path.join(ROOT, "magreete_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", inc, src, "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    return lib
