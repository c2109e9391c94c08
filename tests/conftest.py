import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def _need(path, build_cmd):
    if not os.path.exists(path):
        subprocess.run(build_cmd, check=True, cwd=ROOT)
    return path


@pytest.fixture(scope="session")
def oracle_lib():
    _need(os.path.join(ROOT, "oracle", "liboracle.so"), ["make", "-C", "oracle", "oracle"])
    import oracle  # noqa: F401
    return oracle


@pytest.fixture(scope="session")
def hostmath():
    import ctypes
    path = os.path.join(ROOT, "tests", "libhostmath.so")
    src = os.path.join(ROOT, "tests", "host_math_check.cpp")
    deps = [src] + [os.path.join(ROOT, "ibemproject_amd", "csrc", f)
                    for f in ("ibem3d_math.h", "kelvin_tables.inc")]
    if not os.path.exists(path) or os.path.getmtime(path) < max(os.path.getmtime(d) for d in deps):
        subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-o", path, src], check=True)
    return ctypes.CDLL(path)
