"""CPU: the C-ABI library loads and exports every entry point include/octree_amd.h declares
(no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import pytest

import octree_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "octree_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(oct_[a-z_]+)\s*\(", src)))


def test_header_declares_the_binding_list():
    assert declared_functions() == sorted(A.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(A.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in A.lib().oct_version()


def test_synth_is_deterministic():
    a = A.lidar_tile(20000, seed=4, extent=30.0)
    b = A.lidar_tile(20000, seed=4, extent=30.0)
    assert a.tobytes() == b.tobytes()
    assert a[:, 0].min() >= 651000.0 and a[:, 0].max() < 651030.0
    u = A.uniform_box(1000, seed=1)
    assert u.min() >= 0.0 and u[:, 0].max() < 100.0 and u[:, 2].max() < 20.0
    # millimetre grid
    assert abs(((u * 1000.0) - np.round(u * 1000.0))).max() < 1e-6


import numpy as np  # noqa: E402


@pytest.mark.parametrize("src_name", ["facade_check.cpp", "facade_host_check.cpp"])
def test_facade_headers_compile(src_name):
    """The drop-in C++ facade (reference-named headers) compiles against the C ABI header under the
    reference's language standard (C++20, cmake/CMakeConfig.cmake:61), warnings as errors."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ absent")
    pkg = os.path.join(ROOT, "octrees-benchmark_amd")
    src = os.path.join(pkg, "facade", "examples", src_name)
    r = subprocess.run(["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I", os.path.join(pkg, "facade"),
                        "-I", os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
