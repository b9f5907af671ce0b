"""The C ABI: every function declared in include/*.h is exported by the library that implements it, and
the device library loads on a machine without a GPU (no compute calls are made here)."""
import ctypes as C
import os
import re

import pytest

from photon_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ph[s]?_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("photon_host.h", _lib.HOST_LIB), ("photon_hip.h", _lib.HIP_LIB)])
def test_every_declared_symbol_is_exported(header, lib):
    names = declared(header)
    assert len(names) > 10
    L = _lib.load(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_device_library_loads_without_a_device():
    from photon_amd import gpu
    n = gpu.lib().ph_device_count()
    assert n >= 0


def test_wrapper_header_compiles():
    # The C++ wrapper that keeps the reference's GpuRenderer/Svgf API must compile against the ABI.
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("g++"):
        pytest.skip("no host compiler")
    src = '#include "photon/gpu_renderer.hpp"\nint main() { return photon::cuda_available() ? 0 : 1; }\n'
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.cpp")
        open(p, "w").write(src)
        r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), p], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_device_bvh_build_refuses_inputs_above_its_limit():
    # Checked on the host before any device work, so this runs without a GPU.
    import numpy as np
    from photon_amd import gpu
    limit = int(re.search(r"#define PH_BVH_BUILD_MAX_TRIS (\d+)", open(os.path.join(ROOT, "include", "photon_hip.h")).read()).group(1))
    assert limit == 262144
    with pytest.raises(RuntimeError, match="limited to PH_BVH_BUILD_MAX_TRIS"):
        gpu.bvh_build(np.zeros((3, 3), np.float32), np.zeros((limit + 1, 3), np.int32))
