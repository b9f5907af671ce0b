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


LAYOUTS = [("gpu", "GpuInfo", "ph_gpu_info"), ("gpu", "WavefrontConfig", "ph_wavefront_config"), ("gpu", "RenderStats", "ph_render_stats"),
           ("gpu", "ServerConfig", "ph_server_config"), ("gpu", "JobDesc", "ph_job_desc"), ("gpu", "JobStatus", "ph_job_status"),
           ("gpu", "SvgfInput", "ph_svgf_input"), ("gpu", "BvhBuildStats", "ph_bvh_build_stats"), ("gpu", "CompactGBufferDesc", "ph_compact_gbuffer"),
           ("types", "BvhNode", "ph_bvh_node"), ("types", "TriGeom", "ph_tri_geom"), ("types", "TextureInfo", "ph_texture_info"),
           ("types", "Material", "ph_material"), ("types", "EmissiveTri", "ph_emissive_tri"), ("types", "AliasEntry", "ph_alias_entry"),
           ("types", "PunctualLight", "ph_punctual_light"), ("types", "CameraView", "ph_camera_view"), ("types", "Ray", "ph_ray"),
           ("types", "Hit", "ph_hit"), ("types", "PathOptions", "ph_path_options"), ("types", "RenderSettings", "ph_render_settings"),
           ("types", "SvgfSettings", "ph_svgf_settings"), ("types", "SceneDesc", "ph_scene_desc"), ("types", "RawScene", "phs_raw_scene"),
           ("types", "BvhStats", "phs_bvh_stats")]


def test_ctypes_layouts_match_the_headers():
    """Every ctypes mirror of a C ABI struct has the header's size and, field by field, its offsets (a
    field added on one side only, or in another place, fails here)."""
    import importlib
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("gcc"):
        pytest.skip("no host compiler")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "photon_hip.h"', '#include "photon_host.h"', "int main(void) {"]
    expect = []
    for mod, cls, cname in LAYOUTS:
        S = getattr(importlib.import_module("photon_amd." + mod), cls)
        lines.append(f'    printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(S))
        for name, _ in S._fields_:
            lines.append(f'    printf("%zu\\n", offsetof({cname}, {name}));')
            expect.append(getattr(S, name).offset)
    lines.append("    return 0;\n}")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "t.c"), os.path.join(d, "t")
        open(src, "w").write("\n".join(lines))
        r = subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), src, "-o", exe], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert got == expect
