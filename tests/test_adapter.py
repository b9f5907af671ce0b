"""The reference-side adapter (photon_amd/reference_adapter/): the implementations of the reference's
GpuRenderer and Svgf over the C ABI compile against the reference's own unchanged headers
(src/render/gpu_renderer.h, src/denoise/svgf.h), and every entry point they call is one the device
library exports. Compile only: the reference is read where its checkout exists (this container), never
run, and the test is skipped elsewhere."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
ADAPTER = os.path.join(ROOT, "photon_amd", "reference_adapter")
HIP_LIB = os.path.join(ROOT, "photon_amd", "lib", "libphoton_hip.so")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="the reference checkout is not present")


@pytest.mark.parametrize("name", ["gpu_renderer_amd.cpp", "svgf_amd.cpp"])
def test_adapter_compiles_against_the_reference_headers(name, tmp_path):
    obj = tmp_path / (name + ".o")
    cmd = ["g++", "-std=c++20", "-c", "-Wall", "-Wextra", "-Werror", "-I", REF_SRC, "-I", os.path.join(ROOT, "include"),
           os.path.join(ADAPTER, name), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    undefined = subprocess.run(["nm", "-u", "--demangle=none", str(obj)], capture_output=True, text=True, check=True).stdout.split()
    calls = {s for s in undefined if s.startswith("ph_")}
    assert calls, "the adapter calls no device-library entry point"
    if os.path.exists(HIP_LIB):
        exported = set(subprocess.run(["nm", "-D", "--defined-only", HIP_LIB], capture_output=True, text=True, check=True).stdout.split())
        assert calls <= exported, sorted(calls - exported)
