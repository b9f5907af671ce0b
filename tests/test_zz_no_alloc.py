"""ph_svgf_run_static (the denoiser without an allocation: frame-local planes in the library's static
device array, history in caller memory) against ph_svgf_run on the same G-buffers, called through the C
ABI: every frame of a sequence must be bit-identical, over two sequences in a row, the second on a
smaller frame (so stale contents of the shared array from the larger call cannot leak into it)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd import gpu, host
from photon_amd.gpu import SvgfInput
from photon_amd.types import CameraView, RenderSettings, SvgfSettings

pytestmark = pytest.mark.gpu

MAX_PIXELS = 1600 * 900


def _lib():
    L = gpu.lib()
    L.ph_svgf_run_static.argtypes = [C.POINTER(SvgfInput), C.c_int, C.c_int, C.POINTER(CameraView), C.POINTER(SvgfSettings), C.c_void_p,
                                     C.c_int, C.c_void_p]
    return L


def _frames(scene, w, h, n):
    """G-buffers of `n` frames of a camera orbiting 0.5 degrees a frame (traced by the oracle)."""
    rs = RenderSettings.make(w, h, seed=3)
    cam0 = scene.camera()
    out = []
    for f in range(n):
        a = np.deg2rad(0.5 * f)
        p = np.array(cam0["position"]) - np.array(cam0["target"])
        p = np.array([np.cos(a) * p[0] + np.sin(a) * p[2], p[1], -np.sin(a) * p[0] + np.cos(a) * p[2]]) + np.array(cam0["target"])
        scene.set_camera(p, cam0["target"], cam0["up"], cam0["vfov"])
        out.append((scene.camera_view(w / h), O.render_gbuffer(scene.desc(), scene.camera_view(w / h), rs, f)))
    scene.set_camera(cam0["position"], cam0["target"], cam0["up"], cam0["vfov"])
    return out


def _sequence(L, frames, w, h):
    sv = SvgfSettings.make()
    f = gpu.Svgf()
    f.resize(w, h)
    out_obj = gpu.DeviceArray(w * h * 16)
    out_static = gpu.DeviceArray(w * h * 16)
    history = gpu.DeviceArray(4 * w * h * 16)
    prev = frames[0][0]
    for k, (cam, g) in enumerate(frames):
        dev = [gpu.DeviceArray.from_numpy(p) for p in g]
        f.run(*[d.ptr for d in dev], prev, sv, out_obj.ptr)
        inp = SvgfInput(*[d.ptr for d in dev])
        rc = L.ph_svgf_run_static(C.byref(inp), w, h, C.byref(prev), C.byref(sv), history.ptr, int(k > 0), out_static.ptr)
        assert rc == 0, L.ph_last_error().decode()
        a, b = out_obj.numpy(np.float32), out_static.numpy(np.float32)
        print(f"{w}x{h} frame {k}: bit-identical {a.tobytes() == b.tobytes()}")
        assert a.tobytes() == b.tobytes(), k
        prev = cam
    del f


def test_static_scratch_denoiser_matches_the_allocating_one():
    L = _lib()
    scene = host.Scene.builtin("cornell-c1").build()
    _sequence(L, _frames(scene, 160, 120, 3), 160, 120)
    _sequence(L, _frames(scene, 72, 40, 3), 72, 40)  # smaller: the shared array holds the larger call's planes


def test_static_scratch_refuses_frames_beyond_its_size():
    L = _lib()
    dummy = gpu.DeviceArray(64)
    inp = SvgfInput(*([dummy.ptr] * 5))
    cam, sv = CameraView(), SvgfSettings.make()
    rc = L.ph_svgf_run_static(C.byref(inp), 1601, 900, C.byref(cam), C.byref(sv), dummy.ptr, 0, dummy.ptr)
    assert rc != 0 and "exceeds the static SVGF scratch" in L.ph_last_error().decode()
    assert 1601 * 900 > MAX_PIXELS
