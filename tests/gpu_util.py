"""Helpers shared by the GPU tests: scene/ray fixtures and the parity metrics of the device path."""
import numpy as np

from photon_amd import gpu, host
from photon_amd.types import HIT_DTYPE, RAY_DTYPE


def radiance_agreement(a, b, rel=1e-3, abs_floor=1e-6):
    """Fraction of pixels whose three channels all agree within `rel` relative (or `abs_floor` absolute)."""
    d = np.abs(a - b)
    ok = (d <= rel * np.abs(b)) | (d <= abs_floor)
    return float(np.mean(np.all(ok, axis=-1)))


def bitwise_fraction(a, b):
    return float(np.mean(np.all(a.view(np.uint32) == b.view(np.uint32), axis=-1)))


def trace_on_gpu(renderer, rays, anyhit=False, counts=False):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    n = len(rays)
    rd = gpu.DeviceArray.from_numpy(rays)
    hd = gpu.DeviceArray(n * HIT_DTYPE.itemsize)
    cd = gpu.DeviceArray(n * 12) if counts else None
    renderer.trace(rd.ptr, n, hd.ptr, anyhit=anyhit, counts_dev=cd.ptr if cd else None)
    hits = hd.numpy(HIT_DTYPE)
    if counts:
        return hits, cd.numpy(np.uint32, (n, 3))
    return hits


def uploaded(scene, width=0, height=0, engine=gpu.ENGINE_MEGAKERNEL):
    r = gpu.GpuRenderer()
    r.set_engine(engine)
    ok, err = r.upload(scene)
    assert ok, err
    if width and height:
        r.resize(width, height)
    return r


def sphere_rays(scene, width, height, shadow, seed=5):
    cam = scene.camera_view(width / height)
    return host.camera_rays(cam, width, height), scene.shadow_rays(shadow, seed=seed)
