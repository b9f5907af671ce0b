"""The display path: the tonemapped frame written straight into pinned host memory (ph_host_alloc as the
display target) must be byte-identical to the renderer's display buffer copied to the host
(download_display), for every display mode (gpu_renderer.cu:98-175, DisplayMode gpu_renderer.h:16)."""
import numpy as np
import pytest

from photon_amd import gpu, host
from photon_amd.types import (RenderSettings, DISPLAY_COLOR, DISPLAY_ALBEDO, DISPLAY_NORMAL, DISPLAY_DEPTH, DISPLAY_VARIANCE,
                              DISPLAY_HEATMAP)

from gpu_util import uploaded

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("engine", [gpu.ENGINE_MEGAKERNEL, gpu.ENGINE_WAVEFRONT])
def test_display_into_pinned_host_memory(engine):
    scene = host.Scene.builtin("cornell-c1").build()
    w, h = 67, 45  # not a multiple of any block shape
    rs = RenderSettings.make(w, h, max_depth=8, rr_depth=3, seed=1)
    r = uploaded(scene, w, h, engine)
    r.render(scene.camera_view(w / h), rs, 4)
    pinned = gpu.PinnedArray(w * h * 4)
    for mode in (DISPLAY_COLOR, DISPLAY_ALBEDO, DISPLAY_NORMAL, DISPLAY_DEPTH, DISPLAY_VARIANCE, DISPLAY_HEATMAP):
        for tonemap in (0, 1, 2):
            rs.tonemap = tonemap
            r.display(rs, mode)
            want = r.download_display()
            pinned.numpy(np.uint32)[:] = 0xDEADBEEF
            r.display(rs, mode, target=pinned.ptr)
            got = pinned.numpy(np.uint32, (h, w))
            assert np.array_equal(got, want.reshape(h, w)), (mode, tonemap)
    assert np.any(pinned.numpy(np.uint32) != 0)
    pinned.free()
    r.close()


def test_display_refuses_pageable_and_short_targets():
    """A display target kernels cannot write (a numpy array's pageable memory) or a device buffer smaller
    than the frame fails with a message before any launch. The pointer test itself is checked first, so
    the display call is only made once the library has been seen to refuse such memory."""
    scene = host.Scene.builtin("cornell-c1").build()
    w, h = 256, 128
    rs = RenderSettings.make(w, h, max_depth=4, rr_depth=3, seed=1)
    r = uploaded(scene, w, h, gpu.ENGINE_MEGAKERNEL)
    r.render(scene.camera_view(w / h), rs, 1)
    pageable = np.zeros(w * h, np.uint32)
    assert gpu.pointer_kind(pageable.ctypes.data) == gpu.POINTER_UNUSABLE
    with pytest.raises(RuntimeError, match="pinned host memory"):
        r.display(rs, DISPLAY_COLOR, target=pageable.ctypes.data)
    assert np.all(pageable == 0)
    short = gpu.DeviceArray(w * h * 2)  # half the frame, a whole number of pages
    assert gpu.pointer_kind(short.ptr) == gpu.POINTER_DEVICE
    with pytest.raises(RuntimeError, match="smaller than the frame"):
        r.display(rs, DISPLAY_COLOR, target=short.ptr)
    pinned = gpu.PinnedArray(w * h * 4)
    assert gpu.pointer_kind(pinned.ptr) == gpu.POINTER_PINNED_HOST
    r.display(rs, DISPLAY_COLOR, target=pinned.ptr)
    r.display(rs, DISPLAY_COLOR)  # into the renderer's own buffer, which download_display reads
    assert np.array_equal(pinned.numpy(np.uint32, (h, w)), r.download_display().reshape(h, w))
    short.free()
    pinned.free()
    r.close()
