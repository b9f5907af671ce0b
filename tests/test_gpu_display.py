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
