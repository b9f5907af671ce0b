"""Whole-image pins of the oracle integrator (the reference's test_render.cpp): white furnaces converge
to exactly one, NEE and BSDF-only sampling agree in expectation, NEE cuts the noise, rendering is
deterministic and seed sensitive, and the Cornell box has its expected brightness and colours."""
import numpy as np

from oracle import oracle_py as O
from photon_amd import host
from photon_amd.types import RenderSettings


def lum(img):
    return 0.2126 * img[..., 0] + 0.7152 * img[..., 1] + 0.0722 * img[..., 2]


def render(scene, w, h, spp, seed=0, **opt):
    rs = RenderSettings.make(w, h, seed=seed, **opt)
    return O.render(scene.desc(), scene.camera_view(w / h), rs, spp, threads=8)


def test_white_furnace_converges_to_one():
    s = host.Scene.builtin("furnace")
    s.clear_normals()
    s.build()
    img = render(s, 48, 48, 16, max_depth=64, rr_depth=64, nee=0)
    L = lum(img)
    assert abs(L.mean() - 1) < 0.002 and np.abs(L - 1).max() < 1e-3
    assert abs(lum(render(s, 48, 48, 64)).mean() - 1) < 0.005
    smooth = host.Scene.builtin("furnace").build()
    assert abs(lum(render(smooth, 48, 48, 16, max_depth=64, rr_depth=64, nee=0)).mean() - 1) < 0.01


def test_white_furnace_with_glossy_metal():
    s = host.Scene.builtin("furnace")
    m = s.material(0)
    m.metallic, m.roughness = 1.0, 0.0
    s.set_material(0, m)
    s.build()
    assert abs(lum(render(s, 32, 32, 16, max_depth=64, rr_depth=64)).mean() - 1) < 0.01


def test_light_sampling_does_not_change_the_expectation():
    for name in ("cornell", "materials"):
        s = host.Scene.builtin(name).build()
        a = lum(render(s, 64, 36, 256, seed=1, max_depth=6)).mean()
        b = lum(render(s, 64, 36, 1024, seed=2, max_depth=6, nee=0)).mean()
        assert abs(a - b) < 0.02 * a, (name, a, b)


def test_next_event_estimation_reduces_noise():
    s = host.Scene.builtin("cornell").build()
    ref = lum(render(s, 32, 32, 512, seed=9))
    e_with = np.mean((lum(render(s, 32, 32, 16, seed=3)) - ref) ** 2)
    e_without = np.mean((lum(render(s, 32, 32, 16, seed=3, nee=0)) - ref) ** 2)
    assert e_with * 3 < e_without


def test_rendering_is_deterministic_and_seed_sensitive():
    s = host.Scene.builtin("glass").build()
    a = render(s, 40, 24, 4, seed=5)
    b = render(s, 40, 24, 4, seed=5)
    c = render(s, 40, 24, 4, seed=6)
    assert a.tobytes() == b.tobytes()
    assert (a[..., 0] != c[..., 0]).any()


def test_cornell_box_brightness_and_colours():
    s = host.Scene.builtin("cornell").build()
    img = render(s, 64, 64, 256, seed=4)
    wall = lum(img[18:24, 28:36]).mean()
    assert 0.05 < wall < 0.5
    left, right = img[32, 4], img[32, 59]
    assert left[0] > 2 * left[1] and right[1] > 1.5 * right[0]


def test_single_samples_match_the_accumulated_render():
    # or_render_samples (one path per call, the unit of work of every engine) sums to or_render.
    s = host.Scene.builtin("cornell-c1").build()
    w = h = 16
    rs = RenderSettings.make(w, h, seed=1)
    cam = s.camera_view(1.0)
    total = O.render_sum(s.desc(), cam, rs, 3)
    import ctypes as C
    ys, xs = np.meshgrid(np.arange(h, dtype=np.int32), np.arange(w, dtype=np.int32), indexing="ij")
    acc = np.zeros((h * w, 3), np.float32)
    for k in range(3):
        rgb = np.zeros((h * w, 3), np.float32)
        frames = np.full(h * w, k, np.uint32)
        xr, yr = np.ascontiguousarray(xs.reshape(-1)), np.ascontiguousarray(ys.reshape(-1))
        O.oracle().or_render_samples(C.byref(s.desc()), C.byref(cam), C.byref(rs), h * w, xr.ctypes.data, yr.ctypes.data,
                                     frames.ctypes.data, rgb.ctypes.data_as(C.POINTER(C.c_float)), None)
        acc = acc + rgb
    assert acc.reshape(h, w, 3).tobytes() == total.tobytes()
