"""The shading inputs the built-in scenes leave out, on the device against the oracle:

* the showcase scene (tests/golden/make_golden.py): normal maps with tangents, base-colour,
  metallic-roughness and emissive textures (integrator.h:102-125, lights.h:110-118), point, spot and
  directional lights (lights.h:164-190) and a thin-lens camera (integrator.h:265-280). The oracle
  reproduces the reference CpuRenderer's image of this scene bit for bit (render_showcase.npz);
* every display mode and tone map (gpu_renderer.cu:108-175), on the device's own planes, progressive and
  denoised;
* update_lights (gpu_renderer.cu:299-334) after changing emission and the environment: the device must
  render what a fresh upload of the changed scene renders, and what the oracle renders.

Bars: radiance within 0.1 % relative on at least 99.9 % of pixels; packed display pixels equal on at
least 99.9 % and never more than one code value apart (the sRGB curve's powf may round differently in
the last bit on the two sides).
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd import gpu
from photon_amd.types import (Material, RenderSettings, SvgfSettings, DISPLAY_COLOR, DISPLAY_ALBEDO, DISPLAY_NORMAL, DISPLAY_DEPTH,
                              DISPLAY_VARIANCE, DISPLAY_HEATMAP)

from golden_scenes import make_scene
from gpu_util import bitwise_fraction, radiance_agreement, uploaded

pytestmark = pytest.mark.gpu

ENGINES = [gpu.ENGINE_MEGAKERNEL, gpu.ENGINE_WAVEFRONT]
MODES = [DISPLAY_COLOR, DISPLAY_ALBEDO, DISPLAY_NORMAL, DISPLAY_DEPTH, DISPLAY_VARIANCE, DISPLAY_HEATMAP]


@pytest.fixture(scope="module")
def showcase():
    return make_scene("showcase").build()


def _extent(desc):
    """The renderer's depth-view scale: the length of the scene's bounding-box diagonal (float32)."""
    d = [np.float32(desc.bounds_max[k]) - np.float32(desc.bounds_min[k]) for k in range(3)]
    return max(np.float32(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1]) + np.float32(d[2] * d[2]))), np.float32(1e-3))


@pytest.mark.parametrize("engine", ENGINES)
def test_showcase_matches_oracle(showcase, engine):
    w, h, spp = 96, 60, 16
    rs = RenderSettings.make(w, h, max_depth=8, rr_depth=3, seed=4)
    cam = showcase.camera_view(w / h)
    assert cam.aperture > 0
    r = uploaded(showcase, w, h, engine)
    r.render(cam, rs, spp)
    img = r.download_image()
    ref = O.render(showcase.desc(), cam, rs, spp)
    agree, bits = radiance_agreement(img, ref), bitwise_fraction(img, ref)
    print(f"showcase engine={engine}: {100 * agree:.3f}% within 0.1%, {100 * bits:.3f}% bitwise, means {img.mean():.5f}/{ref.mean():.5f}")
    assert agree >= 0.999
    r.close()


def test_showcase_gbuffer_matches_oracle(showcase):
    w, h = 96, 60
    rs = RenderSettings.make(w, h, seed=6)
    cam = showcase.camera_view(w / h)
    r = uploaded(showcase, w, h)
    r.render_denoised(cam, cam, rs, SvgfSettings.make())
    planes = [gpu.copy_device_to_numpy(r.buffer_ptr(b), w * h * 16, np.float32, (h, w, 4))
              for b in (gpu.BUF_COLOR, gpu.BUF_ALBEDO, gpu.BUF_NORMAL_DEPTH, gpu.BUF_POSITION, gpu.BUF_EMISSION)]
    ref = O.render_gbuffer(showcase.desc(), cam, rs, 0)
    for name, a, b in zip(("color", "albedo", "nd", "position", "emission"), planes, ref):
        frac = radiance_agreement(a, b)
        print(f"{name}: {100 * frac:.3f}% within 0.1%, {100 * bitwise_fraction(a, b):.3f}% bitwise")
        assert frac >= 0.999, name
    r.close()


def _planes(r, w, h):
    n = w * h
    get = lambda b, k=4: gpu.copy_device_to_numpy(r.buffer_ptr(b), n * 4 * k, np.float32)  # noqa: E731
    return dict(accum=get(gpu.BUF_ACCUM), denoised_rgb=get(gpu.BUF_DENOISED), albedo=get(gpu.BUF_ALBEDO), nd=get(gpu.BUF_NORMAL_DEPTH),
                heat=get(gpu.BUF_HEAT, 1))


def _compare_display(r, rs, w, h, samples, denoised, extent, label):
    for mode in MODES:
        for tonemap in (0, 1, 2) if mode == DISPLAY_COLOR else (0,):
            rs.tonemap = tonemap
            r.display(rs, mode)
            got = r.download_display().reshape(-1)
            p = _planes(r, w, h)
            if denoised:
                p["variance"] = gpu.copy_device_to_numpy(r.buffer_ptr(gpu.BUF_VARIANCE), w * h * 16, np.float32)
            want = O.display(mode, w * h, exposure=rs.exposure, tonemap=tonemap, samples=samples, max_depth=extent, denoised=denoised, **p)
            gc = np.stack([(got >> s) & 255 for s in (0, 8, 16)], -1).astype(np.int64)
            wc = np.stack([(want >> s) & 255 for s in (0, 8, 16)], -1).astype(np.int64)
            equal = np.mean(got == want)
            print(f"{label} mode={mode} tonemap={tonemap}: {100 * equal:.3f}% of pixels equal, max code difference {np.abs(gc - wc).max()}")
            assert equal >= 0.999 and np.abs(gc - wc).max() <= 1, (mode, tonemap)
            assert np.any(want != 0xFF000000), (mode, tonemap)  # the view shows something


@pytest.mark.parametrize("engine", ENGINES)
def test_every_display_mode_matches_oracle(showcase, engine):
    w, h, spp = 80, 50, 6
    rs = RenderSettings.make(w, h, max_depth=8, rr_depth=3, seed=2)
    rs.exposure = 1.7
    cam = showcase.camera_view(w / h)
    r = uploaded(showcase, w, h, engine)
    r.render(cam, rs, spp)
    _compare_display(r, rs, w, h, spp, False, _extent(showcase.desc()), f"progressive engine={engine}")
    r.close()


def test_every_display_mode_of_a_denoised_frame_matches_oracle(showcase):
    w, h = 80, 50
    rs = RenderSettings.make(w, h, seed=9)
    cam = showcase.camera_view(w / h)
    r = uploaded(showcase, w, h)
    for _ in range(3):
        r.render_denoised(cam, cam, rs, SvgfSettings.make())
    assert r.samples == 1
    _compare_display(r, rs, w, h, 1, True, _extent(showcase.desc()), "denoised")
    r.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_update_lights_renders_the_changed_scene(engine):
    """Emission and environment changed after upload: update_lights must leave the renderer drawing the
    changed scene, exactly as a fresh upload of it, and as the oracle on it."""
    s = make_scene("showcase").build()
    w, h, spp = 64, 40, 8
    rs = RenderSettings.make(w, h, max_depth=8, rr_depth=3, seed=12)
    cam = s.camera_view(w / h)
    r = uploaded(s, w, h, engine)
    r.render(cam, rs, spp)
    before = r.download_image()
    glow = s.material(1)
    assert glow.emissive_tex >= 0
    glow.emission[:] = (0.5, 4.0, 9.0)
    s.set_material(1, glow)
    ground = s.material(0)
    ground.emission[:] = (0.3, 0.1, 0.0)  # a new emitter: the light tables grow
    s.set_material(0, ground)
    s.set_env_color((0.2, 0.3, 0.5), 1.5)
    s.rebuild_lights()
    ok, err = r.update_lights(s)
    assert ok, err
    assert r.samples == 0  # update_lights resets accumulation (gpu_renderer.cu:332)
    r.render(cam, rs, spp)
    after = r.download_image()
    fresh = uploaded(s, w, h, engine)
    fresh.render(cam, rs, spp)
    assert after.tobytes() == fresh.download_image().tobytes()
    ref = O.render(s.desc(), cam, rs, spp)
    agree = radiance_agreement(after, ref)
    print(f"update_lights engine={engine}: {100 * agree:.3f}% within 0.1% of the oracle, mean {before.mean():.4f} -> {after.mean():.4f}")
    assert agree >= 0.999
    assert not np.allclose(before, after)
    r.close()
    fresh.close()
