"""The compact G-buffer of the real-time workload (fp32 RGB colour/albedo/emission, fp32 depth, fp16
octahedral normal, int32 material) and the denoiser on it.

* The compact sample is the same sample as the float4 G-buffer of render_denoised (the reference's
  sample_kernel): encoding the device's own float4 planes with the oracle's encoder gives the device's
  compact planes bit for bit, which pins the device encoder against its CPU restatement.
* The denoiser on compact frames matches the oracle's compact denoiser (decode, then the float4
  pipeline) within 1e-3 relative, as the float4 denoiser does, over a moving camera.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd import gpu, host
from photon_amd.types import RenderSettings, SvgfSettings

from gpu_util import uploaded

pytestmark = pytest.mark.gpu


def _frames(scene, w, h, frames, seed=7):
    rs = RenderSettings.make(w, h, seed=seed)
    cam0 = scene.camera()
    out, prev = [], None
    for f in range(frames):
        ang = np.deg2rad(0.5 * f)
        p = np.array(cam0["position"], np.float64) - np.array(cam0["target"], np.float64)
        c, s_ = np.cos(ang), np.sin(ang)
        p = np.array([c * p[0] + s_ * p[2], p[1], -s_ * p[0] + c * p[2]]) + np.array(cam0["target"])
        scene.set_camera(p, cam0["target"], cam0["up"], cam0["vfov"])
        cam = scene.camera_view(w / h)
        out.append((cam, prev if prev is not None else cam, O.compact_gbuffer(*O.render_gbuffer(scene.desc(), cam, rs, f))))
        prev = cam
    scene.set_camera(cam0["position"], cam0["target"], cam0["up"], cam0["vfov"])
    return out


def _within(a, b):
    d = np.abs(a - b)
    return np.all((d <= 1e-3 * np.abs(b)) | (d <= 1e-6), axis=-1)


@pytest.mark.parametrize("name,w,h", [("cornell-c1", 100, 75), ("materials", 64, 48), ("glass", 33, 17)])
def test_gbuffer_sample_is_the_per_pixel_render(name, w, h):
    """The real-time G-buffer sample is the per-pixel kernel's one-sample render of frame 0: the same
    colour, albedo, normal and depth, bit for bit, also at sizes that leave partial blocks."""
    scene = host.Scene.builtin(name).build()
    rs = RenderSettings.make(w, h, seed=3)
    cam = scene.camera_view(w / h)
    planes = lambda r, bufs: [gpu.copy_device_to_numpy(r.buffer_ptr(b), w * h * 16, np.float32, (h, w, 4)) for b in bufs]
    r4 = uploaded(scene, w, h)
    r4.render_denoised(cam, cam, rs, SvgfSettings.make())
    color, albedo, nd = planes(r4, (gpu.BUF_COLOR, gpu.BUF_ALBEDO, gpu.BUF_NORMAL_DEPTH))
    r4.close()
    r1 = uploaded(scene, w, h, engine=gpu.ENGINE_MEGAKERNEL)
    r1.render(cam, rs, 1)
    acc, albedo1, nd1 = planes(r1, (gpu.BUF_ACCUM, gpu.BUF_ALBEDO, gpu.BUF_NORMAL_DEPTH))
    r1.close()
    # The accumulation adds the sample to zero, which turns a -0 component into +0; values are compared.
    assert np.array_equal(color[..., :3], acc[..., :3])
    assert np.array_equal(albedo.view(np.uint32), albedo1.view(np.uint32))
    assert np.array_equal(nd.view(np.uint32), nd1.view(np.uint32))


@pytest.mark.parametrize("name", ["cornell-c1", "materials"])
def test_compact_sample_is_the_encoded_float4_sample(name):
    scene = host.Scene.builtin(name).build()
    w, h = 96, 72
    rs = RenderSettings.make(w, h, seed=3)
    cam = scene.camera_view(w / h)
    r4 = uploaded(scene, w, h)
    r4.render_denoised(cam, cam, rs, SvgfSettings.make())
    planes = [gpu.copy_device_to_numpy(r4.buffer_ptr(b), w * h * 16, np.float32, (h, w, 4))
              for b in (gpu.BUF_COLOR, gpu.BUF_ALBEDO, gpu.BUF_NORMAL_DEPTH, gpu.BUF_POSITION, gpu.BUF_EMISSION)]
    r4.close()
    rc = uploaded(scene, w, h)
    gb = gpu.CompactGBuffer(w, h)
    rc.sample_compact(cam, rs, gb)
    got = gb.numpy()
    rc.close()
    want = O.compact_gbuffer(*planes)
    for k in ("color", "albedo", "emission", "depth", "normal", "material"):
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), k
    hit = got["material"] >= 0
    print(f"{name}: {hit.mean():.3f} of pixels hit")
    assert hit.any()
    # And the sample itself is the oracle's, within the workload's tolerance.
    ref = O.compact_gbuffer(*O.render_gbuffer(scene.desc(), cam, rs, 0))
    assert np.mean(got["material"] == ref["material"]) >= 0.999


def test_compact_svgf_matches_oracle():
    scene = host.Scene.builtin("cornell-c1").build()
    w, h = 120, 90
    frames = _frames(scene, w, h, 6)
    sv = SvgfSettings.make()
    ora = O.SvgfOracle(w, h)
    f = gpu.Svgf()
    f.resize(w, h)
    out = gpu.DeviceArray(w * h * 16)
    # A float4 frame first: switching formats must start a new history.
    g4 = O.render_gbuffer(scene.desc(), frames[0][0], RenderSettings.make(w, h, seed=7), 0)
    dev4 = [gpu.DeviceArray.from_numpy(p) for p in g4]
    f.run(*[d.ptr for d in dev4], frames[0][0], sv, out.ptr)
    for k, (cam, prev, c) in enumerate(frames):
        gb = gpu.CompactGBuffer.from_numpy(c, w, h)
        f.run_compact(gb, cam, prev, sv, out.ptr)
        a = out.numpy(np.float32, (h, w, 4))
        b = ora.run_compact(c, cam, prev, sv)
        ok = _within(a, b)
        print(f"frame {k}: {100 * ok.mean():.3f}% within 1e-3")
        assert np.all(np.isfinite(a))
        assert ok.mean() >= 0.999


def test_renderer_compact_frames_denoise():
    """The real-time loop of the workload: compact sample, then the compact denoiser, frame after frame,
    against the oracle denoiser on the device's own compact G-buffers."""
    scene = host.Scene.builtin("cornell-c1").build()
    w, h = 80, 64
    rs = RenderSettings.make(w, h, seed=5)
    r = uploaded(scene, w, h)
    f = gpu.Svgf()
    f.resize(w, h)
    ora = O.SvgfOracle(w, h)
    out = gpu.DeviceArray(w * h * 16)
    gb = gpu.CompactGBuffer(w, h)
    prev = None
    for k, (cam, _, _) in enumerate(_frames(scene, w, h, 3)):
        r.sample_compact(cam, rs, gb)
        f.run_compact(gb, cam, prev or cam, SvgfSettings.make(), out.ptr)
        a = out.numpy(np.float32, (h, w, 4))
        b = ora.run_compact(gb.numpy(), cam, prev or cam, SvgfSettings.make())
        assert _within(a, b).mean() >= 0.999, f"frame {k}"
        prev = cam
    r.close()


@pytest.mark.parametrize("iterations", [0, 1, 3, 6])
def test_compact_svgf_iteration_counts_match_oracle(iterations):
    """Other a-trous iteration counts than the default 5: none (the history and remodulation launches),
    fewer (the last pass remodulates at steps 1 and 4), and 6 (step 32, past the lattice kernel's steps,
    on the gathering fallback)."""
    scene = host.Scene.builtin("cornell-c1").build()
    w, h = 136, 80
    frames = _frames(scene, w, h, 4, seed=11)
    sv = SvgfSettings.make(atrous_iterations=iterations)
    ora = O.SvgfOracle(w, h)
    f = gpu.Svgf()
    f.resize(w, h)
    out = gpu.DeviceArray(w * h * 16)
    for k, (cam, prev, c) in enumerate(frames):
        gb = gpu.CompactGBuffer.from_numpy(c, w, h)
        f.run_compact(gb, cam, prev, sv, out.ptr)
        a = out.numpy(np.float32, (h, w, 4))
        b = ora.run_compact(c, cam, prev, sv)
        ok = _within(a, b)
        print(f"{iterations} iterations, frame {k}: {100 * ok.mean():.3f}% within 1e-3")
        assert np.all(np.isfinite(a))
        assert ok.mean() >= 0.999


@pytest.mark.parametrize("w,h", [(37, 21), (70, 5), (5, 70), (129, 33)])
def test_compact_svgf_odd_sizes_match_oracle(w, h):
    """Frames narrower than a lattice block (64 pixels) or shorter than the widest a-trous step (16 rows),
    and sizes that leave partial tiles and blocks at both edges."""
    scene = host.Scene.builtin("cornell-c1").build()
    frames = _frames(scene, w, h, 3, seed=13)
    sv = SvgfSettings.make()
    ora = O.SvgfOracle(w, h)
    f = gpu.Svgf()
    f.resize(w, h)
    out = gpu.DeviceArray(w * h * 16)
    for k, (cam, prev, c) in enumerate(frames):
        gb = gpu.CompactGBuffer.from_numpy(c, w, h)
        f.run_compact(gb, cam, prev, sv, out.ptr)
        a = out.numpy(np.float32, (h, w, 4))
        b = ora.run_compact(c, cam, prev, sv)
        ok = _within(a, b)
        print(f"{w}x{h} frame {k}: {100 * ok.mean():.3f}% within 1e-3")
        assert np.all(np.isfinite(a))
        assert ok.mean() >= 0.999
