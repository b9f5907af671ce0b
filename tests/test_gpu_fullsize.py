"""Parity at the workloads' full sizes.

The oracle checks what it can finish in seconds: a deterministic subsample of the rays, the whole
image where the scene is small, and sampled pixels where it is not. Properties that do not depend on
size check the rest:
  * any-hit occlusion of a camera ray agrees with the closest hit existing, on every ray;
  * the wavefront engine and the one-thread-per-pixel kernel give bit-identical accumulations, also with
    one sample per sub-stream on more path slots than pixels;
  * the SVGF output is finite and within 1e-3 of the oracle on identical G-buffers.

Workloads:
  * traversal: 262,144-triangle sphere soup (256 spheres of 1024 triangles, 10% alpha-masked), 2048^2
    camera rays and 4,194,304 shadow rays;
  * Cornell box: 512^2, 16 spp, depth 8, RR 3, seed 1;
  * sphere field: 1,048,578 triangles, 1920x1080, 64 spp, depth 16, seed 1;
  * SVGF: 1600x900 G-buffers of the field, two frames with a 0.5 degree orbit.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd import gpu, host
from photon_amd.types import RenderSettings, SvgfSettings

from gpu_util import bitwise_fraction, radiance_agreement, trace_on_gpu, uploaded

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def soup():
    return host.Scene.spheres(256, box=100.0, alpha_fraction=0.1, seed=1).build()


@pytest.fixture(scope="module")
def field():
    return host.Scene.field(1024, seed=1, env_width=2048).build()


def test_traversal_workload(soup):
    cam = soup.camera_view(1.0)
    cam_rays = host.camera_rays(cam, 2048, 2048)
    shadow = soup.shadow_rays(4194304, seed=2)
    r = uploaded(soup)
    closest = trace_on_gpu(r, cam_rays)
    occluded = trace_on_gpu(r, cam_rays, anyhit=True)
    shadow_hits = trace_on_gpu(r, shadow, anyhit=True)
    r.close()
    # Every ray: a camera ray is occluded exactly when it has a closest hit.
    assert np.array_equal(closest["tri"] >= 0, occluded["tri"] >= 0)
    # A deterministic 1/64 subsample of each ray set against the oracle.
    sub = np.arange(0, len(cam_rays), 64)
    o = O.trace(soup.desc(), cam_rays[sub])
    hit = o["tri"] >= 0
    print(f"camera rays: {hit.mean():.3f} hit; subsample {len(sub)}")
    assert np.array_equal(closest["tri"][sub], o["tri"])
    assert np.max(np.abs(closest["t"][sub][hit] - o["t"][hit]) / o["t"][hit], initial=0) <= 1e-6
    sub = np.arange(0, len(shadow), 64)
    o = O.trace(soup.desc(), shadow[sub], anyhit=True)
    print(f"shadow rays: {np.mean(o['tri'] >= 0):.3f} occluded")
    assert np.array_equal(shadow_hits["tri"][sub] >= 0, o["tri"] >= 0)
    assert np.array_equal(shadow_hits["tri"][sub], o["tri"])


def test_cornell_workload_whole_image():
    scene = host.Scene.builtin("cornell-c1").build()
    w = h = 512
    rs = RenderSettings.make(w, h, max_depth=8, rr_depth=3, seed=1)
    cam = scene.camera_view(1.0)
    r = uploaded(scene, w, h, engine=gpu.ENGINE_WAVEFRONT)
    r.render(cam, rs, 16)
    img = r.download_image()
    r.close()
    ref = O.render(scene.desc(), cam, rs, 16)
    agree = radiance_agreement(img, ref)
    print(f"cornell 512^2 x 16: {100 * agree:.4f}% within 0.1%, {100 * bitwise_fraction(img, ref):.3f}% bitwise")
    assert agree >= 0.999


def test_field_workload_sampled_pixels(field):
    W, H, SPP = 1920, 1080, 64
    rs = RenderSettings.make(W, H, max_depth=16, rr_depth=3, seed=1)
    cam = field.camera_view(W / H)
    accs = {}
    for engine in (gpu.ENGINE_WAVEFRONT, gpu.ENGINE_MEGAKERNEL):
        r = uploaded(field, W, H, engine=engine)
        r.render(cam, rs, SPP)
        accs[engine] = gpu.copy_device_to_numpy(r.buffer_ptr(gpu.BUF_ACCUM), W * H * 16, np.float32, (H, W, 4))
        if engine == gpu.ENGINE_WAVEFRONT:
            img = r.download_image()
        r.close()
    # The two engines accumulate every pixel in the same order.
    assert accs[gpu.ENGINE_WAVEFRONT].tobytes() == accs[gpu.ENGINE_MEGAKERNEL].tobytes()
    # One sample per sub-stream on four path slots per pixel (the benchmark's tuned shape): the partial
    # sums are single samples added in order, so the accumulation is still the in-order one, bit for bit.
    r = uploaded(field, W, H, engine=gpu.ENGINE_WAVEFRONT)
    r.set_wavefront_config(sample_split=SPP, path_slots=4 * W * H, block=128, stack_depth=16)
    r.render(cam, rs, SPP)
    acc64 = gpu.copy_device_to_numpy(r.buffer_ptr(gpu.BUF_ACCUM), W * H * 16, np.float32, (H, W, 4))
    r.close()
    assert acc64.tobytes() == accs[gpu.ENGINE_MEGAKERNEL].tobytes()
    # With 8 sub-streams per pixel: the in-order sums of each eighth, added in order.
    r = uploaded(field, W, H, engine=gpu.ENGINE_WAVEFRONT)
    r.set_wavefront_config(sample_split=8)
    r.render(cam, rs, SPP)
    img8 = r.download_image()
    r.close()
    g = np.random.default_rng(5)
    n = 384
    xs, ys = g.integers(0, W, n), g.integers(0, H, n)
    rgb, _ = O.render_samples(field.desc(), cam, rs, np.repeat(xs, SPP), np.repeat(ys, SPP), np.tile(np.arange(SPP), n))
    # The reference accumulates per render call in sample order, then divides.
    sums = np.zeros((n, 3), np.float32)
    for k in range(SPP):
        sums += rgb[k::SPP]
    ref = sums * (np.float32(1.0) / np.float32(SPP))
    parts = np.zeros((8, n, 3), np.float32)
    for k in range(SPP):
        parts[k // 8] += rgb[k::SPP]
    split = np.zeros((n, 3), np.float32)
    for p in range(8):
        split += parts[p]
    ref8 = split * (np.float32(1.0) / np.float32(SPP))
    for name, image, want in (("in order", img, ref), ("8 sub-streams", img8, ref8)):
        got = image[ys, xs]
        agree = radiance_agreement(got[:, None, :], want[:, None, :])
        print(f"field 1080p x 64 {name}, {n} pixels: {100 * agree:.3f}% within 0.1%, {100 * bitwise_fraction(got, want):.2f}% bitwise")
        assert agree == 1.0


def test_svgf_workload_gbuffers(field):
    W, H = 1600, 900
    rs = RenderSettings.make(W, H, max_depth=16, rr_depth=3, seed=1)
    sv = SvgfSettings.make()
    cam0 = field.camera()
    views = []
    for f in range(8):  # the history reaches 4 frames (the end of the spatial variance estimate) and beyond
        a = np.deg2rad(0.5 * f)
        p = np.array(cam0["position"]) - np.array(cam0["target"])
        p = np.array([np.cos(a) * p[0] + np.sin(a) * p[2], p[1], -np.sin(a) * p[0] + np.cos(a) * p[2]]) + np.array(cam0["target"])
        field.set_camera(p, cam0["target"], cam0["up"], cam0["vfov"])
        views.append(field.camera_view(W / H))
    field.set_camera(cam0["position"], cam0["target"], cam0["up"], cam0["vfov"])
    ora = O.SvgfOracle(W, H)
    f = gpu.Svgf()
    f.resize(W, H)
    out = gpu.DeviceArray(W * H * 16)
    prev = views[0]
    for k, cam in enumerate(views):
        g = O.render_gbuffer(field.desc(), cam, rs, k)
        dev = [gpu.DeviceArray.from_numpy(p) for p in g]
        f.run(*[d.ptr for d in dev], prev, sv, out.ptr)
        a = out.numpy(np.float32, (H, W, 4))
        b = ora.run(*g, prev, sv)
        assert np.all(np.isfinite(a))
        d = np.abs(a - b)
        ok = np.all((d <= 1e-3 * np.abs(b)) | (d <= 1e-6), axis=-1)
        print(f"svgf 1600x900 frame {k}: {100 * ok.mean():.4f}% within 1e-3")
        assert ok.mean() >= 0.999
        prev = cam
    long_history = np.mean(ora.plane(4)[..., 3] >= 4)
    print(f"history of 4 frames or more at the last frame: {100 * long_history:.1f}% of pixels")
    assert long_history > 0.5


def test_svgf_workload_compact_gbuffers(field):
    """The real-time workload in its own G-buffer format: the compact sample traced on the device at
    1600x900, denoised on the device and by the oracle from the same compact planes, ten frames of the
    orbit (0.5 degrees a frame), every frame compared."""
    W, H = 1600, 900
    rs = RenderSettings.make(W, H, max_depth=16, rr_depth=3, seed=1)
    sv = SvgfSettings.make()
    cam0 = field.camera()
    r = uploaded(field, W, H)
    f = gpu.Svgf()
    f.resize(W, H)
    out = gpu.DeviceArray(W * H * 16)
    gb = gpu.CompactGBuffer(W, H)
    ora = O.SvgfOracle(W, H)
    prev = None
    for k in range(10):
        a = np.deg2rad(0.5 * k)
        p = np.array(cam0["position"]) - np.array(cam0["target"])
        p = np.array([np.cos(a) * p[0] + np.sin(a) * p[2], p[1], -np.sin(a) * p[0] + np.cos(a) * p[2]]) + np.array(cam0["target"])
        field.set_camera(p, cam0["target"], cam0["up"], cam0["vfov"])
        cam = field.camera_view(W / H)
        r.sample_compact(cam, rs, gb)
        f.run_compact(gb, cam, prev or cam, sv, out.ptr)
        got = out.numpy(np.float32, (H, W, 4))
        want = ora.run_compact(gb.numpy(), cam, prev or cam, sv)
        assert np.all(np.isfinite(got))
        d = np.abs(got - want)
        ok = np.all((d <= 1e-3 * np.abs(want)) | (d <= 1e-6), axis=-1)
        print(f"compact svgf 1600x900 frame {k}: {100 * ok.mean():.4f}% within 1e-3")
        assert ok.mean() >= 0.999
        prev = cam
    assert np.mean(ora.plane(4)[..., 3] >= 4) > 0.5
    field.set_camera(cam0["position"], cam0["target"], cam0["up"], cam0["vfov"])
    r.close()
