"""Pins of the SVGF restatement (the reference has no CPU version, so these are the checks that fix it):
special cases with closed answers (a constant image on a constant G-buffer is a fixed point, misses pass
through, zero a-trous passes leave the temporal result), the reference's own behavioural tests
(test_gpu.cpp:105-164: frame-1 error below half the 1-spp error, later frames better, finite,
non-negative, mean within 5%, stable under camera motion) and the temporal bookkeeping."""
import numpy as np

from oracle import oracle_py as O
from photon_amd import host
from photon_amd.types import CameraView, RenderSettings, SvgfSettings


def lum(img):
    return 0.2126 * img[..., 0] + 0.7152 * img[..., 1] + 0.0722 * img[..., 2]


def flat_gbuffer(w, h, colour, depth=5.0, material=0.0):
    color = np.zeros((h, w, 4), np.float32)
    color[..., :3] = colour
    color[..., 3] = 1
    albedo = np.ones((h, w, 4), np.float32)
    nd = np.zeros((h, w, 4), np.float32)
    nd[..., 2] = 1
    nd[..., 3] = depth
    pos = np.zeros((h, w, 4), np.float32)
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    pos[..., 0], pos[..., 1], pos[..., 2], pos[..., 3] = xs, ys, 0, material
    emission = np.zeros((h, w, 4), np.float32)
    emission[..., 3] = 1
    return color, albedo, nd, pos, emission


def test_constant_image_is_a_fixed_point():
    w, h = 40, 30
    g = flat_gbuffer(w, h, (0.3, 0.5, 0.7))
    f = O.SvgfOracle(w, h)
    cam = CameraView()
    for _ in range(3):
        out = f.run(*g, cam, SvgfSettings.make())
        assert np.allclose(out[..., :3], (0.3, 0.5, 0.7), rtol=1e-6, atol=0)


def test_misses_pass_through_and_emission_is_added_back():
    w, h = 32, 24
    color, albedo, nd, pos, emission = flat_gbuffer(w, h, (0.2, 0.2, 0.2))
    pos[:, :8, 3] = -1  # misses: the raw colour passes through
    color[:, :8, :3] = 7.0
    emission[:, 20:, :3] = 1.5  # emitters: illumination is demodulated without it and remodulated with it
    color[:, 20:, :3] += 1.5
    out = O.SvgfOracle(w, h).run(color, albedo, nd, pos, emission, CameraView(), SvgfSettings.make())
    assert np.array_equal(out[:, :8], color[:, :8])
    assert np.allclose(out[:, 24:, :3], 1.7, rtol=1e-5)


def test_zero_passes_return_the_temporal_estimate():
    w, h = 24, 16
    rng = np.random.default_rng(1)
    color, albedo, nd, pos, emission = flat_gbuffer(w, h, (0, 0, 0))
    color[..., :3] = rng.random((h, w, 3))
    albedo[..., :3] = 0.5
    f = O.SvgfOracle(w, h)
    out = f.run(color, albedo, nd, pos, emission, CameraView(), SvgfSettings.make(atrous_iterations=0))
    # No history, short history: the 7x7 variance pass replaces the colour by its bilateral mean,
    # which remodulation multiplies back by the albedo.
    var = f.plane(6)
    assert np.allclose(out[..., :3], var[..., :3] * 0.5, rtol=1e-6)
    assert (f.plane(4)[..., 3] == 1).all()  # history length 1 after the first frame


def test_history_grows_and_caps_at_64():
    w, h = 16, 16
    g = flat_gbuffer(w, h, (0.4, 0.4, 0.4), material=2.0)
    f = O.SvgfOracle(w, h)
    cam = CameraView()
    cam.position[:] = (7.5, 7.5, -10)
    cam.forward[:] = (0, 0, 1)
    cam.right[:] = (1, 0, 0)
    cam.up[:] = (0, -1, 0)
    cam.tan_half_fov, cam.aspect = 0.5, 1.0
    # Positions on the z = 0 plane seen by this camera; depth = distance to it.
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    g[3][..., 0] = (xs + 0.5 - w / 2) / (w / 2) * 0.5 * 10 + 7.5
    g[3][..., 1] = 7.5 - (ys + 0.5 - h / 2) / (h / 2) * 0.5 * 10
    d = np.sqrt((g[3][..., 0] - 7.5) ** 2 + (g[3][..., 1] - 7.5) ** 2 + 100)
    g[2][..., 3] = d
    lens = []
    for _ in range(70):
        f.run(*g, cam, SvgfSettings.make())
        lens.append(f.plane(4)[h // 2, w // 2, 3])
    assert lens[0] == 1 and lens[1] == 2 and lens[10] == 11 and lens[-1] == 64


def _cornell_frames(w, h, frames, moving):
    s = host.Scene.builtin("cornell").build()
    cam0 = s.camera()
    out, prev = [], None
    for k in range(frames):
        if moving:
            p = np.array(cam0["position"]) + np.array([0.05 * (k + 1), 0, 0])
            s.set_camera(p, np.array(cam0["target"]) + np.array([0.05 * (k + 1), 0, 0]), cam0["up"], cam0["vfov"])
        cam = s.camera_view(w / h)
        g = O.render_gbuffer(s.desc(), cam, RenderSettings.make(w, h, seed=0), k)
        out.append((cam, prev or cam, g))
        prev = cam
    return s, out


def test_denoiser_reduces_error_like_the_reference_expects():
    w = h = 160  # the reference's test resolution
    s, frames = _cornell_frames(w, h, 16, moving=False)
    cam = frames[0][0]
    ref = O.render(s.desc(), cam, RenderSettings.make(w, h, seed=0), 1024)

    def mse(a):
        return np.mean((np.minimum(lum(a), 1) - np.minimum(lum(ref), 1)) ** 2)

    noisy = frames[0][2][0][..., :3]
    f = O.SvgfOracle(w, h)
    first = f.run(*frames[0][2], cam, SvgfSettings.make())[..., :3]
    for cam_k, prev_k, g in frames[1:]:
        out = f.run(*g, prev_k, SvgfSettings.make())[..., :3]
    e_noisy, e_first, e_last = mse(noisy), mse(first), mse(out)
    assert e_first < 0.5 * e_noisy
    assert e_last < e_first
    assert np.isfinite(out).all() and out.min() >= 0
    assert abs(lum(out).mean() - lum(ref).mean()) < 0.05 * lum(ref).mean()


def test_denoiser_handles_camera_motion():
    w, h = 64, 48
    _, frames = _cornell_frames(w, h, 12, moving=True)
    f = O.SvgfOracle(w, h)
    for cam, prev, g in frames:
        out = f.run(*g, prev, SvgfSettings.make())
    assert np.isfinite(out).all() and out[..., :3].min() >= 0 and lum(out).mean() > 0.02
    # Reprojection accepted history for most of the frame.
    assert (f.plane(4)[..., 3] > 4).mean() > 0.5
