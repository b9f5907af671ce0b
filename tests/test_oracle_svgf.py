"""Pins of the SVGF restatement (the reference has no CPU version, so these are the checks that fix it):
special cases with closed answers (a constant image on a constant G-buffer is a fixed point, misses pass
through, zero a-trous passes leave the temporal result, an a-trous pass with unit edge weights is the
dilated B3-spline convolution, the short-history variance estimate a 7x7 box-moment variance, and the
temporal pass over the first 1/alpha frames a running mean, all computed independently with scipy and
numpy), the reference's own behavioural tests
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


# ---- the compact G-buffer of the real-time workload ----------------------------------------------

def test_fp16_conversion_matches_numpy_round_to_nearest_even():
    """The integer fp16 conversion against numpy's float16 (IEEE round to nearest even), over normals,
    subnormals, exact ties, overflow, infinities and NaN; and the widening of every half bit pattern."""
    g = np.random.default_rng(3)
    vals = [0.0, -0.0, 1.0, -1.0, 65504.0, 65520.0, 65519.99, 1e9, -1e9, np.inf, -np.inf, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
            2.0 ** -14, 2.0 ** -14 * (1 - 2.0 ** -11), 1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, 1e-8, -1e-8]
    vals += list(g.standard_normal(3000) * 10.0 ** g.uniform(-9, 5, 3000))
    # Exact halfway cases between neighbouring halves (normal and subnormal range).
    h = g.integers(1, 0x7bff, 2000).astype(np.uint16)
    lo = h.view(np.float16).astype(np.float64)
    hi = (h + 1).view(np.float16).astype(np.float64)
    vals += list(((lo + hi) / 2).astype(np.float32))
    x = np.array(vals, np.float32)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    got = np.array([O.f32_to_f16(v) for v in x], np.uint16)
    assert np.array_equal(got, want)
    nan = O.f32_to_f16(np.float32(np.nan))
    assert (nan & 0x7c00) == 0x7c00 and (nan & 0x3ff) != 0
    bits = np.arange(1 << 16, dtype=np.uint32)
    ref = bits.astype(np.uint16).view(np.float16).astype(np.float32)
    got = np.array([O.f16_to_f32(b) for b in bits], np.float32)
    finite = np.isfinite(ref)
    assert np.array_equal(got[finite].view(np.uint32), ref[finite].view(np.uint32))
    assert np.array_equal(np.isnan(got), np.isnan(ref)) and np.array_equal(np.isinf(got), np.isinf(ref))


def test_octahedral_normals_round_trip():
    # Axis directions survive exactly, in both hemispheres.
    for n in ([0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]):
        assert np.array_equal(O.oct_decode(O.oct_encode(np.array(n, np.float32))), np.array(n, np.float32))
    # Random unit vectors come back within the fp16 grid's angular resolution, and unit length.
    g = np.random.default_rng(4)
    v = g.standard_normal((4000, 3))
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    back = np.array([O.oct_decode(O.oct_encode(n)) for n in v])
    ang = np.arccos(np.clip(np.sum(back.astype(np.float64) * v, axis=1), -1, 1))
    assert ang.max() < 1.5e-3
    assert np.allclose(np.linalg.norm(back, axis=1), 1.0, atol=1e-6)
    # The lower hemisphere is folded, not mirrored: z keeps its sign.
    far = np.abs(v[:, 2]) > 1e-3
    assert np.array_equal(np.sign(back[far, 2]), np.sign(v[far, 2]))
    # A zero normal (a miss) encodes as the origin.
    assert O.oct_encode(np.zeros(3, np.float32)) == 0


def _centre_positions(cam, w, h, depth):
    """Pixel-centre ray positions, restated in numpy float32 with the library's operation order."""
    f32 = np.float32
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    px, py = xs.astype(f32) + f32(0.5), ys.astype(f32) + f32(0.5)
    sx = (f32(2) * px / f32(w) - f32(1)) * f32(cam.tan_half_fov) * f32(cam.aspect)
    sy = (f32(1) - f32(2) * py / f32(h)) * f32(cam.tan_half_fov)
    fwd, right, up, pos = (np.array(list(v), f32) for v in (cam.forward, cam.right, cam.up, cam.position))
    d = fwd + right * sx[..., None]
    d = d + up * sy[..., None]
    l2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    l2 = l2 + d[..., 2] * d[..., 2]
    d = d * (f32(1) / np.sqrt(l2))[..., None]
    return pos + d * depth[..., None]


def test_compact_frames_denoise_like_their_decoded_planes():
    """A compact frame through or_svgf_run_compact equals the float4 pipeline on the planes it decodes
    to (decoded normals, centre-ray positions at the stored depth, w = 1), bit for bit, over two frames
    with a moving camera; the decoded planes here are restated independently in numpy."""
    w, h = 48, 32
    s = host.Scene.builtin("cornell").build()
    cam0 = s.camera()
    g = np.random.default_rng(6)
    fa, fb = O.SvgfOracle(w, h), O.SvgfOracle(w, h)
    prev = None
    for k in range(2):
        p = np.array(cam0["position"]) + np.array([0.03 * k, 0, 0])
        s.set_camera(p, cam0["target"], cam0["up"], cam0["vfov"])
        cam = s.camera_view(w / h)
        color = g.uniform(0, 2, (h, w, 3)).astype(np.float32)
        albedo = g.uniform(0.05, 1, (h, w, 3)).astype(np.float32)
        emission = np.where(g.uniform(size=(h, w, 1)) < 0.05, 1.0, 0.0).astype(np.float32) * np.ones(3, np.float32)
        depth = g.uniform(2, 4, (h, w)).astype(np.float32)
        material = np.where(g.uniform(size=(h, w)) < 0.08, -1, g.integers(0, 3, (h, w))).astype(np.int32)
        nrm = g.standard_normal((h, w, 3))
        nrm = (nrm / np.linalg.norm(nrm, axis=2, keepdims=True)).astype(np.float32)
        codes = np.array([[O.oct_encode(nrm[y, x]) for x in range(w)] for y in range(h)], np.uint32)
        compact = dict(color=color, albedo=albedo, emission=emission, depth=depth, normal=codes, material=material)
        one = np.ones((h, w, 1), np.float32)
        dec = np.array([[O.oct_decode(c) for c in row] for row in codes], np.float32)
        planes = (np.concatenate([color, one], 2), np.concatenate([albedo, one], 2), np.concatenate([dec, depth[..., None]], 2),
                  np.concatenate([_centre_positions(cam, w, h, depth), material[..., None].astype(np.float32)], 2),
                  np.concatenate([emission, one], 2))
        a = fa.run_compact(compact, cam, prev or cam, SvgfSettings.make())
        b = fb.run(*planes, prev or cam, SvgfSettings.make())
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"frame {k}"
        prev = cam


def test_compact_gbuffer_of_float4_planes():
    w, h = 40, 30
    s = host.Scene.builtin("cornell").build()
    cam = s.camera_view(w / h)
    g4 = O.render_gbuffer(s.desc(), cam, RenderSettings.make(w, h, seed=0), 0)
    c = O.compact_gbuffer(*g4)
    color, albedo, nd, pos, emission = g4
    assert np.array_equal(c["color"], color[..., :3]) and np.array_equal(c["albedo"], albedo[..., :3])
    assert np.array_equal(c["emission"], emission[..., :3]) and np.array_equal(c["depth"], nd[..., 3])
    assert np.array_equal(c["material"], pos[..., 3].astype(np.int32))
    hit = c["material"] >= 0
    dec = np.array([O.oct_decode(v) for v in c["normal"][hit]])
    assert np.max(np.abs(dec - nd[..., :3][hit])) < 2e-3


def _b3_taps(step):
    """The a-trous footprint as a dense kernel: the B3 spline (1/16, 1/4, 3/8, 1/4, 1/16) in x and y, taps
    `step` pixels apart."""
    k1 = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])
    k = np.zeros((4 * step + 1, 4 * step + 1))
    k[::step, ::step] = np.outer(k1, k1)
    return k


def _unit_weight_gbuffer(w, h, material):
    """A G-buffer whose edge-stopping weights are all 1: one normal (n.n = 1, so the normal weight is 1 for
    any exponent) and one depth (zero gradient, zero depth difference)."""
    nd = np.zeros((h, w, 4), np.float32)
    nd[..., 2] = 1
    nd[..., 3] = 5.0
    pos = np.zeros((h, w, 4), np.float32)
    pos[..., 3] = material
    return nd, pos


def test_atrous_with_unit_edge_weights_is_the_dilated_b3_convolution():
    """With every edge-stopping weight 1 (flat G-buffer, colour sigma ~1e30 so the luminance weight is
    exp(-1e-30) = 1), an a-trous pass is the normalised B3-spline convolution at its step: colour =
    sum k x / sum k, variance = sum k^2 v / (sum k)^2, over the taps inside the image. Computed here with
    scipy's correlate (zero outside the image, so the same correlate of ones gives the in-image weight)."""
    from scipy.ndimage import correlate
    w, h = 37, 29
    rng = np.random.default_rng(7)
    src = rng.random((h, w, 4)).astype(np.float32)
    src[..., 3] = 0.1 + 0.9 * src[..., 3]  # variances away from 0, so sigma_l is huge
    nd, pos = _unit_weight_gbuffer(w, h, 0.0)
    f = O.SvgfOracle(w, h)
    for step in (1, 2, 4, 8):
        out = f.atrous(nd, pos, src, step, 1e30, 128.0, 1.0)
        k = _b3_taps(step)
        den = correlate(np.ones((h, w)), k, mode="constant", cval=0.0)
        for c in range(3):
            want = correlate(src[..., c].astype(np.float64), k, mode="constant", cval=0.0) / den
            np.testing.assert_allclose(out[..., c], want, rtol=2e-6, atol=0)
        want_v = correlate(src[..., 3].astype(np.float64), k * k, mode="constant", cval=0.0) / (den * den)
        np.testing.assert_allclose(out[..., 3], want_v, rtol=2e-6, atol=0)


def test_atrous_taps_only_the_same_material_and_passes_misses_through():
    """Material ids partition the taps: each pixel's result is the normalised B3 convolution restricted to
    the pixels of its own material (a masked correlate per id), and misses (id < 0) keep their value."""
    from scipy.ndimage import correlate
    w, h = 33, 27
    rng = np.random.default_rng(8)
    src = rng.random((h, w, 4)).astype(np.float32)
    src[..., 3] = 0.1 + 0.9 * src[..., 3]
    mat = rng.integers(-1, 3, size=(h, w)).astype(np.float32)  # ids -1 (miss), 0, 1, 2 scattered
    mat[10:20, 5:25] = 1  # and one large block
    nd, pos = _unit_weight_gbuffer(w, h, mat)
    f = O.SvgfOracle(w, h)
    with np.errstate(invalid="ignore", divide="ignore"):
        for step in (1, 2, 4):
            out = f.atrous(nd, pos, src, step, 1e30, 128.0, 1.0)
            k = _b3_taps(step)
            for m in (0.0, 1.0, 2.0):  # den > 0 wherever mat == m (the centre tap); elsewhere 0/0 is not compared
                mask = (mat == m).astype(np.float64)
                den = correlate(mask, k, mode="constant", cval=0.0)
                sel = mat == m
                for c in range(3):
                    num = correlate(src[..., c] * mask, k, mode="constant", cval=0.0)
                    np.testing.assert_allclose(out[..., c][sel], (num / den)[sel], rtol=2e-6, atol=0)
                num_v = correlate(src[..., 3] * mask, k * k, mode="constant", cval=0.0)
                np.testing.assert_allclose(out[..., 3][sel], (num_v / (den * den))[sel], rtol=2e-6, atol=0)
            assert np.array_equal(out[mat < 0], src[mat < 0])


def test_variance_estimate_with_unit_weights_is_the_7x7_box_moment_variance():
    """While the history is shorter than 4 frames, the variance pass replaces each pixel by a 7x7 bilateral
    mean and estimates variance spatially. With unit edge weights (flat G-buffer, one material) that is a
    box filter over the in-image pixels: colour = box(illum) / box(1), variance = max(0, box(l^2)/n -
    (box(l)/n)^2) * 4 / history length, from the frame's own illumination and moments planes. Pixels
    whose history has reached 4 keep their temporal colour and variance."""
    from scipy.ndimage import correlate
    w, h = 31, 23
    rng = np.random.default_rng(9)
    box = np.ones((7, 7))
    f = O.SvgfOracle(w, h)
    n = correlate(np.ones((h, w)), box, mode="constant", cval=0.0)
    # A still camera and points 10 units from it inside its view: every pixel reprojects onto itself, so
    # the history grows by one a frame, while depth (10) and normal stay the same everywhere.
    cam = CameraView()
    cam.position[:] = (0, 0, -10)
    cam.forward[:] = (0, 0, 1)
    cam.right[:] = (1, 0, 0)
    cam.up[:] = (0, 1, 0)
    cam.tan_half_fov, cam.aspect = 0.5, w / h
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    d = np.stack([(2 * (xs + 0.5) / w - 1) * 0.4, (1 - 2 * (ys + 0.5) / h) * 0.4, np.ones((h, w))], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    for frame in range(1, 6):
        color, albedo, nd, pos, emission = flat_gbuffer(w, h, (0, 0, 0), depth=10.0)
        pos[..., :3] = np.array(cam.position[:]) + 10.0 * d
        color[..., :3] = rng.random((h, w, 3))
        f.run(color, albedo, nd, pos, emission, cam, SvgfSettings.make(atrous_iterations=0))
        illum, moments, est = f.plane(4), f.plane(5), f.plane(6)
        length = illum[..., 3]
        np.testing.assert_allclose(length, frame, rtol=1e-6)  # bilinear history weights may round a hair below
        spatial = length < 4.0
        assert spatial.all() if frame < 4 else (~spatial).any()
        for c in range(3):
            want = correlate(illum[..., c].astype(np.float64), box, mode="constant", cval=0.0) / n
            np.testing.assert_allclose(est[..., c][spatial], want[spatial], rtol=2e-6, atol=0)
        m1 = correlate(moments[..., 0].astype(np.float64), box, mode="constant", cval=0.0) / n
        m2 = correlate(moments[..., 1].astype(np.float64), box, mode="constant", cval=0.0) / n
        want_v = np.maximum(0.0, m2 - m1 * m1) * 4.0 / length
        np.testing.assert_allclose(est[..., 3][spatial], want_v[spatial], rtol=1e-4, atol=1e-7)
        assert np.array_equal(est[..., :3][~spatial], illum[..., :3][~spatial])
        assert np.array_equal(est[..., 3][~spatial], moments[..., 2][~spatial])


def test_temporal_accumulation_is_the_running_mean_while_the_history_is_short():
    """With a still camera whose pixels reproject onto their own centres, the temporal pass blends history
    and the new frame with weight max(alpha, 1/length). For the first 1/alpha frames that is the running
    mean of the demodulated inputs, the moments are the running means of l and l^2, and the temporal
    variance is the population variance of the luminances so far (numpy's mean and var here). The history
    keeps the spatially filtered illumination; a material id of its own per pixel leaves the 7x7 filter
    only the centre tap, so the filter is the identity and the running mean is exact."""
    w, h = 20, 14
    rng = np.random.default_rng(10)
    cam = CameraView()
    cam.position[:] = (1, 2, -10)
    cam.forward[:] = (0, 0, 1)
    cam.right[:] = (1, 0, 0)
    cam.up[:] = (0, 1, 0)
    cam.tan_half_fov, cam.aspect = 0.5, w / h
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    sx = 2 * (xs + 0.5) / w - 1
    sy = 1 - 2 * (ys + 0.5) / h
    d = np.stack([sx * 0.5 * (w / h), sy * 0.5, np.ones((h, w))], -1)  # through each pixel centre
    t = 5 + 3 * rng.random((h, w, 1))  # points at varied depths along those rays
    p = np.array(cam.position[:]) + t * d
    alpha = 0.2
    settings = SvgfSettings.make(atrous_iterations=0, alpha_color=alpha, alpha_moments=alpha)
    f = O.SvgfOracle(w, h)
    seen = []
    for frame in range(1, int(round(1 / alpha)) + 1):
        color, albedo, nd, pos, emission = flat_gbuffer(w, h, (0, 0, 0))
        pos[..., :3] = p
        pos[..., 3] = np.arange(h * w).reshape(h, w)
        nd[..., 3] = np.linalg.norm(p - np.array(cam.position[:]), axis=-1)
        color[..., :3] = rng.random((h, w, 3))
        albedo[..., :3] = 0.25 + 0.75 * rng.random((h, w, 3))
        f.run(color, albedo, nd, pos, emission, cam, settings)
        seen.append(color[..., :3].astype(np.float64) / albedo[..., :3])
        x = np.stack(seen)
        lum_x = 0.2126 * x[..., 0] + 0.7152 * x[..., 1] + 0.0722 * x[..., 2]
        illum, moments = f.plane(4), f.plane(5)
        np.testing.assert_allclose(illum[..., 3], frame, rtol=1e-6)
        np.testing.assert_allclose(illum[..., :3], x.mean(0), rtol=1e-5)
        np.testing.assert_allclose(moments[..., 0], lum_x.mean(0), rtol=1e-5)
        np.testing.assert_allclose(moments[..., 1], (lum_x ** 2).mean(0), rtol=1e-5)
        np.testing.assert_allclose(moments[..., 2], lum_x.var(0), rtol=2e-3, atol=1e-6)
