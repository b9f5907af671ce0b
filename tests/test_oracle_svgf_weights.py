"""Pins of the SVGF restatement's edge-stopping weights and of its reprojection under camera motion
(svgf.cu:37-92 temporal, :96-132 variance, :134-186 a-trous), by special cases whose answers geometry
fixes:

* the a-trous normal weight of a tap whose normal is tilted by a known angle is max(0, cos)^phi_n;
* the a-trous depth weight on a plane of known slope g: the central-difference gradient is g, a tap at
  (dx, dy) differs by g dx s in depth, and its weight is exp(-|g dx s| / (phi_z g |(dx,dy)| s + 1e-3 z +
  1e-4));
* the a-trous luminance weight of a known variance impulse through the 3x3 Gaussian prefilter
  (1/4, 1/8, 1/16, clamped at the image border): exp(-|dl| / (phi_c sqrt(var) + 1e-4)), and the output
  variance sum w^2 v / (sum w)^2;
* the variance pass's 7x7 bilateral weights, max(0, n.n')^phi_n exp(-|dz| / (0.05 z + 1e-3));
* reprojection under a pure camera translation by whole pixels (the history is the shifted frame) and by
  half pixels (four bilinear taps of weight 1/4, renormalised at the border), and the rejection of taps
  whose material, distance (|dz| > 0.05 z + 0.01) or normal (n.n' < 0.9) changed.

Each single-tap case isolates one weight: every other tap has zero colour and the centre's normal and
depth, so its weight is its spline coefficient k, and the pass's colour output k_t w / (1 - k_t + k_t w)
(the 5x5 spline sums to 1 away from the border) gives w back.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd.types import CameraView, SvgfSettings

B3 = np.array([3 / 8, 1 / 4, 1 / 16])  # by |offset|
W, H = 32, 32  # powers of two: pixel centres, projections and shifts are exact in float
CX = CY = 16


def k_tap(dx, dy):
    return B3[abs(dx)] * B3[abs(dy)]


def base_planes(depth=5.0):
    nd = np.zeros((H, W, 4), np.float32)
    nd[..., 2] = 1
    nd[..., 3] = depth
    pos = np.zeros((H, W, 4), np.float32)
    src = np.zeros((H, W, 4), np.float32)
    return nd, pos, src


def weight_from_output(out, k, value=1.0):
    """The isolated tap's edge-stopping weight from the centre's output colour, value k w / (1 - k + k w)
    for a tap of colour `value`."""
    out = np.float64(out)
    return out * (1 - k) / (k * (value - out))


@pytest.mark.parametrize("phi_n", [128.0, 16.0])
def test_atrous_normal_weight_is_the_clamped_cosine_power(phi_n):
    f = O.SvgfOracle(W, H)
    for deg in (0.0, 2.0, 5.0, 11.0, 60.0, 95.0, 150.0):
        nd, pos, src = base_planes()
        src[..., 3] = 1.0
        th = np.radians(deg)
        nd[CY, CX + 1, :3] = (np.sin(th), 0.0, np.cos(th))
        src[CY, CX + 1, :3] = 1.0
        out = f.atrous(nd, pos, src, 1, 1e30, phi_n, 1.0)
        got = weight_from_output(out[CY, CX, 0], k_tap(1, 0))
        want = max(0.0, np.cos(th)) ** phi_n
        assert abs(got - want) <= 2e-5 * max(want, 1e-3), (deg, got, want)


def test_atrous_depth_weight_on_a_slanted_plane():
    """Depth z(x) = z0 + g x: the centre's gradient is g (half the larger central difference), and each
    tap's weight follows from its offset alone."""
    z0, g, phi_z = 40.0, 0.25, 1.3
    f = O.SvgfOracle(W, H)
    for step in (1, 2, 4):
        for tap in ((1, 0), (2, 0), (2, 1), (1, 2), (0, 2), (-2, -2)):
            nd, pos, src = base_planes()
            xs = np.arange(W, dtype=np.float64)
            nd[..., 3] = (z0 + g * (xs - CX))[None, :]
            src[..., 3] = 1.0
            tx, ty = CX + tap[0] * step, CY + tap[1] * step
            src[ty, tx, :3] = 1.0
            out = f.atrous(nd, pos, src, step, 1e30, 128.0, phi_z)
            z = np.float64(nd[CY, CX, 3])
            num = den = 0.0
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    dist = np.hypot(dx, dy) * step
                    wz = np.exp(-abs(g * dx * step) / (phi_z * g * dist + 1e-3 * z + 1e-4))
                    den += k_tap(dx, dy) * wz
                    if (dx, dy) == tap:
                        num = k_tap(dx, dy) * wz
            assert abs(out[CY, CX, 0] - num / den) <= 1e-5 * num / den, (step, tap)


def test_atrous_luminance_weight_of_a_variance_impulse_through_the_prefilter():
    phi_c = 4.0
    f = O.SvgfOracle(W, H)
    # Variances at the centre, at an edge neighbour and at a corner neighbour; 0 elsewhere.
    v_c, v_e, v_d = 0.09, 0.64, 0.36
    var = 0.25 * v_c + 0.125 * v_e + 0.0625 * v_d
    sigma = phi_c * np.sqrt(var) + 1e-4
    for lum_tap, tap, v_tap in ((0.3, (2, 0), 0.5), (1.7, (-1, 2), 0.0), (0.05, (2, -2), 2.0)):
        nd, pos, src = base_planes()
        src[CY, CX, 3] = v_c
        src[CY, CX + 1, 3] = v_e
        src[CY + 1, CX - 1, 3] = v_d
        tx, ty = CX + 2 * tap[0], CY + 2 * tap[1]
        src[ty, tx, :3] = lum_tap  # grey: luminance = the value (the weights sum to 1)
        src[ty, tx, 3] = v_tap
        out = f.atrous(nd, pos, src, 2, phi_c, 128.0, 1.0)
        wl = np.exp(-lum_tap / sigma)
        k = k_tap(*tap)
        assert abs(weight_from_output(out[CY, CX, 0], k, lum_tap) - wl) <= 2e-5 * max(wl, 1e-3), (tap, wl)
        # Output variance: sum (k w)^2 v / (sum k w)^2, with w = 1 except at the tap.
        wsum = 1 - k + k * wl
        vsum = 0.0
        for (dx, dy), v in (((0, 0), v_c), (tap, v_tap)):
            kw = k_tap(dx, dy) * (wl if (dx, dy) == tap else 1.0)
            vsum += kw * kw * v
        # The impulses at (1,0) and (-1,1) are not on the step-2 footprint, so they only enter through sigma.
        assert abs(out[CY, CX, 3] - vsum / wsum ** 2) <= 1e-5 * vsum / wsum ** 2


def test_atrous_variance_prefilter_clamps_at_the_border():
    """At the corner the 3x3 prefilter re-reads the corner pixel for its out-of-image taps: a variance v
    there gives (1/4 + 2/8 + 1/16) v = 0.5625 v, not the renormalised in-image mean (0.444 v)."""
    f = O.SvgfOracle(W, H)
    nd, pos, src = base_planes()
    v = 0.81
    src[0, 0, 3] = v
    src[0, 2, :3] = 0.6
    out = f.atrous(nd, pos, src, 2, 4.0, 128.0, 1.0)
    sigma = 4.0 * np.sqrt(0.5625 * v) + 1e-4
    wl = np.exp(-0.6 / sigma)
    # In-image taps of the corner at step 2: offsets (0..2) x (0..2) of the spline.
    ks = {(dx, dy): k_tap(dx, dy) for dx in range(3) for dy in range(3)}
    want = ks[(1, 0)] * wl * 0.6 / (sum(ks.values()) - ks[(1, 0)] + ks[(1, 0)] * wl)
    assert abs(out[0, 0, 0] - want) <= 1e-5 * want


def _gbuffer(colour, nd, pos):
    color = np.zeros((H, W, 4), np.float32)
    color[..., :3] = colour
    albedo = np.ones((H, W, 4), np.float32)
    emission = np.zeros((H, W, 4), np.float32)
    return color, albedo, nd, pos, emission


def test_variance_pass_bilateral_weights():
    """First frame (history 1): the 7x7 estimate weights each same-material neighbour by
    max(0, n.n')^phi_n exp(-|dz| / (0.05 z + 1e-3)). One lit neighbour with a tilted normal and another
    depth gives its weight back from the estimated colour."""
    f = O.SvgfOracle(W, H)
    for deg, dz, tap in ((8.0, 0.0, (3, 0)), (0.0, 0.4, (-2, 3)), (20.0, 1.1, (1, -1)), (100.0, 0.0, (2, 2))):
        nd = np.zeros((H, W, 4), np.float32)
        nd[..., 2] = 1
        nd[..., 3] = 12.0
        pos = np.zeros((H, W, 4), np.float32)
        tx, ty = CX + tap[0], CY + tap[1]
        th = np.radians(deg)
        nd[ty, tx, :3] = (0.0, np.sin(th), np.cos(th))
        nd[ty, tx, 3] = 12.0 + dz
        colour = np.zeros((H, W, 3), np.float32)
        colour[ty, tx] = 1.0
        f.reset()
        f.run(*_gbuffer(colour, nd, pos), CameraView(), SvgfSettings.make(atrous_iterations=0))
        est = f.plane(6)
        want = max(0.0, np.cos(th)) ** 128 * np.exp(-dz / (0.05 * 12.0 + 1e-3))
        # 49 taps in the image: 48 of weight 1 and the lit one.
        got = np.float64(est[CY, CX, 0])
        w_got = got * 48 / (1 - got)
        assert abs(w_got - want) <= 2e-5 * max(want, 1e-3), (deg, dz, w_got, want)
        # Variance of the single-frame moments (l, l^2) with those weights, times 4 / length (= 4).
        m1 = want / (48 + want)
        assert abs(est[CY, CX, 3] - 4 * max(0.0, m1 - m1 * m1)) <= 1e-5 * max(4 * (m1 - m1 * m1), 1e-6)


# ---- reprojection -------------------------------------------------------------------------------

D, TAN = 16.0, 0.5  # the plane's distance and tan(fov/2): a pixel is 2 D TAN / W world units wide
PIX = 2 * D * TAN / W


def plane_camera(shift_x=0.0, shift_y=0.0):
    c = CameraView()
    c.position[:] = (shift_x, shift_y, 0.0)
    c.forward[:] = (0, 0, -1)
    c.right[:] = (1, 0, 0)
    c.up[:] = (0, 1, 0)
    c.tan_half_fov, c.aspect = TAN, 1.0
    return c


def plane_frame(cam, material=None):
    """Positions, normals and depths of the plane z = -D seen through each pixel centre of `cam`."""
    ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    sx = 2 * (xs + 0.5) / W - 1
    sy = 1 - 2 * (ys + 0.5) / H
    p = np.stack([cam.position[0] + sx * D * TAN, cam.position[1] + sy * D * TAN, np.full((H, W), -D)], -1)
    pos = np.zeros((H, W, 4), np.float32)
    pos[..., :3] = p
    pos[..., 3] = 0.0 if material is None else material
    nd = np.zeros((H, W, 4), np.float32)
    nd[..., 2] = 1
    nd[..., 3] = np.linalg.norm(p - np.array(cam.position[:]), axis=-1)
    return nd, pos


def warm_history(f, image, frames=90):
    """`frames` still frames of the same image. The first three are filtered spatially (history < 4); from
    then on each frame blends the history towards the image by at least 0.2, so after 90 frames the
    history is the image to within 0.8^86 (1e-8) and its length is capped at 64. The variance pass no
    longer filters spatially (length >= 4), so the next frame's temporal output is what it shows."""
    cam = plane_camera()
    nd, pos = plane_frame(cam)
    for _ in range(frames):
        f.run(*_gbuffer(image, nd, pos), cam, SvgfSettings.make(atrous_iterations=0))
    return cam


def lum3(a):
    return 0.2126 * a[..., 0] + 0.7152 * a[..., 1] + 0.0722 * a[..., 2]


@pytest.mark.parametrize("kx,ky", [(1, 0), (0, -2), (3, 1)])
def test_whole_pixel_translation_reads_the_shifted_history(kx, ky):
    rng = np.random.default_rng(kx * 7 + ky + 10)
    img = rng.random((H, W, 3)).astype(np.float32)
    f = O.SvgfOracle(W, H)
    prev = warm_history(f, img)
    cam = plane_camera(kx * PIX, ky * PIX)
    nd, pos = plane_frame(cam)
    cur = rng.random((H, W, 3)).astype(np.float32)
    f.run(*_gbuffer(cur, nd, pos), prev, SvgfSettings.make(atrous_iterations=0))
    illum = f.plane(4)
    # Moving the camera right by k pixels moves the image left: pixel x now sees what x + k saw; up by k
    # moves it down: pixel y sees what y - k saw.
    ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    sx, sy = xs + kx, ys - ky
    inside = (sx >= 0) & (sx < W) & (sy >= 0) & (sy < H)
    hist = np.zeros_like(img)
    hist[inside] = img[sy[inside], sx[inside]]
    a = 0.2  # max(alpha, 1 / 64)
    want = hist + (cur - hist) * a
    np.testing.assert_allclose(illum[..., :3][inside], want[inside], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(illum[..., 3][inside], 64.0, rtol=1e-6)
    # Off-image history: a fresh start.
    assert np.array_equal(illum[..., :3][~inside], cur[~inside]) and np.all(illum[..., 3][~inside] == 1)


def test_half_pixel_translation_averages_four_taps():
    rng = np.random.default_rng(5)
    img = rng.random((H, W, 3)).astype(np.float32)
    f = O.SvgfOracle(W, H)
    prev = warm_history(f, img)
    cam = plane_camera(0.5 * PIX, 0.5 * PIX)
    nd, pos = plane_frame(cam)
    cur = rng.random((H, W, 3)).astype(np.float32)
    f.run(*_gbuffer(cur, nd, pos), prev, SvgfSettings.make(atrous_iterations=0))
    illum, moments = f.plane(4), f.plane(5)
    # The previous pixel coordinates are (x + 1/2, y - 1/2): taps x, x+1 and y-1, y, each of weight 1/4;
    # taps outside the image are skipped and the rest renormalised.
    hist = np.zeros((H, W, 3))
    wsum = np.zeros((H, W))
    hm = np.zeros((H, W, 2))
    lum_img = lum3(img.astype(np.float64))
    for ox, oy in ((0, -1), (1, -1), (0, 0), (1, 0)):
        ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
        tx, ty = xs + ox, ys + oy
        ok = (tx >= 0) & (tx < W) & (ty >= 0) & (ty < H)
        hist[ok] += 0.25 * img[ty[ok], tx[ok]]
        hm[ok, 0] += 0.25 * lum_img[ty[ok], tx[ok]]
        hm[ok, 1] += 0.25 * lum_img[ty[ok], tx[ok]] ** 2
        wsum[ok] += 0.25
    hist /= wsum[..., None]
    hm /= wsum[..., None]
    a = 0.2  # max(alpha, 1 / 64)
    want = hist + (cur - hist) * a
    np.testing.assert_allclose(illum[..., :3], want, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(illum[..., 3], 64.0, rtol=1e-6)
    lc = lum3(cur.astype(np.float64))
    np.testing.assert_allclose(moments[..., 0], hm[..., 0] + (lc - hm[..., 0]) * a, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(moments[..., 1], hm[..., 1] + (lc * lc - hm[..., 1]) * a, rtol=1e-5, atol=1e-6)
    # The moments of this synthetic history are those of single frames (l, l^2), so the blended second
    # moment minus the squared first is the variance the pass reports.
    np.testing.assert_allclose(moments[..., 2], np.maximum(0, moments[..., 1] - moments[..., 0] ** 2), rtol=1e-4, atol=1e-6)


def test_reprojection_rejects_changed_material_distance_and_normal():
    """A still camera: every pixel reprojects onto its own centre (one tap). A tap is dropped when the
    material differs, when the distance moved by more than 5 % + 0.01 (4 % is kept), or when the normals'
    cosine falls below 0.9 (0.91 is kept); a dropped history restarts at length 1 with the new value."""
    rng = np.random.default_rng(6)
    img = rng.random((H, W, 3)).astype(np.float32)
    f = O.SvgfOracle(W, H)
    cam = warm_history(f, img)
    nd, pos = plane_frame(cam)
    cur = rng.random((H, W, 3)).astype(np.float32)
    cases = {}
    # Material changed.
    pos[2, 3, 3] = 7.0
    cases[(2, 3)] = False
    # Surface moved along the pixel's ray: 6 % farther is rejected, 4 % closer kept.
    for (y, x), scale, keep in (((5, 5), 1.06, False), ((5, 9), 0.96, True), ((9, 5), 1.048, True)):
        c = np.array(cam.position[:])
        pos[y, x, :3] = c + (pos[y, x, :3] - c) * scale
        nd[y, x, 3] = nd[y, x, 3] * scale
        cases[(y, x)] = keep
    # Normal tilted: cos 0.89 rejected, 0.91 kept.
    for (y, x), cosv, keep in (((12, 12), 0.89, False), ((12, 20), 0.91, True)):
        nd[y, x, :3] = (np.sqrt(1 - cosv * cosv), 0.0, cosv)
        cases[(y, x)] = keep
    f.run(*_gbuffer(cur, nd, pos), cam, SvgfSettings.make(atrous_iterations=0))
    illum = f.plane(4)
    a = 0.2  # max(alpha, 1 / 64)
    for (y, x), keep in cases.items():
        if keep:
            np.testing.assert_allclose(illum[y, x, :3], img[y, x] + (cur[y, x] - img[y, x]) * a, rtol=1e-5)
            assert abs(illum[y, x, 3] - 64) < 1e-4, (y, x)
        else:
            assert np.array_equal(illum[y, x, :3], cur[y, x]) and illum[y, x, 3] == 1, (y, x)
    # Every untouched pixel kept its history.
    mask = np.ones((H, W), bool)
    for yx in cases:
        mask[yx] = False
    np.testing.assert_allclose(illum[..., 3][mask], 64.0, rtol=1e-6)
