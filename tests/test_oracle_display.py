"""The oracle's display_kernel (every mode and tone map, gpu_renderer.cu:108-175 with color.h:8-41)
pinned by special cases the definitions fix, each chosen so that a plausible slip fails it:

* sRGB encoding (IEC 61966-2-1): code values of known linear inputs, on both sides of the linear toe;
* tone maps: Reinhard x/(1+x) at x = 1 and 3, exposure scaling, the clamp of the linear map, and Hill's
  ACES fit, which for grey input reduces to the scalar rational function (x(x+a)-b)/(x(cx+d)+e) because
  both matrices have unit row sums (a transposed matrix breaks grey-in/grey-out);
* debug views (normal, depth, variance, heat map) are display values, decoded before packing, so their
  code values are round(255 v): a missing decode shows as the sRGB curve;
* the variance view of a progressive image is the variance of the mean of the samples (numpy's
  var(ddof=1)/n), and of a denoised frame the denoiser's variance plane.
"""
import numpy as np
import pytest

from oracle import oracle_py as O

COLOR, ALBEDO, NORMAL, DEPTH, VARIANCE, HEAT = range(6)
LINEAR, ACES, REINHARD = range(3)


def codes(px):
    px = np.asarray(px, np.uint32)
    return np.stack([px & 255, (px >> 8) & 255, (px >> 16) & 255], -1).astype(np.int64)


def srgb_code(x):
    """The sRGB transfer function's code value of linear x in [0, 1] (IEC 61966-2-1), in double."""
    x = np.clip(np.asarray(x, np.float64), 0, 1)
    v = np.where(x <= 0.0031308, 12.92 * x, 1.055 * x ** (1 / 2.4) - 0.055)
    return np.floor(v * 255 + 0.5).astype(np.int64)


def colour(rgb, **kw):
    rgb = np.asarray(rgb, np.float32).reshape(-1, 3)
    accum = np.concatenate([rgb, np.zeros((len(rgb), 1), np.float32)], 1)
    return codes(O.display(COLOR, len(rgb), accum=accum, **kw))


def test_srgb_codes_of_known_values():
    # 0.5 -> 0.7354 (188), 0.18 -> 0.4614 (118), 0.001 is on the linear toe (12.92 x -> 3), 1 -> 255, 0 -> 0.
    got = colour([[0.5, 0.18, 0.001], [1.0, 0.0, 2.0]])
    assert got.tolist() == [[188, 118, 3], [255, 0, 255]]
    xs = np.linspace(0, 1, 997, dtype=np.float32)
    got = colour(np.stack([xs, xs, xs], 1))[:, 0]
    want = srgb_code(xs)
    assert np.abs(got - want).max() <= 1 and np.mean(got == want) > 0.99


def test_progressive_colour_divides_by_samples_and_applies_exposure():
    # A sum of 4 samples of 0.25 each at exposure 2 shows as 0.5.
    assert colour([[1.0, 1.0, 1.0]], samples=4, exposure=2.0)[0].tolist() == [188] * 3
    # The denoised frame is shown as it is, without the division by samples.
    den = np.array([[0.5, 0.5, 0.5, 0]], np.float32)
    out = codes(O.display(COLOR, 1, samples=7, denoised=True, denoised_rgb=den, accum=np.full((1, 4), 9.0, np.float32)))
    assert out[0].tolist() == [188] * 3


def test_reinhard_and_linear_tone_maps():
    # Reinhard: x/(1+x); 1 -> 0.5 (188), 3 -> 0.75 (225); linear clamps at 1.
    assert colour([[1.0, 3.0, 0.0]], tonemap=REINHARD)[0].tolist() == [188, 225, 0]
    assert colour([[4.0, 0.5, -1.0]], tonemap=LINEAR)[0].tolist() == [255, 188, 0]


def test_aces_fit_on_grey_reduces_to_its_scalar_curve():
    xs = np.array([0.0, 0.05, 0.18, 0.5, 1.0, 2.0, 8.0, 1e6], np.float32)
    got = colour(np.stack([xs, xs, xs], 1), tonemap=ACES)
    assert np.all(got[:, 0] == got[:, 1]) and np.all(got[:, 1] == got[:, 2])  # grey in, grey out
    x = xs.astype(np.float64)
    f = (x * (x + 0.0245786) - 0.000090537) / (x * (0.983729 * x + 0.4329510) + 0.238081)
    want = srgb_code(f)
    assert np.abs(got[:, 0] - want).max() <= 1, (got[:, 0], want)
    assert got[0, 0] == 0 and got[-1, 0] == 255
    # A saturated red stays red-dominant, and the channels differ (the matrices mix them).
    r = colour([[4.0, 0.0, 0.0]], tonemap=ACES)[0]
    assert r[0] > r[1] and r[0] > r[2] and r[1] != r[2]


def test_albedo_view_is_srgb_encoded():
    alb = np.array([[0.5, 0.18, 1.0, 0.0]], np.float32)
    assert codes(O.display(ALBEDO, 1, albedo=alb))[0].tolist() == [188, 118, 255]


def test_normal_view_shows_half_offset_normals_and_black_misses():
    nd = np.array([[1, 0, 0, 2.0], [0, -1, 0, 0.5], [0.3, 0.3, 0.3, 0.0]], np.float32)
    got = codes(O.display(NORMAL, 3, nd=nd))
    assert got.tolist() == [[255, 128, 128], [128, 0, 128], [0, 0, 0]]


def test_depth_view_is_one_minus_relative_depth():
    nd = np.zeros((4, 4), np.float32)
    nd[:, 3] = [2.5, 10.0, 20.0, 0.0]  # 1 - d/10: 0.75, 0, 0 (clamped), miss
    got = codes(O.display(DEPTH, 4, nd=nd, max_depth=10.0))[:, 0]
    assert got.tolist() == [191, 0, 0, 0]


def test_variance_view_of_a_progressive_image():
    rng = np.random.default_rng(3)
    for n in (2, 4, 16):
        lum = rng.uniform(0, 3, n).astype(np.float32)
        accum = np.array([[lum.sum(), lum.sum(), lum.sum(), np.sum(lum.astype(np.float32) ** 2)]], np.float32)
        got = codes(O.display(VARIANCE, 1, samples=n, accum=accum))[0, 0]
        var = np.var(lum.astype(np.float64), ddof=1) / n
        t = np.sqrt(var)
        assert abs(got - np.floor(255 * t / (1 + t) + 0.5)) <= 1, (n, got)
    # One sample has no variance estimate.
    assert codes(O.display(VARIANCE, 1, samples=1, accum=np.array([[5, 5, 5, 25]], np.float32)))[0, 0] == 0


def test_variance_view_of_a_denoised_frame_reads_the_variance_plane():
    var = np.array([[9, 9, 9, 1.0], [0, 0, 0, 0.0]], np.float32)  # sqrt(1) -> 1/(1+1) = 0.5 -> 128
    got = codes(O.display(VARIANCE, 2, samples=1, denoised=True, variance=var))[:, 0]
    assert got.tolist() == [128, 0]


@pytest.mark.parametrize("t,want", [(0.0, [13, 13, 89]), (1.0, [217, 26, 13]), (0.7, [240, 198, 24]), (0.5, [121, 185, 64]),
                                    (1 / 6, [6, 83, 96]), (2.0, [217, 26, 13])])
def test_heat_map_ramp(t, want):
    # The ramp's stops at 0, 1/3, 2/3 and 1 are (0.05,0.05,0.35), (0,0.6,0.4), (0.95,0.85,0.1), (0.85,0.1,0.05);
    # t = 1/2 is halfway from the second to the third stop, 1/6 halfway from the first to the second, and
    # 0.7 is a tenth of the way from the third to the fourth.
    got = codes(O.display(HEAT, 1, heat=np.array([t * 160.0], np.float32), heat_scale=1.0 / 160.0))[0]
    assert np.abs(got - np.array(want)).max() <= 1, got
