"""Light-sampling pins (the reference's test_lights.cpp): environment pdf self-consistency, E[1/pdf] = 4 pi,
E[L/pdf] against a quadrature, emissive-triangle density equal to the hit-point density, NEE against
hemisphere irradiance, and the inverse-square law of point lights."""
import ctypes as C

import numpy as np

from oracle import oracle_py as O
from photon_amd import host
from photon_amd.types import LIGHT_POINT, Material, PunctualLight, RAY_DTYPE

FP = C.POINTER(C.c_float)


def P(a):
    return a.ctypes.data_as(FP)


def env_scene(image):
    s = host.Scene()
    m = Material.default()
    m.metallic = 0
    s.add_mesh(np.array([[-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1]], np.float32), np.array([0, 1, 2, 0, 2, 3], np.int32), s.add_material(m))
    if image:
        sun = np.array([0.5, 0.6, -0.3])
        s.set_sky(sun / np.linalg.norm(sun), 3000.0, 256)
        # rotation 0.7 as in the reference test: set through the image path
        raw = s.raw()
        w, h = raw.env_width, raw.env_height
        tex = np.ctypeslib.as_array(raw.env_texels, (h * w * 3,)).copy().reshape(h, w, 3)
        s.set_env_image(tex, 1.0, 0.7)
    else:
        s.set_env_color((0.5, 0.7, 1.0))
    return s.build()


def test_environment_sampling_density_is_consistent():
    for image in (False, True):
        s = env_scene(image)
        d = s.desc()
        n = 400000
        rng = np.random.default_rng(3)
        u = rng.random((n, 2), dtype=np.float32)
        dirs, pdf, ok = np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros(n, np.int32)
        O.oracle().or_env_sample(C.byref(d), n, P(u), P(dirs), P(pdf), ok.ctypes.data)
        g = ok == 1
        rad, pe = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        O.oracle().or_env_eval(C.byref(d), n, P(dirs), P(rad), P(pe))
        mism = np.sum(np.abs(pe[g] - pdf[g]) > 2e-3 * pdf[g] + 1e-6)
        assert mism < n / 1000
        assert abs(np.sum(1.0 / pdf[g]) / n - 4 * np.pi) < 4 * np.pi * 0.01
        lum = 0.2126 * rad[:, 0] + 0.7152 * rad[:, 1] + 0.0722 * rad[:, 2]
        est = np.sum(lum[g] / pdf[g]) / n
        gt, gp = 512, 1024
        yy, xx = np.meshgrid((np.arange(gt) + 0.5) / gt, (np.arange(gp) + 0.5) / gp, indexing="ij")
        uv = np.stack([xx.reshape(-1), yy.reshape(-1)], 1).astype(np.float32)
        gd = np.zeros((len(uv), 3), np.float32)
        O.oracle().or_warp(5, len(uv), P(uv), P(gd))
        gr = np.zeros_like(gd)
        O.oracle().or_env_eval(C.byref(d), len(uv), P(gd), P(gr), None)
        gl = 0.2126 * gr[:, 0] + 0.7152 * gr[:, 1] + 0.0722 * gr[:, 2]
        ref = np.sum(gl * np.sin(yy.reshape(-1) * np.pi)) * (np.pi / gt) * (2 * np.pi / gp)
        assert abs(est - ref) < 0.02 * ref, (image, est, ref)


def test_emissive_density_matches_hit_density():
    s = host.Scene.builtin("cornell").build()
    d = s.desc()
    rng = np.random.default_rng(4)
    n = 20000
    p = np.stack([0.5 + 4.5 * rng.random(n), 0.01 + 3 * rng.random(n), 0.2 + 5 * rng.random(n)], 1).astype(np.float32)
    u = rng.random((n, 4), dtype=np.float32)
    out = np.zeros((n, 12), np.float32)
    O.oracle().or_sample_light(C.byref(d), n, P(p), P(u), P(out))
    wi, dist, pdf, delta, tri = out[:, 0:3], out[:, 3], out[:, 7], out[:, 8], out[:, 9].astype(np.int32)
    keep = (pdf > 0) & (delta == 0)
    rays = np.zeros(int(keep.sum()), RAY_DTYPE)
    rays["o"], rays["d"], rays["tmax"] = p[keep], wi[keep], 1e30
    hits = O.trace(d, rays)
    vis = hits["tri"] == tri[keep]
    assert vis.sum() > 5000
    assert np.all(np.abs(hits["t"][vis] - dist[keep][vis]) <= 1e-3 * dist[keep][vis])
    ep = np.zeros(int(vis.sum()), np.float32)
    tri_v = np.ascontiguousarray(hits["tri"][vis])
    wi_v = np.ascontiguousarray(wi[keep][vis])
    t_v = np.ascontiguousarray(hits["t"][vis])
    O.oracle().or_emissive_pdf(C.byref(d), len(ep), tri_v.ctypes.data, P(wi_v), P(t_v), P(ep))
    assert np.all(np.abs(ep - pdf[keep][vis]) <= 1e-3 * pdf[keep][vis])


def test_light_sampling_and_hemisphere_sampling_agree():
    s = host.Scene.builtin("cornell").build()
    d = s.desc()
    for pt in ([2.7, 0, 2.8], [0.6, 0, 4.9], [4.5, 0, 0.8]):
        n = 400000
        rng = np.random.default_rng(7)
        p = np.tile(np.array(pt, np.float32), (n, 1))
        o = np.zeros((1, 3), np.float32)
        O.oracle().or_offset_ray(1, P(np.array([pt], np.float32)), P(np.array([[0, 1, 0]], np.float32)), P(o))
        out = np.zeros((n, 12), np.float32)
        O.oracle().or_sample_light(C.byref(d), n, P(p), P(rng.random((n, 4), dtype=np.float32)), P(out))
        wi, dist, rad, pdf = out[:, 0:3], out[:, 3], out[:, 4:7], out[:, 7]
        use = (pdf > 0) & (wi[:, 1] > 0)
        sh = np.zeros(int(use.sum()), RAY_DTYPE)
        sh["o"], sh["d"], sh["tmax"] = o[0], wi[use], dist[use] * np.float32(1 - 1e-4)
        occ = O.trace(d, sh, anyhit=True)["tri"] >= 0
        lum = 0.2126 * rad[use, 0] + 0.7152 * rad[use, 1] + 0.0722 * rad[use, 2]
        nee = np.sum(np.where(occ, 0, lum * wi[use, 1] / pdf[use])) / n
        # The light covers ~0.5% of the hemisphere: 2M cosine samples keep the estimator's error near 1%.
        nh = 2000000
        hd = np.zeros((nh, 3), np.float32)
        O.oracle().or_warp(1, nh, P(rng.random((nh, 2), dtype=np.float32)), P(hd))
        hemi_dirs = np.stack([hd[:, 0], hd[:, 2], -hd[:, 1]], 1)  # local z -> world +y (any ONB around +y)
        rays = np.zeros(nh, RAY_DTYPE)
        rays["o"], rays["d"], rays["tmax"] = o[0], hemi_dirs, 1e30
        hits = O.trace(d, rays)
        em = np.zeros((nh, 3), np.float32)
        O.oracle().or_hit_emission(C.byref(d), nh, rays.ctypes.data, hits.ctypes.data, P(em))
        hemi = np.sum((0.2126 * em[:, 0] + 0.7152 * em[:, 1] + 0.0722 * em[:, 2]).astype(np.float64)) * np.pi / nh
        assert abs(nee - hemi) < 0.03 * hemi + 1e-3, (pt, nee, hemi)


def test_point_light_inverse_square():
    s = host.Scene()
    m = Material.default()
    m.metallic = 0
    s.add_mesh(np.array([[-10, 0, -10], [-10, 0, 10], [10, 0, 10], [10, 0, -10]], np.float32), np.array([0, 1, 2, 0, 2, 3], np.int32),
               s.add_material(m))
    light = PunctualLight()
    light.type = LIGHT_POINT
    light.position[:] = (0, 2, 0)
    light.intensity[:] = (8, 8, 8)
    s.add_punctual(light)
    s.build()
    d = s.desc()
    out = np.zeros((2, 12), np.float32)
    p = np.array([[0, 0, 0], [1.5, 0, 0]], np.float32)
    u = np.array([[0.5, 0.3, 0.3, 0.1], [0.5, 0.3, 0.3, 0.1]], np.float32)
    O.oracle().or_sample_light(C.byref(d), 2, P(p), P(u), P(out))
    lum = 0.2126 * out[:, 4] + 0.7152 * out[:, 5] + 0.0722 * out[:, 6]
    assert out[0, 8] == 1.0 and abs(out[0, 3] - 2) < 1e-5 and abs(out[0, 7] - 1.0) < 1e-6
    assert abs(lum[0] - 2.0) < 1e-4 and abs(lum[1] - 8.0 / 6.25) < 1e-4
