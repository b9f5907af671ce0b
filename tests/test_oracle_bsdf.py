"""BSDF pins (the reference's test_bsdf.cpp): the sampled directions follow `pdf` (chi-square over the
sphere), `pdf` integrates to the fraction of valid samples, white-furnace energy bounds per material,
reciprocity, and finiteness at degenerate angles."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_py as O

FP = C.POINTER(C.c_float)


def P(a):
    return np.ascontiguousarray(a, np.float32).ctypes.data_as(FP)


CASES = {
    "diffuse": ((0.8, 0.8, 0.8), 0, 1, 0),
    "plastic rough": ((0.5, 0.2, 0.1), 0, 0.5, 0),
    "plastic glossy": ((0.5, 0.2, 0.1), 0, 0.25, 0),
    "metal rough": ((0.95, 0.7, 0.3), 1, 0.6, 0),
    "metal glossy": ((0.95, 0.7, 0.3), 1, 0.3, 0),
    "half metal": ((0.6, 0.6, 0.6), 0.5, 0.4, 0),
    "rough glass": ((1, 1, 1), 0, 0.5, 1),
    "frosted glass": ((1, 1, 1), 0, 0.3, 1),
    "semi transmissive": ((0.7, 0.7, 0.7), 0, 0.6, 0.5),
}


def params(base, metallic, roughness, transmission, ior=1.5):
    return np.array([*base, metallic, roughness, transmission, ior], np.float32)


def direction(ct, phi):
    s = np.sqrt(np.maximum(0, 1 - ct * ct))
    return np.stack([s * np.cos(phi), s * np.sin(phi), ct], -1).astype(np.float32)


def sample(p, wo, u):
    n = len(u)
    wos = np.ascontiguousarray(np.broadcast_to(wo, (n, 3)), np.float32)
    wi, f, pdf, flags = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros(n, np.int32)
    O.oracle().or_bsdf_sample(P(p), n, P(wos), P(u), P(wi), P(f), P(pdf), flags.ctypes.data)
    return wi, f, pdf, flags


def evaluate(p, wo, wi):
    n = len(wi)
    wos = np.ascontiguousarray(np.broadcast_to(wo, (n, 3)), np.float32)
    f, pdf = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
    O.oracle().or_bsdf_eval(P(p), n, P(wos), P(np.ascontiguousarray(wi, np.float32)), P(f), P(pdf))
    return f, pdf


def chi_square(p, wo, n=400000, nt=20, nph=24):
    rng = np.random.default_rng(1234)
    wi, f, pdf, flags = sample(p, wo, rng.random((n, 4), dtype=np.float32))
    valid = (flags & 1) == 1
    w = wi[valid] / np.linalg.norm(wi[valid], axis=1, keepdims=True)
    phi = np.mod(np.arctan2(w[:, 1], w[:, 0]), 2 * np.pi)
    it = np.minimum(nt - 1, ((w[:, 2] + 1) * 0.5 * nt).astype(int))
    ip = np.minimum(nph - 1, (phi / (2 * np.pi) * nph).astype(int))
    observed = np.bincount(it * nph + ip, minlength=nt * nph).astype(float)

    def integral(sub):
        a = (np.arange(sub) + 0.5) / sub
        ct = -1 + 2 * (np.arange(nt)[:, None] + a[None, :]) / nt  # (nt, sub)
        ph = 2 * np.pi * (np.arange(nph)[:, None] + a[None, :]) / nph  # (nph, sub)
        CT = np.broadcast_to(ct[:, None, :, None], (nt, nph, sub, sub))
        PH = np.broadcast_to(ph[None, :, None, :], (nt, nph, sub, sub))
        _, pd = evaluate(p, wo, direction(CT.reshape(-1), PH.reshape(-1)))
        return pd.reshape(nt, nph, sub * sub).mean(-1) * (2.0 / nt) * (2 * np.pi / nph)

    coarse, fine = integral(8), integral(16)
    expected = fine.copy()
    bad = np.abs(coarse - fine) > 0.01 * fine + 1e-7
    if bad.any():
        refined = integral(64)
        expected[bad] = refined[bad]
    expected = expected.reshape(-1) * n
    small = expected < 5
    chi2 = np.sum((observed[~small] - expected[~small]) ** 2 / expected[~small])
    dof = int((~small).sum())
    pe, po = expected[small].sum(), observed[small].sum()
    if pe >= 5:
        chi2 += (po - pe) ** 2 / pe
        dof += 1
    return chi2, dof, expected.sum() / n, valid.mean()


@pytest.mark.parametrize("name", list(CASES))
def test_sampling_matches_pdf(name):
    base, m, r, t = CASES[name]
    p = params(base, m, r, t)
    for ct in (0.9, 0.5, 0.15, -0.6):
        if ct < 0 and t <= 0:
            continue
        wo = direction(np.float32(ct), np.float32(0.7))
        chi2, dof, integral, valid = chi_square(p, wo)
        bound = dof + 6 * np.sqrt(2.0 * dof) + 20
        assert chi2 < bound, (name, ct, chi2, bound)
        assert abs(integral - valid) < 0.02, (name, ct, integral, valid)


def radiance_scale(transmitted, wo_z, ior):
    return np.where(transmitted, np.where(wo_z > 0, ior * ior, 1.0 / (ior * ior)), 1.0)


@pytest.mark.parametrize("name", list(CASES))
def test_energy_conservation(name):
    base, m, r, t = CASES[name]
    p = params((1, 1, 1), m, r, t)
    rng = np.random.default_rng(99)
    for ct in (1.0, 0.7, 0.3, 0.05):
        wo = direction(np.float32(ct), np.float32(1.3))
        wi, f, pdf, flags = sample(p, wo, rng.random((200000, 4), dtype=np.float32))
        valid = (flags & 1) == 1
        lum = 0.2126 * f[:, 0] + 0.7152 * f[:, 1] + 0.0722 * f[:, 2]
        scale = radiance_scale((flags & 2) == 2, wo[2], 1.5)
        e = np.sum(np.where(valid, lum * np.abs(wi[:, 2]) / np.where(valid, pdf, 1) * scale, 0)) / len(f)
        assert e <= 1.02, (name, ct, e)
        assert e >= (0.75 if ct > 0.2 else 0.3), (name, ct, e)


def test_smooth_mirror_and_glass_conserve_energy_exactly():
    rng = np.random.default_rng(7)
    for t in (0.0, 1.0):
        p = params((1, 1, 1), 1 - t, 0.0, t)
        for ct in (0.9, 0.4, -0.8):
            if ct < 0 and t == 0:
                continue
            wo = direction(np.float32(ct), np.float32(0.2))
            wi, f, pdf, flags = sample(p, wo, rng.random((100000, 4), dtype=np.float32))
            valid = (flags & 1) == 1
            lum = 0.2126 * f[:, 0] + 0.7152 * f[:, 1] + 0.0722 * f[:, 2]
            scale = radiance_scale((flags & 2) == 2, wo[2], 1.5)
            e = np.sum(np.where(valid, lum * np.abs(wi[:, 2]) / np.where(valid, pdf, 1) * scale, 0)) / len(f)
            assert abs(e - 1) < 0.03, (t, ct, e)


def test_reflection_is_reciprocal():
    rng = np.random.default_rng(5)
    for name, (base, m, r, t) in CASES.items():
        if t > 0:
            continue
        p = params(base, m, r, t)
        u = rng.random((4000, 2), dtype=np.float32)
        a = np.zeros((2000, 3), np.float32)
        d = np.zeros((2000, 3), np.float32)
        O.oracle().or_warp(1, 2000, P(u[:2000]), P(a))
        O.oracle().or_warp(1, 2000, P(u[2000:]), P(d))
        keep = (a[:, 2] >= 0.05) & (d[:, 2] >= 0.05)
        a, d = a[keep], d[keep]
        f1 = np.stack([evaluate(p, a[i], d[i:i + 1])[0][0] for i in range(len(a))])
        f2 = np.stack([evaluate(p, d[i], a[i:i + 1])[0][0] for i in range(len(a))])
        mx = np.maximum(f1.max(1), 1e-3)
        assert (np.abs(f1 - f2).max(1) <= 1e-3 * mx + 1e-5).all(), name


def test_values_are_finite_at_degenerate_angles():
    rng = np.random.default_rng(8)
    for name, (base, m, r, t) in CASES.items():
        p = params(base, m, r, t)
        n = 20000
        wo = np.zeros((n, 3), np.float32)
        O.oracle().or_warp(2, n, P(rng.random((n, 2), dtype=np.float32)), P(wo))
        wo[::100] = [0, 0, 1]
        wo[1::100] = [1 / np.sqrt(1 + 1e-14), 0, 1e-7 / np.sqrt(1 + 1e-14)]
        wi, f, pdf, flags = (np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros(n, np.int32))
        O.oracle().or_bsdf_sample(P(p), n, P(wo), P(rng.random((n, 4), dtype=np.float32)), P(wi), P(f), P(pdf), flags.ctypes.data)
        v = (flags & 1) == 1
        assert np.isfinite(f[v]).all() and np.isfinite(pdf[v]).all() and (pdf[v] > 0).all() and (f[v] >= 0).all(), name
        wj = np.zeros((n, 3), np.float32)
        O.oracle().or_warp(2, n, P(rng.random((n, 2), dtype=np.float32)), P(wj))
        fe, pe = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        O.oracle().or_bsdf_eval(P(p), n, P(wo), P(wj), P(fe), P(pe))
        assert np.isfinite(fe).all() and np.isfinite(pe).all() and (pe >= 0).all(), name
