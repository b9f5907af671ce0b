"""Pins of the oracle's sampling arithmetic by known answers, published inverses, closed forms and
round trips (the reference's test_math.cpp checks, plus the published PCG32 and hash vectors)."""
import ctypes as C

import numpy as np

from oracle import oracle_py as O

FP = C.POINTER(C.c_float)


def f32p(a):
    return a.ctypes.data_as(FP)


def test_pcg32_known_answer():
    # pcg32-demo (O'Neill, pcg-c-basic): pcg32_srandom_r(&rng, 42u, 54u) produces these first outputs.
    # Rng(seed, stream) has exactly the structure of pcg32_srandom_r(initstate=seed, initseq=stream).
    out = np.zeros(6, np.uint32)
    O.oracle().or_rng_u32(42, 54, 6, out.ctypes.data)
    assert [hex(v) for v in out] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b", "0xcbed606e"]


def test_hash_is_lowbias32_by_its_published_inverse():
    # lowbias32 (C. Wellons, hash-prospector) and its published inverse; a wrong constant or shift in the
    # forward hash breaks the round trip.
    rng = np.random.default_rng(1)
    x = rng.integers(0, 2**32, 100000, dtype=np.uint64).astype(np.uint32)
    x[:3] = [0, 1, 0xFFFFFFFF]
    h = np.zeros_like(x)
    O.oracle().or_hash_u32(x.ctypes.data, len(x), h.ctypes.data)
    y = h.astype(np.uint64)
    M = np.uint64(0xFFFFFFFF)
    y ^= y >> np.uint64(16)
    y = (y * np.uint64(0x43021123)) & M
    y ^= (y >> np.uint64(15)) ^ (y >> np.uint64(30))
    y = (y * np.uint64(0x1d69e2a5)) & M
    y ^= y >> np.uint64(16)
    assert np.array_equal(y.astype(np.uint32), x)
    assert h[0] == 0  # lowbias32(0) = 0


def test_uniform_floats_are_in_range_and_uniform():
    n = 160000
    u = np.zeros(n, np.float32)
    O.oracle().or_rng_float(42, 7, n, f32p(u))
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.005
    hist = np.bincount((u * 16).astype(int), minlength=16)
    chi2 = ((hist - n / 16) ** 2 / (n / 16)).sum()
    assert chi2 < 40
    a, b = np.zeros(100, np.uint32), np.zeros(100, np.uint32)
    O.oracle().or_rng_u32(1, 1, 100, a.ctypes.data)
    O.oracle().or_rng_u32(1, 2, 100, b.ctypes.data)
    assert np.sum(a == b) < 3


def _warp(kind, inp, out_dim):
    inp = np.ascontiguousarray(inp, np.float32)
    n = inp.shape[0]
    out = np.zeros((n, out_dim), np.float32)
    assert O.oracle().or_warp(kind, n, f32p(inp), f32p(out)) == 0
    return out


def test_warp_densities_integrate_to_the_domain_measure():
    rng = np.random.default_rng(3)
    u = rng.random((200000, 2), dtype=np.float32)
    d = _warp(1, u, 3)
    assert (d[:, 2] >= 0).all()
    z = d[:, 2][d[:, 2] > 1e-4]
    assert abs(np.sum(np.pi / z) / len(u) - 2 * np.pi) < 0.05  # E[1/pdf] = 2 pi for cosine sampling
    b = _warp(3, u, 2)
    assert (b >= 0).all() and (b.sum(1) <= 1.0001).all()
    assert abs(b[:, 0].mean() - 1 / 3) < 0.005 and abs(b[:, 1].mean() - 1 / 3) < 0.005
    s = _warp(2, u, 3)
    assert np.allclose(np.linalg.norm(s, axis=1), 1, atol=1e-5)
    assert np.abs(s.mean(0)).max() < 0.01
    disk = _warp(0, u, 2)
    assert (np.linalg.norm(disk, axis=1) <= 1 + 1e-6).all()
    assert abs(np.mean(np.sum(disk ** 2, 1)) - 0.5) < 0.005  # uniform disk: E[r^2] = 1/2


def test_concentric_map_special_cases():
    # The centre maps to the origin; the mid-edges map to the unit circle's axis points.
    out = _warp(0, np.array([[0.5, 0.5], [1.0, 0.5], [0.5, 1.0], [0.0, 0.5]], np.float32), 2)
    assert np.allclose(out, [[0, 0], [1, 0], [0, 1], [-1, 0]], atol=1e-6)


def test_equirect_round_trip_and_poles():
    rng = np.random.default_rng(5)
    d = _warp(2, rng.random((10000, 2), dtype=np.float32), 3)
    uv = _warp(4, d, 2)
    assert (uv >= 0).all() and (uv <= 1).all()
    back = _warp(5, uv, 3)
    assert np.abs(back - d).max() < 1e-4
    top = _warp(4, np.array([[0, 1, 0]], np.float32), 2)
    assert abs(top[0, 1]) < 1e-6  # +y is the top row
    # -z is the image centre (u = 0.5)
    c = _warp(4, np.array([[0, 0, -1]], np.float32), 2)
    assert abs(c[0, 0] - 0.5) < 1e-6 and abs(c[0, 1] - 0.5) < 1e-6


def test_frame_is_orthonormal():
    rng = np.random.default_rng(1)
    n = _warp(2, rng.random((10000, 2), dtype=np.float32), 3)
    n[0] = [0, 0, -1]
    n[1] = [0, 0, 1]
    v = _warp(2, rng.random((10000, 2), dtype=np.float32), 3)
    local, world, basis = (np.zeros((10000, 3), np.float32) for _ in range(3))
    basis = np.zeros((10000, 9), np.float32)
    O.oracle().or_frame(10000, f32p(n), f32p(v), f32p(local), f32p(world), f32p(basis))
    t, b, nn = basis[:, 0:3], basis[:, 3:6], basis[:, 6:9]
    for x, y in ((t, b), (t, nn), (b, nn)):
        assert np.abs(np.sum(x * y, 1)).max() < 1e-5
    for x in (t, b):
        assert np.abs(np.linalg.norm(x, axis=1) - 1).max() < 1e-5
    assert np.abs(world - v).max() < 1e-5
    # The normal maps to +z.
    lz = np.zeros((10000, 3), np.float32)
    O.oracle().or_frame(10000, f32p(n), f32p(n), f32p(lz), None, None)
    assert np.abs(lz[:, 2] - 1).max() < 1e-5


def test_offset_ray_leaves_the_surface():
    rng = np.random.default_rng(11)
    n = 10000
    scale = 10.0 ** (rng.random(n) * 8 - 4)
    p = ((rng.random((n, 3)) * 2 - 1) * scale[:, None]).astype(np.float32)
    nrm = _warp(2, rng.random((n, 2), dtype=np.float32), 3)
    q = np.zeros_like(p)
    O.oracle().or_offset_ray(n, f32p(p), f32p(nrm), f32p(q))
    d = (q - p).astype(np.float64)
    assert (np.sum(d * nrm, 1) > 0).all()
    assert (np.linalg.norm(d, axis=1) < 1e-3 * np.maximum(scale, 1)).all()


def test_colour_transfer_and_packing():
    c = (np.arange(256) / 255.0).astype(np.float32)
    lin, back = np.zeros_like(c), np.zeros_like(c)
    O.oracle().or_srgb(256, f32p(c), 0, f32p(lin))
    O.oracle().or_srgb(256, f32p(lin), 1, f32p(back))
    assert np.abs(back - c).max() < 1e-5
    rgb = np.array([[1, 1, 1], [0, 0, 0], [100, 100, 100]], np.float32)
    tm, packed = np.zeros_like(rgb), np.zeros(3, np.uint32)
    O.oracle().or_tonemap(3, f32p(rgb), 1.0, 0, f32p(tm), packed.ctypes.data)
    assert packed[0] == 0xFFFFFFFF and packed[1] == 0xFF000000
    O.oracle().or_tonemap(3, f32p(rgb), 1.0, 1, f32p(tm), packed.ctypes.data)
    assert abs(tm[1, 0]) < 1e-3 and 0.95 < tm[2, 0] <= 1.0  # ACES fit: black stays black, bright saturates


def test_alias_table_matches_weights():
    from photon_amd.types import AliasEntry
    w = np.array([1, 0, 3, 0.5, 5.5], np.float32)
    tab = (AliasEntry * 5)()
    if O.ref_available():
        O.ref().ref_alias_table(f32p(w), 5, C.cast(tab, C.c_void_p))
    # The oracle's own copy of the table comes from a built scene: five emissive quads with these powers.
    from photon_amd import host
    from photon_amd.types import Material
    s = host.Scene()
    for k, wk in enumerate(w):
        m = Material.default()
        m.metallic = 0.0
        m.emission[:] = (float(wk),) * 3 if wk > 0 else (0.0, 0.0, 0.0)
        mi = s.add_material(m)
        s.add_mesh(np.array([[k, 0, 0], [k + 1, 0, 0], [k, 1, 0]], np.float32), np.array([0, 1, 2], np.int32), mi)
    s.build()
    d = s.desc()
    n = d.num_emissive
    assert n == 4  # the zero-power triangle is not an emitter
    pmf = np.array([d.emissive_alias[i].pmf for i in range(n)])
    prob = np.array([d.emissive_alias[i].prob for i in range(n)])
    alias = np.array([d.emissive_alias[i].alias for i in range(n)])
    nz = w[w > 0]
    assert np.allclose(pmf, nz / nz.sum(), atol=1e-6)
    # Exact pmf of the table: bin i keeps prob_i / n and receives (1 - prob_j) / n from bins aliasing to it.
    got = prob / n
    for j in range(n):
        got[alias[j]] += (1 - prob[j]) / n
    assert np.allclose(got, pmf, atol=1e-6)
    if O.ref_available():
        assert np.allclose([tab[i].pmf for i in range(5)], w / w.sum(), atol=1e-6)
