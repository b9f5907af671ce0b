"""BVH construction and traversal pins (the reference's test_bvh.cpp): traversal equals brute force over
all triangles, the tree satisfies its structure invariants, the SAH keeps node visits sub-linear, and
alpha-masked triangles are skipped."""
import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd import host
from photon_amd.types import ALPHA_MASK, Material, RAY_DTYPE

from golden_scenes import soup_scene


def brute_force(desc, rays):
    """Moller-Trumbore against every triangle (numpy float32, the reference's evaluation order)."""
    a = O.desc_arrays(desc)
    tri = a["tris"].view(np.float32).reshape(-1, 12)
    v0, e1, e2 = tri[:, 0:3], tri[:, 4:7], tri[:, 8:11]
    out_t = np.full(len(rays), 1e30, np.float32)
    out_i = np.full(len(rays), -1, np.int32)
    for k, r in enumerate(rays):
        o, d, tmax = r["o"].astype(np.float32), r["d"].astype(np.float32), np.float32(r["tmax"])
        p = np.stack([d[1] * e2[:, 2] - d[2] * e2[:, 1], d[2] * e2[:, 0] - d[0] * e2[:, 2], d[0] * e2[:, 1] - d[1] * e2[:, 0]], 1)
        det = e1[:, 0] * p[:, 0] + e1[:, 1] * p[:, 1] + e1[:, 2] * p[:, 2]
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            inv = np.float32(1) / det
            s = o - v0
            u = (s[:, 0] * p[:, 0] + s[:, 1] * p[:, 1] + s[:, 2] * p[:, 2]) * inv
            q = np.stack([s[:, 1] * e1[:, 2] - s[:, 2] * e1[:, 1], s[:, 2] * e1[:, 0] - s[:, 0] * e1[:, 2], s[:, 0] * e1[:, 1] - s[:, 1] * e1[:, 0]], 1)
            v = (d[0] * q[:, 0] + d[1] * q[:, 1] + d[2] * q[:, 2]) * inv
            t = (e2[:, 0] * q[:, 0] + e2[:, 1] * q[:, 1] + e2[:, 2] * q[:, 2]) * inv
        ok = (np.abs(det) >= 1e-14) & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t > 0) & (t < tmax)
        if ok.any():
            i = np.flatnonzero(ok)[np.argmin(t[ok])]
            out_t[k], out_i[k] = t[i], i
    return out_t, out_i


def random_rays(n, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, RAY_DTYPE)
    rays["o"] = (rng.random((n, 3)) * 1.6 - 0.3).astype(np.float32)
    z = 1 - 2 * rng.random(n)
    r = np.sqrt(np.maximum(0, 1 - z * z))
    phi = 2 * np.pi * rng.random(n)
    rays["d"] = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)
    rays["tmax"] = np.where(np.arange(n) % 4 == 0, 0.3, 1e30).astype(np.float32)
    return rays


def test_traversal_matches_brute_force():
    s = soup_scene().build()
    rays = random_rays(4000, 7)
    hits = O.trace(s.desc(), rays)
    bt, bi = brute_force(s.desc(), rays)
    hit = hits["tri"] >= 0
    assert np.array_equal(hit, bi >= 0)
    assert np.all(np.abs(hits["t"][hit] - bt[hit]) <= 1e-5 * np.maximum(1, bt[hit]))
    anyhit = O.trace(s.desc(), rays, anyhit=True)
    assert np.array_equal(anyhit["tri"] >= 0, bi >= 0)
    assert hit.mean() > 0.1


def test_structure_invariants():
    s = host.Scene.spheres(8, alpha_fraction=0.0, seed=3).build()
    order = s.order()
    assert np.array_equal(np.sort(order), np.arange(s.num_tris))
    nodes = s.nodes()
    a = O.desc_arrays(s.desc())
    tri = a["tris"].view(np.float32).reshape(-1, 12)
    v = np.stack([tri[:, 0:3], tri[:, 0:3] + tri[:, 4:7], tri[:, 0:3] + tri[:, 8:11]], 1)
    covered = np.zeros(s.num_tris, int)
    leaves = 0
    for side in ("l", "r"):
        cnt = nodes[side + "count"]
        first = nodes["left" if side == "l" else "right"]
        lo, hi = nodes[side + "min"], nodes[side + "max"]
        for k in np.flatnonzero(cnt > 0):
            leaves += 1
            assert cnt[k] <= 8
            idx = np.arange(first[k], first[k] + cnt[k])
            covered[idx] += 1
            pts = v[idx].reshape(-1, 3)
            assert (pts >= lo[k] - 1e-5).all() and (pts <= hi[k] + 1e-5).all()
        # Interior children point at later nodes (depth-first numbering).
        inner = (cnt == 0) & (first >= 0)
        assert (first[inner] > np.flatnonzero(inner)).all()
    assert (covered == 1).all()
    st = s.stats()
    assert leaves == st.leaves and st.nodes == len(nodes) and st.sah_cost > 0


def test_sah_keeps_node_visits_sublinear():
    small = host.Scene.spheres(1, seed=4).build()
    large = host.Scene.spheres(64, seed=4).build()

    def visits(sc):
        cam = sc.camera_view(1.0)
        rays = host.camera_rays(cam, 64, 64)
        _, c = O.trace(sc.desc(), rays, counters=True)
        return c[:, 0].mean()

    a, b = visits(small), visits(large)
    assert b < a * 8  # 64x the triangles


def test_threaded_and_serial_builds_agree():
    s1 = host.Scene.spheres(40, seed=6).build(threads=1)
    s8 = host.Scene.spheres(40, seed=6).build(threads=8)
    assert s1.nodes().tobytes() == s8.nodes().tobytes()
    assert np.array_equal(s1.order(), s8.order())


def test_alpha_masked_triangles_are_skipped():
    s = host.Scene()
    masked = Material.default()
    masked.base_color[:] = (1, 1, 1, 0)
    masked.alpha_mode = ALPHA_MASK
    m0 = s.add_material(masked)
    m1 = s.add_material(Material.default())
    quad = lambda z: np.array([[-1, -1, z], [1, -1, z], [1, 1, z], [-1, 1, z]], np.float32)
    s.add_mesh(quad(1), np.array([0, 1, 2, 0, 2, 3], np.int32), m0)
    s.add_mesh(quad(0), np.array([0, 1, 2, 0, 2, 3], np.int32), m1)
    s.build()
    rays = np.zeros(1, RAY_DTYPE)
    rays["o"] = [0.1, 0.2, 5]
    rays["d"] = [0, 0, -1]
    rays["tmax"] = 1e30
    h = O.trace(s.desc(), rays)
    assert h["tri"][0] >= 0 and abs(h["t"][0] - 5) < 1e-4
    a = O.desc_arrays(s.desc())
    assert a["tris"].view(np.int32).reshape(-1, 12)[h["tri"][0], 3] == m1


def test_empty_scene_and_single_leaf():
    s = host.Scene()
    s.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([0, 1, 2], np.int32), s.add_material(Material.default()))
    s.build()
    nodes = s.nodes()
    assert len(nodes) == 1 and nodes["lcount"][0] == 1 and nodes["right"][0] == -1
    rays = np.zeros(2, RAY_DTYPE)
    rays["o"] = [[0.2, 0.2, 1], [5, 5, 1]]
    rays["d"] = [[0, 0, -1], [0, 0, -1]]
    rays["tmax"] = 1e30
    h = O.trace(s.desc(), rays)
    assert h["tri"].tolist() == [0, -1]
