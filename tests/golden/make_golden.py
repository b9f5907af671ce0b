"""Writes the golden fixtures of tests/golden/ from the reference's own CPU build.

The reference is compiled in place from its sources by `make -C oracle ref` (oracle/_ref, never
committed). This script feeds it the authored arrays produced by this repository's generators, lets the
reference build its BVH and light tables and render, and stores the outputs as small .npz files:

  bvh_<scene>.npz      BVH nodes, BVH order, BVH-order triangles, alias table, scene scalars
  hits_soup.npz        (t, tri, u, v) of seeded closest-hit and any-hit rays on a random soup
  render_<scene>.npz   per-pixel average radiance of the reference CpuRenderer (and the G-buffer)

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle_py as O  # noqa: E402
from photon_amd import host  # noqa: E402
from photon_amd.types import RAY_DTYPE, RenderSettings  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# Renders fixed by the fixtures: (scene, width, height, spp, max_depth, seed).
RENDERS = [("cornell-c1", 64, 64, 16, 8, 1), ("cornell", 48, 48, 8, 8, 0), ("materials", 48, 27, 4, 8, 2), ("glass", 48, 27, 4, 8, 3),
           ("field-small", 64, 36, 2, 16, 1), ("showcase", 64, 40, 8, 8, 4)]


def make_scene(name):
    if name == "field-small":
        return host.Scene.field(48, seed=5, env_width=256)
    if name == "soup":
        return soup_scene()
    if name == "showcase":
        return showcase_scene()
    return host.Scene.builtin(name)


def _pack(rgba):
    """(H, W, 4) uint8 -> (H, W) uint32 RGBA8 texels, R in the low byte."""
    a = np.asarray(rgba, np.uint32)
    return a[..., 0] | (a[..., 1] << 8) | (a[..., 2] << 16) | (a[..., 3] << 24)


def _uv_sphere(center, radius, rings, segments):
    """Latitude-longitude sphere with per-vertex normals and UVs (u around, v from pole to pole)."""
    th = np.linspace(0, np.pi, rings + 1)
    ph = np.linspace(0, 2 * np.pi, segments + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    uv = np.stack([P / (2 * np.pi), T / np.pi], -1).reshape(-1, 2)
    idx = []
    for r in range(rings):
        for k in range(segments):
            a, b = r * (segments + 1) + k, (r + 1) * (segments + 1) + k
            idx += [a, a + 1, b, a + 1, b + 1, b]
    return (np.asarray(center) + radius * n).astype(np.float32), np.asarray(idx, np.int32), n.astype(np.float32), uv.astype(np.float32)


def showcase_scene():
    """Every shading input the built-in scenes leave out (integrator.h:102-125, lights.h:164-190,
    integrator.h:265-280): a normal-mapped ground with base-colour (sRGB) and metallic-roughness textures,
    a normal-mapped sphere with an emissive texture, a metal box whose roughness and metalness come from
    a clamped texture, point, spot and directional lights, a dim constant environment, and a thin-lens
    camera (aperture 0.08, focused at 6)."""
    from photon_amd.types import Material, PunctualLight
    rng = np.random.default_rng(11)
    s = host.Scene()
    n = 64
    y, x = np.mgrid[0:n, 0:n]
    checker = ((x // 8 + y // 8) % 2).astype(np.float64)
    base = np.stack([60 + 170 * checker, 90 + 120 * (1 - checker), 40 + 60 * checker, np.full_like(checker, 255)], -1)
    t_base = s.add_texture(_pack(base.astype(np.uint8)), repeat=True, srgb=True)
    # Metallic-roughness: G roughness, B metalness (glTF); stripes of rough dielectric and smooth metal.
    stripe = ((x // 16) % 2).astype(np.float64)
    mr = np.stack([np.zeros_like(stripe), 40 + 200 * (1 - stripe) * (0.5 + 0.5 * np.sin(y / 5.0)), 255 * stripe, np.full_like(stripe, 255)], -1)
    t_mr = s.add_texture(_pack(np.clip(mr, 0, 255).astype(np.uint8)), repeat=True, srgb=False)
    # Normal map: bumps of height sin(x)sin(y), stored as (n+1)/2.
    hx = np.cos(2 * np.pi * x / 16) * np.sin(2 * np.pi * y / 16)
    hy = np.sin(2 * np.pi * x / 16) * np.cos(2 * np.pi * y / 16)
    nm = np.stack([-0.6 * hx, -0.6 * hy, np.ones_like(hx)], -1)
    nm /= np.linalg.norm(nm, axis=-1, keepdims=True)
    t_nm = s.add_texture(_pack(np.concatenate([np.round((nm + 1) * 127.5), np.full((n, n, 1), 255)], -1).astype(np.uint8)), repeat=True, srgb=False)
    # Emissive texture: warm rings.
    r = np.hypot(x - 32, y - 32)
    em = np.stack([255 * (0.5 + 0.5 * np.cos(r / 3)), 160 * (0.5 + 0.5 * np.cos(r / 3)), 40 + 0 * r, np.full_like(r, 255)], -1)
    t_em = s.add_texture(_pack(np.clip(em, 0, 255).astype(np.uint8)), repeat=True, srgb=True)
    # A clamped (non-repeating) metallic-roughness texture with random blocks.
    blocks = rng.integers(0, 256, (8, 8, 4)).astype(np.uint8).repeat(4, 0).repeat(4, 1)
    blocks[..., 3] = 255
    t_blk = s.add_texture(_pack(blocks), repeat=False, srgb=False)

    ground = Material.default()
    ground.base_color[:] = (0.9, 0.9, 0.9, 1.0)
    ground.metallic, ground.roughness = 1.0, 1.0
    ground.base_color_tex, ground.metallic_roughness_tex, ground.normal_tex = t_base, t_mr, t_nm
    ground.normal_scale = 0.8
    m_ground = s.add_material(ground)
    gp = np.array([[-6, 0, -6], [6, 0, -6], [6, 0, 6], [-6, 0, 6]], np.float32)
    guv = np.array([[0, 0], [3, 0], [3, 3], [0, 3]], np.float32)
    s.add_mesh(gp, np.array([0, 2, 1, 0, 3, 2], np.int32), m_ground, normals=np.tile([0, 1, 0], (4, 1)).astype(np.float32), uvs=guv)

    glow = Material.default()
    glow.base_color[:] = (0.7, 0.7, 0.75, 1.0)
    glow.metallic, glow.roughness = 0.0, 0.45
    glow.emission[:] = (3.0, 3.0, 3.0)
    glow.emissive_tex, glow.normal_tex, glow.normal_scale = t_em, t_nm, 1.3
    m_glow = s.add_material(glow)
    p, i, nn, uv = _uv_sphere((-1.3, 1.0, 0.2), 1.0, 12, 24)
    s.add_mesh(p, i, m_glow, normals=nn, uvs=uv * np.float32(2.0))

    box = Material.default()
    box.base_color[:] = (0.95, 0.64, 0.54, 1.0)
    box.metallic_roughness_tex = t_blk
    m_box = s.add_material(box)
    c = np.array([1.6, 0.7, -0.6], np.float32)
    corners = np.array([[x_, y_, z_] for x_ in (-1, 1) for y_ in (-1, 1) for z_ in (-1, 1)], np.float32) * np.float32(0.7) + c
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    bp, bi, buv = [], [], []
    for f in faces:
        k = len(bp)
        bp += [corners[j] for j in f]
        buv += [[-0.2, -0.2], [1.2, -0.2], [1.2, 1.2], [-0.2, 1.2]]
        bi += [k, k + 1, k + 2, k, k + 2, k + 3]
    s.add_mesh(np.array(bp, np.float32), np.array(bi, np.int32), m_box, uvs=np.array(buv, np.float32))

    glass = Material.default()
    glass.base_color[:] = (0.9, 0.95, 1.0, 1.0)
    glass.metallic, glass.roughness, glass.transmission, glass.ior = 0.0, 0.15, 1.0, 1.45
    m_glass = s.add_material(glass)
    p, i, nn, uv = _uv_sphere((0.5, 0.55, 1.6), 0.55, 8, 16)
    s.add_mesh(p, i, m_glass, normals=nn, uvs=uv)

    point = PunctualLight()
    point.type, point.position[:], point.intensity[:] = 0, (0.0, 3.5, 2.5), (18.0, 16.0, 12.0)
    spot = PunctualLight()
    spot.type, spot.position[:], spot.direction[:] = 1, (3.0, 4.0, 1.0), tuple(np.array([-0.55, -0.8, -0.25]) / np.linalg.norm([-0.55, -0.8, -0.25]))
    spot.intensity[:] = (40.0, 30.0, 60.0)
    spot.cos_inner, spot.cos_outer = float(np.cos(np.radians(12))), float(np.cos(np.radians(25)))
    sun = PunctualLight()
    sun.type, sun.direction[:], sun.intensity[:] = 2, tuple(np.array([0.3, -1.0, -0.45]) / np.linalg.norm([0.3, -1.0, -0.45])), (1.2, 1.1, 1.0)
    for light in (point, spot, sun):
        s.add_punctual(light)
    s.set_env_color((0.04, 0.05, 0.08), 1.0)
    s.set_camera((0.5, 2.4, 6.5), (0.0, 0.6, 0.0), vfov=50.0, aperture=0.08, focus_distance=6.0)
    return s


def soup_scene():
    """The reference test soup's recipe (test_bvh.cpp:14-28) drawn with numpy: 2000 triangles."""
    rng = np.random.default_rng(20240601)
    n = 2000
    c = rng.random((n, 3), dtype=np.float32)
    size = np.where(np.arange(n) % 50 == 0, 0.5, 0.02 + 0.05 * rng.random(n)).astype(np.float32)
    pos = c[:, None, :] + (rng.random((n, 3, 3), dtype=np.float32) - 0.5) * size[:, None, None]
    s = host.Scene()
    from photon_amd.types import Material
    m = s.add_material(Material.default())
    s.add_mesh(pos.reshape(-1, 3).astype(np.float32), np.arange(n * 3, dtype=np.int32), m)
    return s


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def soup_rays(n=4096, seed=7):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, RAY_DTYPE)
    rays["o"] = (rng.random((n, 3)) * 1.6 - 0.3).astype(np.float32)
    z = 1 - 2 * rng.random(n)
    r = np.sqrt(np.maximum(0, 1 - z * z))
    phi = 2 * np.pi * rng.random(n)
    rays["d"] = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)
    rays["tmax"] = np.where(np.arange(n) % 4 == 0, 0.3, 1e30).astype(np.float32)
    return rays


def main():
    O.build(ref=True)
    assert O.ref_available(), "the reference CPU build is required to write fixtures"
    for name in ("cornell-c1", "cornell", "furnace", "soup", "field-small", "showcase"):
        s = make_scene(name)
        ref = O.RefScene(raw=s.raw()).build(threads=4)
        a = O.desc_arrays(ref.desc())
        a["order"] = ref.order()
        a["nodes"] = a["nodes"].view(np.uint8)
        # Small arrays are stored whole; every array is also pinned by its SHA-256 digest.
        keep = {k: v for k, v in a.items() if v.nbytes <= 65536}
        digests = np.array([f"{k}:{sha(v)}" for k, v in sorted(a.items())])
        np.savez_compressed(os.path.join(OUT, f"bvh_{name}.npz"), digests=digests, **keep)
        print("bvh", name, len(a["nodes"]) // 64, sorted(keep))
    s = soup_scene()
    ref = O.RefScene(raw=s.raw()).build(threads=4)
    rays = soup_rays()
    np.savez_compressed(os.path.join(OUT, "hits_soup.npz"), rays=rays, closest=ref.trace(rays), any=ref.trace(rays, anyhit=True))
    for name, w, h, spp, depth, seed in RENDERS:
        s = make_scene(name)
        ref = O.RefScene(raw=s.raw()).build(threads=4)
        cam = ref.camera_view(w / h)
        rs = RenderSettings.make(w, h, max_depth=depth, rr_depth=3, seed=seed)
        img = ref.render(cam, rs, spp, threads=8)
        g = ref.render_gbuffer(cam, rs, 0)
        np.savez_compressed(os.path.join(OUT, f"render_{name}.npz"), image=img, camera=np.frombuffer(bytes(cam), np.uint8),
                            settings=np.array([w, h, spp, depth, seed]), gbuffer=np.stack(g))
        print("render", name, img.mean())


if __name__ == "__main__":
    main()
