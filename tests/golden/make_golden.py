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
           ("field-small", 64, 36, 2, 16, 1)]


def make_scene(name):
    if name == "field-small":
        return host.Scene.field(48, seed=5, env_width=256)
    if name == "soup":
        return soup_scene()
    return host.Scene.builtin(name)


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
    for name in ("cornell-c1", "cornell", "furnace", "soup", "field-small"):
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
