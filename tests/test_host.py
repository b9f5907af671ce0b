"""Host producers: the workload generators produce the configured scenes, ray sets have the stated
properties, and the C ABI rejects bad input with an error instead of crashing."""
import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd import host
from photon_amd.types import Material


def test_cornell_workload_scene():
    s = host.Scene.builtin("cornell-c1").build()
    assert s.num_tris == 36  # 5 walls x 2, two boxes x 12, light quad x 2
    assert s.num_materials == 6
    gold, glass = s.material(4), s.material(5)
    assert gold.metallic == 1 and abs(gold.roughness - 0.3) < 1e-7 and abs(gold.base_color[1] - 0.78) < 1e-7
    assert glass.transmission == 1 and abs(glass.ior - 1.5) < 1e-7 and abs(glass.roughness - 0.1) < 1e-7
    light = s.material(3)
    assert list(light.emission) == [17, 12, 4]
    assert abs(s.camera()["vfov"] - 60) < 1e-6
    d = s.desc()
    assert d.num_emissive == 2 and abs(d.p_emissive - 1.0) < 1e-6


def test_traversal_workload_scene():
    s = host.Scene.spheres(256, box=100.0, alpha_fraction=0.1, seed=1)
    assert s.num_tris == 262144
    s.build()
    a = O.desc_arrays(s.desc())
    flags = a["tris"].view(np.int32).reshape(-1, 12)[:, 7]
    frac = (flags & 1).mean()
    assert 0.095 < frac < 0.105
    assert s.desc().num_textures == 1
    raw = s.raw()
    assert raw.tex_width[0] == 64 and raw.tex_height[0] == 64


def test_path_tracing_workload_scene():
    s = host.Scene.field(1024, seed=1, env_width=64)
    assert s.num_tris == 1024 * 1024 + 2
    mats = [s.material(i) for i in range(1, s.num_materials)]
    metal = np.mean([m.metallic == 1 for m in mats])
    glass = np.mean([m.transmission == 1 for m in mats])
    emitters = sum(m.emission[0] > 0 for m in mats)
    assert emitters == 64
    assert 0.25 < metal < 0.35 and 0.15 < glass < 0.25
    rough = np.array([m.roughness for m in mats])
    assert rough.min() >= 0.02 and rough.max() <= 1.0
    e = np.array([m.emission[0] for m in mats if m.emission[0] > 0])
    assert e.min() >= 1 and e.max() <= 50


def test_gradient_sky_and_sun():
    s = host.Scene.field(4, seed=1, env_width=2048)
    raw = s.raw()
    w, h = raw.env_width, raw.env_height
    assert (w, h) == (2048, 1024)
    tex = np.ctypeslib.as_array(raw.env_texels, (h, w, 3))
    lum = 0.2126 * tex[..., 0] + 0.7152 * tex[..., 1] + 0.0722 * tex[..., 2]
    sky = lum[lum < 1e3]
    assert abs(sky.min() - 0.2) < 1e-3 and sky.max() <= 1.0 + 1e-3 and sky.max() > 0.99
    sun = int((lum > 1e3).sum())
    # A 0.5 degree disk covers ~6.3 texels of (360/2048) x (180/1024) degrees near its elevation.
    assert 2 <= sun <= 16


def test_shadow_rays_have_the_stated_length():
    s = host.Scene.spheres(16, seed=2).build()
    rays = s.shadow_rays(1000, seed=3)
    assert np.allclose(np.linalg.norm(rays["d"], axis=1), 1, atol=1e-6)
    assert (rays["tmax"] > 0).all()


def test_camera_rays_cover_the_view():
    s = host.Scene.builtin("cornell-c1")
    cam = s.camera_view(1.0)
    rays = host.camera_rays(cam, 8, 8)
    assert np.allclose(np.linalg.norm(rays["d"], axis=1), 1, atol=1e-6)
    assert (rays["tmax"] == np.float32(1e30)).all()
    centre = rays["d"].reshape(8, 8, 3)[3:5, 3:5].mean((0, 1))
    assert np.dot(centre / np.linalg.norm(centre), np.array(cam.forward)) > 0.99


def test_bad_input_is_rejected():
    s = host.Scene()
    with pytest.raises(RuntimeError):
        s.add_mesh(np.zeros((3, 3), np.float32), np.array([0, 1, 5], np.int32), 0)
    with pytest.raises(RuntimeError):
        host.Scene.builtin("no-such-scene")
    with pytest.raises(RuntimeError):
        s.desc()  # not built
    with pytest.raises(RuntimeError):
        s.build(bins=100)


def test_rebuild_lights_follows_material_edits():
    s = host.Scene.builtin("cornell").build()
    m = s.material(0)
    m.emission[:] = (1.0, 1.0, 1.0)
    s.set_material(0, m)
    s.rebuild_lights()
    assert s.desc().num_emissive > 2
