"""Scenes that come in through the glTF importer, rendered by the device path and checked against the
CPU oracle on the same imported scene: the Cornell workload written to glTF and read back, and a
textured, alpha-masked glTF whose PNG texture is decoded by the host library (the path the reference's
Sponza-like assets take, src/scene/gltf.cpp:215-247)."""
import base64
import io
import json

import numpy as np
import pytest
from PIL import Image

from oracle import oracle_py as O
from photon_amd import gpu, host
from photon_amd.types import RenderSettings

from gpu_util import bitwise_fraction, radiance_agreement, uploaded

pytestmark = pytest.mark.gpu


def _render_both(scene, w, h, spp, engine):
    rs = RenderSettings.make(w, h, max_depth=8, rr_depth=3, seed=1)
    cam = scene.camera_view(w / h)
    r = uploaded(scene, w, h, engine)
    r.render(cam, rs, spp)
    img = r.download_image()
    ref = O.render(scene.desc(), cam, rs, spp)
    return img, ref


@pytest.mark.parametrize("engine", [gpu.ENGINE_MEGAKERNEL, gpu.ENGINE_WAVEFRONT])
def test_cornell_through_gltf(tmp_path, engine):
    src = host.Scene.builtin("cornell-c1")
    src.save_gltf(tmp_path / "cornell.gltf")
    s = host.Scene.from_gltf(tmp_path / "cornell.gltf").build()
    assert s.num_tris == src.num_tris
    img, ref = _render_both(s, 96, 96, 8, engine)
    agree = radiance_agreement(img, ref)
    print(f"engine {engine}: {100 * agree:.3f}% within 0.1%, {100 * bitwise_fraction(img, ref):.2f}% bitwise, mean {img.mean():.5f}")
    assert agree >= 0.999 and img.mean() > 0.01


@pytest.mark.parametrize("fmt", ["PNG", "JPEG"])
def test_textured_alpha_masked_gltf(tmp_path, fmt):
    """A floor, a back wall and an alpha-masked leaf quad with a base-colour texture (PNG with a disc of
    alpha, or a baseline JPEG, which is opaque), lit by an emissive quad: primary and shadow rays run the
    alpha test against the texels the host library decoded."""
    g = np.random.default_rng(12)
    tex = np.zeros((64, 64, 4), np.uint8)
    tex[..., :3] = g.integers(40, 255, (64, 64, 3))
    yy, xx = np.mgrid[0:64, 0:64]
    tex[..., 3] = np.where((xx - 32) ** 2 + (yy - 32) ** 2 < 26 ** 2, 255, 0)  # a disc
    png = io.BytesIO()
    if fmt == "PNG":
        Image.fromarray(tex, "RGBA").save(png, "PNG")
    else:
        Image.fromarray(tex[..., :3], "RGB").save(png, "JPEG", quality=90)
    quads = [  # corner, edge u, edge v, material
        ((-4, 0, -4), (8, 0, 0), (0, 0, 8), 0),
        ((-4, 0, -4), (0, 8, 0), (8, 0, 0), 0),
        ((-1.5, 0.5, 0), (3, 0, 0), (0, 3, 0), 1),
        ((-1, 7.9, -1), (2, 0, 0), (0, 0, 2), 2),  # faces down
    ]
    pos, uv = [], []
    for c, u, v, _ in quads:
        c, u, v = map(np.array, (c, u, v))
        pos += [c, c + u, c + u + v, c + v]
        uv += [(0, 1), (1, 1), (1, 0), (0, 0)]
    prims = []
    bin_bytes = np.array(pos, np.float32).tobytes() + np.array(uv, np.float32).tobytes()
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": 16 * 12}, {"buffer": 0, "byteOffset": 16 * 12, "byteLength": 16 * 8}]
    accessors = [{"bufferView": 0, "componentType": 5126, "count": 16, "type": "VEC3"},
                 {"bufferView": 1, "componentType": 5126, "count": 16, "type": "VEC2"}]
    for q, (_, _, _, mat) in enumerate(quads):
        ib = np.array([4 * q, 4 * q + 1, 4 * q + 2, 4 * q, 4 * q + 2, 4 * q + 3], np.uint32).tobytes()
        views.append({"buffer": 0, "byteOffset": len(bin_bytes), "byteLength": 24})
        accessors.append({"bufferView": len(views) - 1, "componentType": 5125, "count": 6, "type": "SCALAR"})
        bin_bytes += ib
        prims.append({"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": len(accessors) - 1, "material": mat})
    doc = {"asset": {"version": "2.0"},
           "buffers": [{"byteLength": len(bin_bytes), "uri": "data:application/octet-stream;base64," + base64.b64encode(bin_bytes).decode()}],
           "bufferViews": views, "accessors": accessors,
           "images": [{"uri": f"data:image/{fmt.lower()};base64," + base64.b64encode(png.getvalue()).decode()}],
           "textures": [{"source": 0}],
           "materials": [{"pbrMetallicRoughness": {"baseColorFactor": [0.7, 0.7, 0.7, 1], "metallicFactor": 0, "roughnessFactor": 0.8}},
                         {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicFactor": 0, "roughnessFactor": 0.5},
                          "alphaMode": "MASK", "alphaCutoff": 0.5, "doubleSided": True},
                         {"pbrMetallicRoughness": {"baseColorFactor": [0, 0, 0, 1], "metallicFactor": 0},
                          "emissiveFactor": [1, 0.9, 0.8], "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 20}}}],
           "meshes": [{"primitives": prims}],
           "cameras": [{"type": "perspective", "perspective": {"yfov": 0.9, "znear": 0.01}}],
           "nodes": [{"mesh": 0}, {"camera": 0, "translation": [0, 3, 9]}],
           "scenes": [{"nodes": [0, 1]}]}
    (tmp_path / "leaf.gltf").write_text(json.dumps(doc))
    s = host.Scene.from_gltf(tmp_path / "leaf.gltf").build()
    assert s.raw().num_textures == 1 and s.material(1).base_color_tex == 0
    img, ref = _render_both(s, 80, 64, 8, gpu.ENGINE_WAVEFRONT)
    agree = radiance_agreement(img, ref)
    print(f"textured glTF ({fmt}): {100 * agree:.3f}% within 0.1%, {100 * bitwise_fraction(img, ref):.2f}% bitwise, mean {img.mean():.4f}")
    assert agree >= 0.999 and img.mean() > 0.01
