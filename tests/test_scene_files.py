"""Scene files and images of the host library: the JSON reader, glTF import and export, the PNG codec,
PFM and Radiance HDR.

The glTF cases restate the reference's own tests (tests/test_gltf.cpp:47-180): the JSON parser case, the
round trip of the built-in "materials" scene with a spot light, node transforms and mirrored winding,
and error reporting. Expected values beyond those come from independent checkers: Python's json module,
Pillow's PNG codec, and encoders written here from the PNG, Radiance and PFM format definitions."""
import base64
import io
import json
import math
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

from photon_amd import host
from photon_amd.types import PunctualLight

LIGHT_SPOT = 1


# ---- JSON ---------------------------------------------------------------------------------------------

def test_json_parser_reference_case():
    """test_gltf.cpp:47-65: values, escapes, UTF-8, and an error on a malformed array."""
    text = '{"a": [1, 2.5e1, -3], "b": {"c": "xé\\n"}, "t": true, "n": null}'
    assert json.loads(host.json_canonical(text)) == json.loads(text)
    with pytest.raises(RuntimeError, match=r"json: .*\(line 1, column 12\)"):
        host.json_canonical('{"a": [1, 2}')


def _random_value(g, depth):
    k = g.integers(0, 7 if depth < 4 else 4)
    if k == 0:
        return None
    if k == 1:
        return bool(g.integers(0, 2))
    if k == 2:
        return float(g.normal() * 10.0 ** g.integers(-30, 30)) if g.integers(0, 2) else int(g.integers(-2**40, 2**40))
    if k == 3:
        alphabet = ['a', 'Z', '0', ' ', '"', '\\', '/', '\n', '\t', '\b', '\f', '\r', 'é', '€', '\U0001F600', '\x01']
        return ''.join(alphabet[i] for i in g.integers(0, len(alphabet), g.integers(0, 12)))
    if k in (4, 5):
        return [_random_value(g, depth + 1) for _ in range(g.integers(0, 5))]
    return {f"k{g.integers(0, 50)}": _random_value(g, depth + 1) for _ in range(g.integers(0, 5))}


def test_json_matches_python_json_on_random_documents():
    g = np.random.default_rng(3)
    for _ in range(300):
        v = _random_value(g, 0)
        for ascii_only in (True, False):
            text = json.dumps(v, ensure_ascii=ascii_only, indent=int(g.integers(0, 3)) or None)
            assert json.loads(host.json_canonical(text)) == json.loads(text), text


@pytest.mark.parametrize("bad", ["", "[1,]", '{"a" 1}', "[1 2]", '"abc', "tru", '{"a":}', "1 2", "[", "-", "1.", ".5", "1e", "nul",
                                 "01", '{"a":1,}', '"\\x"', '"\\u12"'])
def test_json_rejects_what_python_rejects(bad):
    with pytest.raises(ValueError):
        json.loads(bad)
    with pytest.raises(RuntimeError, match="json: "):
        host.json_canonical(bad)


def test_json_nesting_limit():
    """The reader refuses more than 512 nested levels, as the reference's does (src/scene/json.cpp:104)."""
    assert host.json_canonical("[" * 500 + "]" * 500) == "[" * 500 + "]" * 500
    with pytest.raises(RuntimeError, match="nesting too deep"):
        host.json_canonical("[" * 600 + "]" * 600)


# ---- glTF ---------------------------------------------------------------------------------------------

def _raw_arrays(scene):
    r = scene.raw()
    nv, nt = r.num_vertices, r.num_tris
    pos = np.ctypeslib.as_array(r.positions, (nv * 3,)).reshape(-1, 3).copy()
    idx = np.ctypeslib.as_array(r.indices, (nt * 3,)).reshape(-1, 3).copy()
    mat = np.ctypeslib.as_array(r.tri_material, (nt,)).copy()
    nrm = np.ctypeslib.as_array(r.normals, (nv * 3,)).reshape(-1, 3).copy() if r.normals else None
    uv = np.ctypeslib.as_array(r.uvs, (nv * 2,)).reshape(-1, 2).copy() if r.uvs else None
    tan = np.ctypeslib.as_array(r.tangents, (nv * 4,)).reshape(-1, 4).copy() if r.tangents else None
    lights = [r.punctual[i] for i in range(r.num_punctual)]
    return dict(pos=pos, idx=idx, mat=mat, nrm=nrm, uv=uv, tan=tan, lights=lights, r=r)


def _area(a):
    p = a["pos"][a["idx"]].astype(np.float64)
    return 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1).sum()


def _lum(e):
    return 0.2126 * e[0] + 0.7152 * e[1] + 0.0722 * e[2]


def test_gltf_round_trip(tmp_path):
    """test_gltf.cpp:67-119 on the built-in "materials" scene with a spot light added; in addition every
    triangle comes back with bit-identical vertices and its material."""
    a = host.Scene.builtin("materials")
    light = PunctualLight()
    light.type = LIGHT_SPOT
    light.position[:] = (1, 4, 2)
    d = np.array([0, -1, 0.2], np.float32)
    d = d / np.linalg.norm(d)
    light.direction[:] = d
    light.intensity[:] = (30, 20, 10)
    light.cos_inner = math.cos(0.3)
    light.cos_outer = math.cos(0.5)
    a.add_punctual(light)
    path = tmp_path / "roundtrip.gltf"
    a.save_gltf(path)
    b = host.Scene.from_gltf(path)
    A, B = _raw_arrays(a), _raw_arrays(b)
    assert b.num_tris == a.num_tris and b.num_materials == a.num_materials
    assert abs(_area(B) - _area(A)) <= 1e-3 * _area(A)
    for i in range(a.num_materials):
        x, y = a.material(i), b.material(i)
        for f in ("metallic", "roughness", "transmission", "ior"):
            assert abs(getattr(x, f) - getattr(y, f)) <= 1e-5, f
        assert abs(x.base_color[0] - y.base_color[0]) <= 1e-5
        assert abs(_lum(x.emission) - _lum(y.emission)) <= 1e-4
    assert np.array_equal(np.bincount(A["mat"], minlength=a.num_materials), np.bincount(B["mat"], minlength=a.num_materials))
    # Triangles per material, as vertex triples: identical bits.
    for m in range(a.num_materials):
        ta = A["pos"][A["idx"][A["mat"] == m]].reshape(-1, 9)
        tb = B["pos"][B["idx"][B["mat"] == m]].reshape(-1, 9)
        assert np.array_equal(ta[np.lexsort(ta.T)], tb[np.lexsort(tb.T)])
    cams = b.cameras()
    assert len(cams) == 1
    ca, cb = a.camera(), b.camera()
    assert np.linalg.norm(np.subtract(cb["position"], ca["position"])) <= 1e-4
    fa = np.subtract(ca["target"], ca["position"])
    fb = np.subtract(cb["target"], cb["position"])
    assert abs(np.dot(fa / np.linalg.norm(fa), fb / np.linalg.norm(fb)) - 1) <= 1e-5
    assert abs(cb["vfov"] - ca["vfov"]) <= 1e-3
    assert len(B["lights"]) == 1
    m = B["lights"][0]
    assert m.type == LIGHT_SPOT
    assert np.linalg.norm(np.subtract(list(m.position), (1, 4, 2))) <= 1e-4
    assert abs(np.dot(list(m.direction), d) - 1) <= 1e-5
    assert abs(m.intensity[0] - 30) <= 1e-3
    assert abs(m.cos_outer - light.cos_outer) <= 1e-5


def _doc_with_buffer(bin_bytes, body):
    doc = {"asset": {"version": "2.0"},
           "buffers": [{"byteLength": len(bin_bytes), "uri": "data:application/octet-stream;base64," + base64.b64encode(bin_bytes).decode()}]}
    doc.update(body)
    return doc


def _write(path, doc):
    path.write_text(json.dumps(doc))
    return path


def test_gltf_node_transforms_and_winding(tmp_path):
    """test_gltf.cpp:121-172: a triangle instanced by a translated child of a scaled parent and by a
    mirrored node, whose winding must be flipped; emissive strength."""
    pos = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32).tobytes()
    idx = np.array([0, 1, 2], np.uint16).tobytes() + b"\0\0"
    doc = _doc_with_buffer(pos + idx, {
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36}, {"buffer": 0, "byteOffset": 36, "byteLength": 6}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3", "min": [0, 0, 0], "max": [1, 1, 0]},
                      {"bufferView": 1, "componentType": 5123, "count": 3, "type": "SCALAR"}],
        "materials": [{"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.4, 0.6, 1], "metallicFactor": 0.0, "roughnessFactor": 0.3},
                       "emissiveFactor": [1, 0.5, 0], "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4}}}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1, "material": 0}]}],
        "nodes": [{"scale": [2, 2, 2], "children": [1]}, {"translation": [0, 0, 3], "mesh": 0}, {"scale": [-1, 1, 1], "mesh": 0}],
        "scenes": [{"nodes": [0, 2]}], "scene": 0})
    s = host.Scene.from_gltf(_write(tmp_path / "transforms.gltf", doc))
    a = _raw_arrays(s)
    assert s.num_tris == 2
    tri = a["pos"][a["idx"]]
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.abs(tri[0, 0] - (0, 0, 6)).max() <= 1e-5 and np.abs(tri[0, 1] - (2, 0, 6)).max() <= 1e-5
    assert abs(n[0, 2] - 1) <= 1e-5 and abs(n[1, 2] - 1) <= 1e-5
    assert np.any(np.linalg.norm(tri[1] - (-1, 0, 0), axis=1) < 1e-5)
    assert s.num_materials == 1
    m = s.material(0)
    assert abs(m.roughness - 0.3) <= 1e-6 and abs(m.emission[0] - 4) <= 1e-5 and abs(m.emission[1] - 2) <= 1e-5


def test_gltf_errors_are_reported(tmp_path):
    """test_gltf.cpp:174-187, plus malformed files the loader must refuse with a message."""
    with pytest.raises(RuntimeError, match="cannot open"):
        host.Scene.from_gltf(tmp_path / "does_not_exist.gltf")
    (tmp_path / "broken.gltf").write_text('{"asset": {"version": "2.0"}, "meshes": [')
    with pytest.raises(RuntimeError, match="json: .*line 1"):
        host.Scene.from_gltf(tmp_path / "broken.gltf")
    cases = {
        "only glTF 2.0": {"asset": {"version": "1.0"}},
        "has no data": {"asset": {"version": "2.0"}, "buffers": [{"byteLength": 4}]},
        "cannot read buffer": {"asset": {"version": "2.0"}, "buffers": [{"byteLength": 4, "uri": "missing.bin"}]},
        "invalid data URI": {"asset": {"version": "2.0"}, "buffers": [{"byteLength": 4, "uri": "data:application/octet-stream;base64,@@@@"}]},
    }
    pos = np.zeros(9, np.float32).tobytes()
    base = {"bufferViews": [{"buffer": 0, "byteLength": 36}], "meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}],
            "nodes": [{"mesh": 0}], "scenes": [{"nodes": [0]}]}
    cases["out of the buffer's range"] = _doc_with_buffer(pos, dict(base, accessors=[{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"}]))
    cases["unsupported accessor type"] = _doc_with_buffer(pos, dict(base, accessors=[{"bufferView": 0, "componentType": 5126, "count": 3, "type": "MAT3"}]))
    bad_idx = dict(base, accessors=[{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                                    {"bufferView": 0, "componentType": 5125, "count": 3, "type": "SCALAR"}],
                   meshes=[{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}])
    cases["vertex index out of range"] = _doc_with_buffer(np.array([0, 1, 7, 0, 0, 0, 0, 0, 0], np.uint32).tobytes(), bad_idx)
    cases["too deep"] = {"asset": {"version": "2.0"}, "nodes": [{"children": [0]}], "scenes": [{"nodes": [0]}]}
    for i, (msg, doc) in enumerate(cases.items()):
        with pytest.raises(RuntimeError, match=msg):
            host.Scene.from_gltf(_write(tmp_path / f"bad{i}.gltf", doc))


def _glb(doc, bin_bytes):
    j = json.dumps(doc).encode()
    j += b" " * (-len(j) % 4)
    b = bin_bytes + b"\0" * (-len(bin_bytes) % 4)
    body = struct.pack("<I4s", len(j), b"JSON") + j + struct.pack("<I4s", len(b), b"BIN\0") + b
    return struct.pack("<4sII", b"glTF", 2, 12 + len(body)) + body


def _quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_glb_interleaved_strips_fans_and_normalized_accessors(tmp_path):
    """A .glb with an interleaved vertex buffer (byteStride), normalized integer UVs and tangents, a strip,
    a fan and a rotated, non-uniformly scaled node. Expected arrays are computed here in float64."""
    g = np.random.default_rng(4)
    nv = 7
    p = g.uniform(-1, 1, (nv, 3)).astype(np.float32)
    nrm = g.normal(size=(nv, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    uv16 = g.integers(0, 65536, (nv, 2)).astype(np.uint16)
    tan8 = g.integers(-128, 128, (nv, 4)).astype(np.int8)
    tan8[:, 3] = np.where(g.uniform(size=nv) < 0.5, -127, 127)
    # Interleaved: position (12), normal (12), uv (4), tangent (4) -> stride 32.
    inter = b"".join(p[i].tobytes() + nrm[i].tobytes() + uv16[i].tobytes() + tan8[i].tobytes() for i in range(nv))
    strip = np.array([0, 1, 2, 3, 4, 5, 6], np.uint8).tobytes()
    fan = np.array([6, 0, 1, 2, 3], np.uint32).tobytes()
    bin_bytes = inter + strip + b"\0" + fan
    q = np.array([0.1, 0.7, -0.2, 0.5])
    q /= np.linalg.norm(q)
    t, s = np.array([1.0, -2.0, 0.5]), np.array([2.0, 0.5, 1.5])
    doc = {"asset": {"version": "2.0"}, "buffers": [{"byteLength": len(bin_bytes)}],
           "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": len(inter), "byteStride": 32},
                           {"buffer": 0, "byteOffset": len(inter), "byteLength": 7},
                           {"buffer": 0, "byteOffset": len(inter) + 8, "byteLength": 20}],
           "accessors": [{"bufferView": 0, "byteOffset": 0, "componentType": 5126, "count": nv, "type": "VEC3"},
                         {"bufferView": 0, "byteOffset": 12, "componentType": 5126, "count": nv, "type": "VEC3"},
                         {"bufferView": 0, "byteOffset": 24, "componentType": 5123, "normalized": True, "count": nv, "type": "VEC2"},
                         {"bufferView": 0, "byteOffset": 28, "componentType": 5120, "normalized": True, "count": nv, "type": "VEC4"},
                         {"bufferView": 1, "componentType": 5121, "count": 7, "type": "SCALAR"},
                         {"bufferView": 2, "componentType": 5125, "count": 5, "type": "SCALAR"}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2, "TANGENT": 3}, "indices": 4, "mode": 5},
                                      {"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2, "TANGENT": 3}, "indices": 5, "mode": 6},
                                      {"attributes": {"POSITION": 0}, "mode": 1}]}],
           "nodes": [{"translation": t.tolist(), "rotation": q.tolist(), "scale": s.tolist(), "mesh": 0}]}
    path = tmp_path / "mesh.glb"
    path.write_bytes(_glb(doc, bin_bytes))
    sc = host.Scene.from_gltf(path)
    a = _raw_arrays(sc)
    R = _quat_matrix(q)
    M = R * s[None, :]
    # Strip (0,1,2) (2,1,3) (2,3,4) (4,3,5) (4,5,6); fan (6,0,1) (6,1,2) (6,2,3); the line primitive is dropped.
    strip_tris = [(0, 1, 2), (2, 1, 3), (2, 3, 4), (4, 3, 5), (4, 5, 6)]
    fan_tris = [(6, 0, 1), (6, 1, 2), (6, 2, 3)]
    want_idx = np.array(strip_tris + [(i + nv, j + nv, k + nv) for i, j, k in fan_tris])
    assert np.array_equal(a["idx"], want_idx)
    want_p = p.astype(np.float64) @ M.T + t
    assert np.abs(a["pos"] - np.vstack([want_p, want_p])).max() <= 1e-5
    Ninv = np.linalg.inv(M).T
    want_n = nrm.astype(np.float64) @ Ninv.T
    want_n /= np.linalg.norm(want_n, axis=1, keepdims=True)
    assert np.abs(a["nrm"] - np.vstack([want_n, want_n])).max() <= 1e-5
    want_uv = uv16.astype(np.float64) / 65535.0
    assert np.abs(a["uv"] - np.vstack([want_uv, want_uv])).max() <= 1e-7
    tv = np.maximum(tan8.astype(np.float64) / 127.0, -1.0)
    want_t = tv[:, :3] @ M.T
    want_t /= np.linalg.norm(want_t, axis=1, keepdims=True)
    assert np.abs(a["tan"][:, :3] - np.vstack([want_t, want_t])).max() <= 1e-5
    assert np.array_equal(a["tan"][:, 3], np.concatenate([tv[:, 3], tv[:, 3]]).astype(np.float32))
    # No material on the primitives: one shared default material.
    assert sc.num_materials == 1 and set(a["mat"]) == {0}
    m = sc.material(0)
    assert (m.base_color[0], m.metallic, m.roughness) == (np.float32(0.8), 0.0, np.float32(0.6))


def _rgba_u32(a):
    a = a.astype(np.uint32)
    return a[..., 0] | (a[..., 1] << 8) | (a[..., 2] << 16) | (a[..., 3] << 24)


def test_gltf_textures_samplers_and_undecodable_images(tmp_path, capfd):
    g = np.random.default_rng(8)
    img = g.integers(0, 256, (5, 7, 4), dtype=np.uint8)
    png = io.BytesIO()
    Image.fromarray(img, "RGBA").save(png, "PNG")
    jpg = io.BytesIO()
    Image.fromarray(img[..., :3], "RGB").save(jpg, "JPEG", quality=95)
    progressive = io.BytesIO()
    Image.fromarray(img[..., :3], "RGB").save(progressive, "JPEG", progressive=True)
    pos = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32).tobytes()
    uv = np.array([0, 0, 1, 0, 0, 1], np.float32).tobytes()
    doc = _doc_with_buffer(pos + uv, {
        "bufferViews": [{"buffer": 0, "byteLength": 36}, {"buffer": 0, "byteOffset": 36, "byteLength": 24}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                      {"bufferView": 1, "componentType": 5126, "count": 3, "type": "VEC2"}],
        "images": [{"uri": "data:image/png;base64," + base64.b64encode(png.getvalue()).decode()},
                   {"uri": "data:image/jpeg;base64," + base64.b64encode(jpg.getvalue()).decode()},
                   {"uri": "data:image/jpeg;base64," + base64.b64encode(progressive.getvalue()).decode()}],
        "samplers": [{"wrapS": 33071, "wrapT": 33071}],
        "textures": [{"source": 0, "sampler": 0}, {"source": 0}, {"source": 1}, {"source": 2}],
        "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicRoughnessTexture": {"index": 0}},
                       "emissiveTexture": {"index": 0}, "normalTexture": {"index": 1, "scale": 0.5},
                       "alphaMode": "MASK", "alphaCutoff": 0.25, "doubleSided": True},
                      {"pbrMetallicRoughness": {"baseColorTexture": {"index": 2}}, "alphaMode": "BLEND", "alphaCutoff": 0.9},
                      {"pbrMetallicRoughness": {"baseColorTexture": {"index": 3}}}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "material": 0},
                                   {"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "material": 1},
                                   {"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "material": 2}]}],
        "nodes": [{"mesh": 0}], "scenes": [{"nodes": [0]}]})
    s = host.Scene.from_gltf(_write(tmp_path / "tex.gltf", doc))
    r = s.raw()
    # base colour (sRGB) and emissive (sRGB) share one texture; metal-roughness (linear) gets its own;
    # the normal texture uses the other sampler (repeat); the baseline JPEG decodes; the progressive one
    # is skipped with a warning.
    m0, m1, m2 = s.material(0), s.material(1), s.material(2)
    assert (m0.base_color_tex, m0.emissive_tex, m0.metallic_roughness_tex, m0.normal_tex) == (0, 0, 1, 2)
    assert m1.base_color_tex == 3 and m2.base_color_tex == -1
    assert r.num_textures == 4
    jt = np.ctypeslib.as_array(r.tex_texels[3], (35,)).reshape(5, 7)
    want = np.asarray(Image.open(io.BytesIO(jpg.getvalue())).convert("RGBA")).astype(np.int32)
    assert np.abs(_channels(jt) - want).max() <= 4
    assert [r.tex_srgb[i] for i in range(3)] == [1, 0, 0] and [r.tex_repeat[i] for i in range(3)] == [0, 0, 1]
    for i in range(3):
        assert (r.tex_width[i], r.tex_height[i]) == (7, 5)
        texels = np.ctypeslib.as_array(r.tex_texels[i], (35,)).reshape(5, 7)
        assert np.array_equal(texels, _rgba_u32(img))
    assert m0.alpha_mode == 1 and abs(m0.alpha_cutoff - 0.25) < 1e-7 and m0.double_sided == 1 and m0.normal_scale == 0.5
    assert m1.alpha_mode == 2 and m1.alpha_cutoff == 0.5
    assert "warning: texture skipped" in capfd.readouterr().err
    s.build()


# ---- PNG ----------------------------------------------------------------------------------------------

def _png_chunk(t, body):
    return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)


def _pack_row(samples, depth):
    if depth == 8:
        return samples.astype(np.uint8).tobytes()
    if depth == 16:
        return samples.astype(">u2").tobytes()
    bits = np.unpackbits(samples.astype(np.uint8)[:, None], axis=1)[:, 8 - depth:].reshape(-1)
    return np.packbits(bits).tobytes()


def _filter_row(raw, prev, bpp, ftype):
    out = bytearray(len(raw))
    for i in range(len(raw)):
        a = raw[i - bpp] if i >= bpp else 0
        b = prev[i] if prev is not None else 0
        c = prev[i - bpp] if (prev is not None and i >= bpp) else 0
        if ftype == 0:
            pred = 0
        elif ftype == 1:
            pred = a
        elif ftype == 2:
            pred = b
        elif ftype == 3:
            pred = (a + b) >> 1
        else:
            p = a + b - c
            pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
            pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
        out[i] = (raw[i] - pred) & 0xFF
    return bytes(out)


def _encode_png(samples, colour, depth, interlace, g, palette=None, trns=None):
    """A PNG written from the format definition: samples (h, w, channels) at `depth` bits, every row with
    a random filter, optionally Adam7-interlaced."""
    h, w, nc = samples.shape
    bpp = max(1, nc * depth // 8)
    if interlace:
        passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
    else:
        passes = [(0, 0, 1, 1)]
    data = b""
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        prev = None
        for row in sub:
            raw = _pack_row(row.reshape(-1), depth)
            ft = int(g.integers(0, 5))
            data += bytes([ft]) + _filter_row(raw, prev, bpp, ft)
            prev = raw
    out = b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, colour, 0, 0, interlace))
    if palette is not None:
        out += _png_chunk(b"PLTE", palette.astype(np.uint8).tobytes())
    if trns is not None:
        out += _png_chunk(b"tRNS", trns)
    z = zlib.compress(data)
    # IDAT split across chunks.
    out += _png_chunk(b"IDAT", z[: len(z) // 2]) + _png_chunk(b"IDAT", z[len(z) // 2:])
    return out + _png_chunk(b"IEND", b"")


@pytest.mark.parametrize("interlace", [0, 1])
@pytest.mark.parametrize("colour,depth", [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8),
                                          (4, 8), (4, 16), (6, 8), (6, 16)])
def test_png_decoder_against_the_format(colour, depth, interlace):
    g = np.random.default_rng(colour * 100 + depth * 2 + interlace)
    h, w = int(g.integers(1, 19)), int(g.integers(1, 23))
    nc = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[colour]
    top = (1 << depth) - 1
    palette = trns = None
    if colour == 3:
        n = int(g.integers(1, top + 2))
        s = g.integers(0, n, (h, w, 1))
        palette = g.integers(0, 256, (n, 3))
        alpha = g.integers(0, 256, int(g.integers(0, n + 1)))
        trns = alpha.astype(np.uint8).tobytes()
        a = np.concatenate([alpha, np.full(n - len(alpha), 255)])
        want = np.concatenate([palette[s[..., 0]], a[s[..., 0]][..., None]], axis=2)
    else:
        s = g.integers(0, top + 1, (h, w, nc))
        to8 = (lambda v: v >> 8) if depth == 16 else (lambda v: v * (255 // top))
        if colour in (0, 4):
            grey = to8(s[..., 0])
            alpha = to8(s[..., 1]) if colour == 4 else np.full((h, w), 255)
            want = np.stack([grey, grey, grey, alpha], axis=2)
        else:
            alpha = to8(s[..., 3]) if colour == 6 else np.full((h, w), 255)
            want = np.concatenate([to8(s[..., :3]), alpha[..., None]], axis=2)
        if colour in (0, 2):  # a transparent colour key, present in the image
            key = s[int(g.integers(0, h)), int(g.integers(0, w))]
            trns = b"".join(struct.pack(">H", int(k)) for k in key)
            want[..., 3] = np.where(np.all(s == key, axis=2), 0, 255)
    png = _encode_png(s, colour, depth, interlace, g, palette, trns)
    assert np.array_equal(host.decode_image(png), _rgba_u32(want))


@pytest.mark.parametrize("mode", ["RGBA", "RGB", "L", "LA", "P", "1", "I;16"])
def test_png_decoder_against_pillow(mode):
    g = np.random.default_rng(len(mode))
    src = g.integers(0, 256, (13, 17, 4), dtype=np.uint8)
    im = Image.fromarray(src, "RGBA")
    if mode == "P":
        im = im.convert("RGB").quantize(37)
    elif mode == "I;16":
        im = Image.fromarray((g.integers(0, 65536, (13, 17))).astype(np.uint16))
    else:
        im = im.convert(mode)
    b = io.BytesIO()
    im.save(b, "PNG")
    got = host.decode_image(b.getvalue())
    if mode == "I;16":
        v = (np.asarray(im).astype(np.uint32) >> 8)
        want = np.stack([v, v, v, np.full_like(v, 255)], axis=2)
    else:
        want = np.asarray(Image.open(io.BytesIO(b.getvalue())).convert("RGBA"))
    assert np.array_equal(got, _rgba_u32(want))


def test_png_writer_round_trips_through_pillow(tmp_path):
    g = np.random.default_rng(2)
    for h, w in [(1, 1), (9, 31), (64, 3)]:
        px = g.integers(0, 2**32, (h, w), dtype=np.uint64).astype(np.uint32)
        if h == 64:
            px[:] = px[0]  # repeated rows exercise the "up" filter
        host.save_png(tmp_path / "o.png", px)
        im = Image.open(tmp_path / "o.png")
        assert im.mode == "RGBA" and im.size == (w, h)
        assert np.array_equal(_rgba_u32(np.asarray(im)), px)
        assert np.array_equal(host.decode_image((tmp_path / "o.png").read_bytes()), px)


def _channels(texels):
    t = np.asarray(texels, np.uint32)
    return np.stack([(t >> s) & 255 for s in (0, 8, 16, 24)], -1).astype(np.int32)


def _jpeg_test_image(h, w, seed):
    g = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    smooth = np.stack([(xx * 2) % 256, (yy * 3) % 256, ((xx + yy) * 2) % 256], -1)
    return np.clip(smooth * 0.7 + g.integers(0, 70, (h, w, 3)), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("subsampling", [0, 1, 2])  # 4:4:4, 4:2:2, 4:2:0
@pytest.mark.parametrize("quality", [30, 75, 95, 100])
def test_jpeg_decoder_against_pillow(subsampling, quality):
    """Baseline JPEG against Pillow's decoder (libjpeg): inverse DCTs and colour conversions differ in
    their last bits, so agreement is within a few code values, nearly everywhere exact."""
    img = _jpeg_test_image(77, 131, subsampling * 7 + quality)
    b = io.BytesIO()
    Image.fromarray(img).save(b, "JPEG", quality=quality, subsampling=subsampling)
    got = _channels(host.decode_image(b.getvalue()))
    want = np.asarray(Image.open(io.BytesIO(b.getvalue())).convert("RGBA")).astype(np.int32)
    d = np.abs(got - want)
    assert d.max() <= 4 and d.mean() <= 0.1, (d.max(), d.mean())


@pytest.mark.parametrize("kind", ["grey", "restart", "tiny", "odd"])
def test_jpeg_variants(kind):
    b = io.BytesIO()
    if kind == "grey":
        Image.fromarray(_jpeg_test_image(40, 57, 1)[..., 0]).save(b, "JPEG", quality=85)
    elif kind == "restart":
        Image.fromarray(_jpeg_test_image(64, 96, 2)).save(b, "JPEG", quality=80, restart_marker_blocks=3)
    elif kind == "tiny":
        Image.fromarray(_jpeg_test_image(1, 1, 3)).save(b, "JPEG", quality=90, subsampling=2)
    else:
        Image.fromarray(_jpeg_test_image(17, 33, 4)).save(b, "JPEG", quality=90, subsampling=2)
    got = _channels(host.decode_image(b.getvalue()))
    want = np.asarray(Image.open(io.BytesIO(b.getvalue())).convert("RGBA")).astype(np.int32)
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 4


def test_image_decode_rejects_other_formats():
    jpg = io.BytesIO()
    Image.fromarray(np.zeros((16, 16, 3), np.uint8)).save(jpg, "JPEG", progressive=True)
    with pytest.raises(RuntimeError, match="progressive JPEG"):
        host.decode_image(jpg.getvalue())
    cmyk = io.BytesIO()
    Image.fromarray(np.zeros((16, 16, 4), np.uint8), "CMYK").save(cmyk, "JPEG")
    with pytest.raises(RuntimeError, match="JPEG"):
        host.decode_image(cmyk.getvalue())
    with pytest.raises(RuntimeError, match="unsupported image format"):
        host.decode_image(b"GIF89a....")
    png = _encode_png(np.zeros((4, 4, 3), np.int64), 2, 8, 0, np.random.default_rng(0))
    with pytest.raises(RuntimeError, match="PNG"):
        host.decode_image(png[:60])



def _jpeg_segments(data):
    """(marker, offset of the segment's first payload byte, payload length) of each header segment before
    the first scan's entropy-coded data."""
    out, p = [], 2
    while p + 4 <= len(data):
        marker, length = data[p + 1], (data[p + 2] << 8) | data[p + 3]
        out.append((marker, p + 4, length - 2))
        if marker == 0xDA:
            break
        p += 2 + length
    return out


@pytest.mark.parametrize("field", ["scan_dc", "scan_ac", "dht", "dqt", "frame_tq"])
@pytest.mark.parametrize("selector", [4, 9, 15])
def test_jpeg_rejects_table_selectors_beyond_three(field, selector):
    """JPEG names at most four Huffman tables per class and four quantisation tables (T.81 B.2.2-B.2.4):
    a selector of 4 or more is refused with a message instead of indexing past the decoder's tables."""
    b = io.BytesIO()
    Image.fromarray(_jpeg_test_image(16, 16, 5)).save(b, "JPEG", quality=80)
    data = bytearray(b.getvalue())
    seg = {m: (o, n) for m, o, n in _jpeg_segments(data)}
    if field in ("scan_dc", "scan_ac"):
        o = seg[0xDA][0] + 2  # Cs, then Td/Ta of the scan's first component
        data[o] = (data[o] & 0x0F) | (selector << 4) if field == "scan_dc" else (data[o] & 0xF0) | selector
    elif field == "dht":
        o = seg[0xC4][0]
        data[o] = (data[o] & 0xF0) | selector
    elif field == "dqt":
        o = seg[0xDB][0]
        data[o] = (data[o] & 0xF0) | selector
    else:
        o = seg[0xC0][0] + 6 + 2  # P, Y, X, Nf, then C/HV/Tq of the first component
        data[o] = selector
    with pytest.raises(RuntimeError, match="selector"):
        host.decode_image(bytes(data))


def test_gltf_rejects_strides_whose_span_wraps(tmp_path):
    """An accessor's span stride*(count-1) + element must not wrap around 2^64: a stride beyond glTF's
    limit of 252 bytes is refused, and so is a span past the buffer's end."""
    pos = np.zeros(9, np.float32).tobytes()
    base = {"meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}], "nodes": [{"mesh": 0}], "scenes": [{"nodes": [0]}]}
    for i, (stride, count, msg) in enumerate([(2**39, 2**25 + 1, "byteStride"), (253, 3, "byteStride"),
                                               (8, 3, "byteStride"), (252, 3, "out of the buffer")]):
        body = dict(base, bufferViews=[{"buffer": 0, "byteLength": 36, "byteStride": stride}],
                    accessors=[{"bufferView": 0, "componentType": 5126, "count": count, "type": "VEC3"}])
        with pytest.raises(RuntimeError, match=msg):
            host.Scene.from_gltf(_write(tmp_path / f"s{i}.gltf", _doc_with_buffer(pos, body)))


# ---- PFM and HDR ----------------------------------------------------------------------------------------

def test_pfm_layout_and_round_trip(tmp_path):
    """save_pfm writes "PF\\n<w> <h>\\n-1.0\\n" and the rows bottom to top (src/scene/image.cpp:88-96)."""
    g = np.random.default_rng(6)
    img = g.normal(size=(5, 8, 3)).astype(np.float32)
    host.save_pfm(tmp_path / "a.pfm", img)
    data = (tmp_path / "a.pfm").read_bytes()
    header = b"PF\n8 5\n-1.0\n"
    assert data[: len(header)] == header
    assert data[len(header):] == img[::-1].astype("<f4").tobytes()
    assert np.array_equal(host.load_pfm(tmp_path / "a.pfm"), img)
    # Big-endian (positive scale) and single-channel files.
    (tmp_path / "b.pfm").write_bytes(b"PF\n8 5\n1.0\n" + img[::-1].astype(">f4").tobytes())
    assert np.array_equal(host.load_pfm(tmp_path / "b.pfm"), img)
    grey = img[..., 0].copy()
    (tmp_path / "c.pfm").write_bytes(b"Pf 8 5 -1\n" + grey[::-1].astype("<f4").tobytes())
    assert np.array_equal(host.load_pfm(tmp_path / "c.pfm"), np.repeat(grey[..., None], 3, axis=2))
    (tmp_path / "d.pfm").write_bytes(header + b"\0" * 10)
    with pytest.raises(RuntimeError, match="PFM"):
        host.load_pfm(tmp_path / "d.pfm")


def _rgbe(rgb):
    """Radiance encoding of linear RGB (exponent of the largest component)."""
    m = rgb.max(axis=2)
    e = np.where(m > 1e-32, np.floor(np.log2(np.maximum(m, 1e-32))) + 1, 0)
    scale = np.where(m > 1e-32, 256.0 / np.exp2(e), 0)
    mant = np.clip(np.floor(rgb * scale[..., None]), 0, 255)
    return np.concatenate([mant, np.where(m > 1e-32, e + 128, 0)[..., None]], axis=2).astype(np.uint8)


def _rle_plane(v):
    out, i = bytearray(), 0
    while i < len(v):
        j = i
        while j < len(v) and j - i < 127 and v[j] == v[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + j - i, v[i]])
            i = j
        else:
            k = i
            while k < len(v) and k - i < 128 and not (k + 2 < len(v) and v[k] == v[k + 1] == v[k + 2]):
                k += 1
            k = max(k, i + 1)
            out += bytes([k - i]) + bytes(v[i:k])
            i = k
    return bytes(out)


@pytest.mark.parametrize("rle", [False, True])
def test_load_environment_hdr(tmp_path, rle):
    g = np.random.default_rng(7)
    h, w = 6, 40
    rgb = np.exp(g.normal(0, 3, (h, w, 3)))
    rgb[2, 5:20] = rgb[2, 5]  # runs
    rgb[3, 0] = 0.0
    px = _rgbe(rgb)
    body = b""
    for y in range(h):
        if rle:
            body += bytes([2, 2, w >> 8, w & 255]) + b"".join(_rle_plane(px[y, :, c]) for c in range(4))
        else:
            body += px[y].tobytes()
    (tmp_path / "e.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1\n\n-Y %d +X %d\n" % (h, w) + body)
    s = host.Scene.builtin("cornell")
    s.load_environment(tmp_path / "e.hdr", intensity=2.0, rotation=0.5)
    r = s.raw()
    assert (r.env_width, r.env_height) == (w, h) and r.env_intensity == 2.0 and r.env_rotation == 0.5
    got = np.ctypeslib.as_array(r.env_texels, (h * w * 3,)).reshape(h, w, 3)
    e = px[..., 3:4].astype(np.int32)
    want = np.where(e > 0, px[..., :3] * np.exp2((e - 136).astype(np.float64)), 0).astype(np.float32)
    assert np.array_equal(got, want)
    s.build()


def test_load_environment_png_is_linearised(tmp_path):
    g = np.random.default_rng(9)
    img = g.integers(0, 256, (4, 8, 3), dtype=np.uint8)
    Image.fromarray(img, "RGB").save(tmp_path / "e.png")
    s = host.Scene.builtin("cornell")
    s.load_environment(tmp_path / "e.png")
    r = s.raw()
    got = np.ctypeslib.as_array(r.env_texels, (4 * 8 * 3,)).reshape(4, 8, 3)
    c = img.astype(np.float64) / 255.0
    want = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    assert np.abs(got - want).max() <= 1e-6
    with pytest.raises(RuntimeError, match="cannot open"):
        s.load_environment(tmp_path / "missing.hdr")


def test_gltf_writer_writes_each_primitive_its_own_vertices(tmp_path):
    """Meshes of interleaved materials whose vertex arrays hold unused vertices: each primitive carries
    only the vertices its triangles use, and the triangles come back bit for bit with their materials."""
    from photon_amd.types import Material
    g = np.random.default_rng(13)
    s = host.Scene()
    mats = [s.add_material(Material.default()) for _ in range(3)]
    p = g.uniform(-1, 1, (40, 3)).astype(np.float32)
    idx = g.integers(0, 40, (60, 3)).astype(np.int32)
    for t in range(60):
        s.add_mesh(p, idx[t], mats[t % 3])
    s.save_gltf(tmp_path / "shared.gltf")
    b = host.Scene.from_gltf(tmp_path / "shared.gltf")
    A, B = _raw_arrays(s), _raw_arrays(b)
    assert b.num_tris == 60 and len(B["pos"]) == len(np.unique(A["idx"]))
    for m in mats:
        ta = A["pos"][A["idx"][A["mat"] == m]].reshape(-1, 9)
        tb = B["pos"][B["idx"][B["mat"] == m]].reshape(-1, 9)
        assert np.array_equal(ta[np.lexsort(ta.T)], tb[np.lexsort(tb.T)])


def test_gltf_rejects_negative_and_huge_offsets(tmp_path):
    """Offsets, strides and counts are checked before any read: negative or out-of-range values fail with
    a message instead of wrapping around the buffer bounds."""
    pos = np.zeros(9, np.float32).tobytes()
    base = {"meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}], "nodes": [{"mesh": 0}], "scenes": [{"nodes": [0]}]}
    bad = {
        "byteOffset": dict(bufferViews=[{"buffer": 0, "byteOffset": -8, "byteLength": 36}],
                           accessors=[{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"}]),
        "byteStride": dict(bufferViews=[{"buffer": 0, "byteLength": 36, "byteStride": -12}],
                           accessors=[{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"}]),
        "accessor count": dict(bufferViews=[{"buffer": 0, "byteLength": 36}],
                               accessors=[{"bufferView": 0, "componentType": 5126, "count": -3, "type": "VEC3"}]),
    }
    for i, (msg, body) in enumerate(bad.items()):
        with pytest.raises(RuntimeError, match=msg):
            host.Scene.from_gltf(_write(tmp_path / f"w{i}.gltf", _doc_with_buffer(pos, dict(base, **body))))
    huge = dict(base, bufferViews=[{"buffer": 0, "byteLength": 36}],
                accessors=[{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3", "byteOffset": 2**62}])
    with pytest.raises(RuntimeError, match="byteOffset"):
        host.Scene.from_gltf(_write(tmp_path / "huge.gltf", _doc_with_buffer(pos, huge)))


def test_decoders_survive_corrupted_files(tmp_path):
    """Seeded corruptions (bytes changed, cut out or inserted) of valid PNG and JPEG files and of a glTF
    document either decode or fail with a message; none may crash the process."""
    g = np.random.default_rng(21)
    img = g.integers(0, 256, (23, 37, 3), dtype=np.uint8)
    seeds = []
    for kw in (dict(format="PNG"), dict(format="JPEG", quality=80), dict(format="JPEG", quality=60, subsampling=2),
               dict(format="JPEG", restart_marker_blocks=2)):
        b = io.BytesIO()
        Image.fromarray(img).save(b, **kw)
        seeds.append(b.getvalue())
    for it in range(400):
        s = bytearray(seeds[it % len(seeds)])
        for _ in range(int(g.integers(1, 8))):
            k, p = int(g.integers(0, 3)), int(g.integers(0, len(s)))
            if k == 0:
                s[p] = int(g.integers(0, 256))
            elif k == 1:
                del s[p:p + int(g.integers(1, 16))]
            else:
                s[p:p] = bytes(g.integers(0, 256, int(g.integers(1, 8))).astype(np.uint8))
        try:
            host.decode_image(bytes(s))
        except RuntimeError:
            pass
    host.Scene.builtin("cornell").save_gltf(tmp_path / "c.gltf")
    text = (tmp_path / "c.gltf").read_text()
    for it in range(100):
        t = list(text)
        for _ in range(int(g.integers(1, 6))):
            t[int(g.integers(0, len(t)))] = "0123456789-,:[]{}\"e. "[int(g.integers(0, 21))]
        (tmp_path / "m.gltf").write_text("".join(t))
        try:
            host.Scene.from_gltf(tmp_path / "m.gltf")
        except RuntimeError:
            pass


def test_gltf_data_uri_base64_forms(tmp_path):
    """A buffer's base64 with line breaks and spaces anywhere, or without its '=' padding, decodes to the
    same bytes; any other character is refused."""
    pos = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32).tobytes()
    idx = np.array([0, 1, 2], np.uint16).tobytes() + b"\0\0"  # 44 bytes: the encoding ends in '=' padding
    body = {"bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36}, {"buffer": 0, "byteOffset": 36, "byteLength": 6}],
            "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3", "min": [0, 0, 0], "max": [1, 1, 0]},
                          {"bufferView": 1, "componentType": 5123, "count": 3, "type": "SCALAR"}],
            "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}],
            "nodes": [{"mesh": 0}], "scenes": [{"nodes": [0]}], "scene": 0}
    enc = base64.b64encode(pos + idx).decode()
    assert enc.endswith("=")

    def load(text, name):
        doc = _doc_with_buffer(pos + idx, body)
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + text
        s = host.Scene.from_gltf(_write(tmp_path / name, doc))
        r = s.raw()
        return np.ctypeslib.as_array(r.positions, (r.num_vertices * 3,)).copy()

    want = load(enc, "plain.gltf")
    assert np.array_equal(want, np.frombuffer(pos, np.float32))
    spaced = "".join(c + ("\n" if k % 7 == 6 else " " if k % 11 == 10 else "") for k, c in enumerate(enc))
    assert np.array_equal(load(spaced, "spaced.gltf"), want)
    assert np.array_equal(load(enc.rstrip("="), "unpadded.gltf"), want)
    with pytest.raises(RuntimeError, match="invalid data URI"):
        load(enc[:20] + "*" + enc[20:], "bad.gltf")
