"""Python view of the host scene producers (include/photon_host.h).

`Scene` mirrors the reference's host Scene (src/scene/scene.h:41-95): build authored geometry, call
`build()` to get the BVH and light tables, then hand `desc()` to the device library or the oracle.
"""
import ctypes as C

import numpy as np

from . import _lib
from .types import (BvhStats, CameraView, Material, PunctualLight, RawScene, Ray, SceneDesc, RAY_DTYPE, NODE_DTYPE,
                    fptr, iptr, uptr, vptr)

_sig_done = False


def lib():
    global _sig_done
    L = _lib.load(_lib.HOST_LIB)
    if not _sig_done:
        P = C.c_void_p
        L.phs_last_error.restype = C.c_char_p
        L.phs_scene_create.argtypes = [C.POINTER(P)]
        L.phs_scene_destroy.argtypes = [P]
        L.phs_scene_destroy.restype = None
        L.phs_scene_builtin.argtypes = [P, C.c_char_p]
        L.phs_scene_spheres.argtypes = [P, C.c_int, C.c_float, C.c_float, C.c_uint64]
        L.phs_scene_field.argtypes = [P, C.c_int, C.c_uint64, C.c_int]
        L.phs_scene_add_mesh.argtypes = [P, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int32), C.c_int, C.c_int,
                                         C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.phs_scene_add_material.argtypes = [P, C.POINTER(Material)]
        L.phs_scene_set_material.argtypes = [P, C.c_int, C.POINTER(Material)]
        L.phs_scene_get_material.argtypes = [P, C.c_int, C.POINTER(Material)]
        L.phs_scene_add_texture.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_uint32), C.c_int, C.c_int]
        L.phs_scene_add_punctual.argtypes = [P, C.POINTER(PunctualLight)]
        L.phs_scene_set_env_color.argtypes = [P, C.POINTER(C.c_float), C.c_float]
        L.phs_scene_set_env_image.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_float, C.c_float]
        L.phs_scene_set_sky.argtypes = [P, C.POINTER(C.c_float), C.c_float, C.c_int]
        L.phs_scene_clear_normals.argtypes = [P]
        L.phs_scene_set_camera.argtypes = [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                           C.c_float, C.c_float, C.c_float]
        L.phs_scene_get_camera.argtypes = [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                           C.POINTER(C.c_float)]
        L.phs_scene_camera_view.argtypes = [P, C.c_float, C.POINTER(CameraView)]
        L.phs_scene_build.argtypes = [P, C.c_int, C.c_int, C.c_float, C.c_int]
        L.phs_scene_rebuild_lights.argtypes = [P]
        L.phs_scene_desc.argtypes = [P, C.POINTER(SceneDesc)]
        L.phs_scene_num_tris.argtypes = [P]
        L.phs_scene_num_materials.argtypes = [P]
        L.phs_scene_order.argtypes = [P]
        L.phs_scene_order.restype = C.POINTER(C.c_int32)
        L.phs_scene_stats.argtypes = [P, C.POINTER(BvhStats)]
        L.phs_scene_device_bytes.argtypes = [P]
        L.phs_scene_device_bytes.restype = C.c_int64
        L.phs_scene_raw.argtypes = [P, C.POINTER(RawScene)]
        L.phs_camera_rays.argtypes = [C.POINTER(CameraView), C.c_int, C.c_int, P]
        L.phs_sphere_shadow_rays.argtypes = [P, C.c_int, C.c_uint64, P]
        L.phs_scene_load_gltf.argtypes = [P, C.c_char_p]
        L.phs_scene_save_gltf.argtypes = [P, C.c_char_p]
        L.phs_scene_num_cameras.argtypes = [P]
        L.phs_scene_camera_at.argtypes = [P, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                          C.POINTER(C.c_float)]
        L.phs_scene_load_environment.argtypes = [P, C.c_char_p, C.c_float, C.c_float]
        L.phs_image_decode.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int), P, C.c_size_t]
        L.phs_image_save_png.argtypes = [C.c_char_p, C.c_int, C.c_int, P]
        L.phs_image_save_pfm.argtypes = [C.c_char_p, C.c_int, C.c_int, P]
        L.phs_image_load_pfm.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), P, C.c_size_t]
        L.phs_json_canonical.argtypes = [C.c_char_p, P, C.c_size_t]
        L.phs_json_canonical.restype = C.c_int64
        _sig_done = True
    return L


def _check(rc):
    if rc < 0:
        raise RuntimeError(lib().phs_last_error().decode())
    return rc


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


class Scene:
    """A host scene; `build()` produces the device inputs (BVH, BVH-order triangles, light tables)."""

    def __init__(self):
        self._h = C.c_void_p()
        _check(lib().phs_scene_create(C.byref(self._h)))
        self.built = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().phs_scene_destroy(h)
            self._h = None

    # -- construction ----------------------------------------------------------------------------
    @classmethod
    def builtin(cls, name):
        s = cls()
        _check(lib().phs_scene_builtin(s._h, name.encode()))
        return s

    @classmethod
    def spheres(cls, count, box=100.0, alpha_fraction=0.0, seed=1):
        s = cls()
        _check(lib().phs_scene_spheres(s._h, count, box, alpha_fraction, seed))
        return s

    @classmethod
    def field(cls, count=1024, seed=1, env_width=2048):
        s = cls()
        _check(lib().phs_scene_field(s._h, count, seed, env_width))
        return s

    def add_mesh(self, positions, indices, material, normals=None, uvs=None):
        p = np.ascontiguousarray(positions, np.float32).reshape(-1)
        i = np.ascontiguousarray(indices, np.int32).reshape(-1)
        n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1)
        u = None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1)
        _check(lib().phs_scene_add_mesh(self._h, fptr(p), p.size // 3, iptr(i), i.size, material, fptr(n), fptr(u)))
        self.built = False

    def add_material(self, m):
        return _check(lib().phs_scene_add_material(self._h, C.byref(m)))

    def set_material(self, index, m):
        _check(lib().phs_scene_set_material(self._h, index, C.byref(m)))

    def material(self, index):
        m = Material()
        _check(lib().phs_scene_get_material(self._h, index, C.byref(m)))
        return m

    def add_texture(self, texels_rgba8, repeat=True, srgb=False):
        t = np.ascontiguousarray(texels_rgba8, np.uint32)
        h, w = t.shape
        return _check(lib().phs_scene_add_texture(self._h, w, h, uptr(t), int(repeat), int(srgb)))

    def add_punctual(self, light):
        return _check(lib().phs_scene_add_punctual(self._h, C.byref(light)))

    def set_env_color(self, rgb, intensity=1.0):
        _check(lib().phs_scene_set_env_color(self._h, _f3(rgb), intensity))

    def set_env_image(self, rgb, intensity=1.0, rotation=0.0):
        a = np.ascontiguousarray(rgb, np.float32)
        h, w, _ = a.shape
        _check(lib().phs_scene_set_env_image(self._h, w, h, fptr(a.reshape(-1)), intensity, rotation))

    def set_sky(self, sun, sun_intensity=30000.0, width=1024):
        _check(lib().phs_scene_set_sky(self._h, _f3(sun), sun_intensity, width))

    def clear_normals(self):
        _check(lib().phs_scene_clear_normals(self._h))

    def set_camera(self, position, target, up=(0, 1, 0), vfov=45.0, aperture=0.0, focus_distance=0.0):
        _check(lib().phs_scene_set_camera(self._h, _f3(position), _f3(target), _f3(up), vfov, aperture, focus_distance))

    def camera(self):
        p, t, u, f = _f3((0, 0, 0)), _f3((0, 0, 0)), _f3((0, 0, 0)), C.c_float()
        _check(lib().phs_scene_get_camera(self._h, p, t, u, C.byref(f)))
        return dict(position=tuple(p), target=tuple(t), up=tuple(u), vfov=f.value)

    # -- scene files ---------------------------------------------------------------------------------
    @classmethod
    def from_gltf(cls, path):
        s = cls()
        s.load_gltf(path)
        return s

    def load_gltf(self, path):
        """Appends a .gltf / .glb file (src/scene/gltf.cpp:486-497)."""
        _check(lib().phs_scene_load_gltf(self._h, str(path).encode()))
        self.built = False

    def save_gltf(self, path):
        _check(lib().phs_scene_save_gltf(self._h, str(path).encode()))

    def cameras(self):
        out = []
        for i in range(_check(lib().phs_scene_num_cameras(self._h))):
            p, t, u, f = _f3((0, 0, 0)), _f3((0, 0, 0)), _f3((0, 0, 0)), C.c_float()
            _check(lib().phs_scene_camera_at(self._h, i, p, t, u, C.byref(f)))
            out.append(dict(position=tuple(p), target=tuple(t), up=tuple(u), vfov=f.value))
        return out

    def load_environment(self, path, intensity=1.0, rotation=0.0):
        _check(lib().phs_scene_load_environment(self._h, str(path).encode(), intensity, rotation))
        self.built = False

    def camera_view(self, aspect):
        v = CameraView()
        _check(lib().phs_scene_camera_view(self._h, aspect, C.byref(v)))
        return v

    # -- build ------------------------------------------------------------------------------------
    def build(self, bins=0, max_leaf=0, traversal_cost=0.0, threads=0):
        _check(lib().phs_scene_build(self._h, bins, max_leaf, traversal_cost, threads))
        self.built = True
        return self

    def rebuild_lights(self):
        _check(lib().phs_scene_rebuild_lights(self._h))

    def desc(self):
        d = SceneDesc()
        _check(lib().phs_scene_desc(self._h, C.byref(d)))
        return d

    def raw(self):
        r = RawScene()
        _check(lib().phs_scene_raw(self._h, C.byref(r)))
        return r

    @property
    def num_tris(self):
        return _check(lib().phs_scene_num_tris(self._h))

    @property
    def num_materials(self):
        return _check(lib().phs_scene_num_materials(self._h))

    def order(self):
        p = lib().phs_scene_order(self._h)
        return np.ctypeslib.as_array(p, (self.num_tris,)).copy()

    def stats(self):
        s = BvhStats()
        _check(lib().phs_scene_stats(self._h, C.byref(s)))
        return s

    def device_bytes(self):
        return _check(lib().phs_scene_device_bytes(self._h))

    def nodes(self):
        d = self.desc()
        return np.frombuffer((C.c_uint8 * (64 * d.num_nodes)).from_address(C.addressof(d.nodes.contents)),
                             dtype=NODE_DTYPE).copy()

    def shadow_rays(self, count, seed=1):
        out = np.zeros(count, RAY_DTYPE)
        _check(lib().phs_sphere_shadow_rays(self._h, count, seed, vptr(out)))
        return out


def camera_rays(view, width, height):
    """Pixel-centre pinhole rays, row-major (the traversal workload's camera rays)."""
    out = np.zeros(width * height, RAY_DTYPE)
    _check(lib().phs_camera_rays(C.byref(view), width, height, vptr(out)))
    return out


# -- images and JSON -------------------------------------------------------------------------------------
def decode_image(data):
    """PNG bytes -> (h, w) uint32 RGBA8 texels, R in the low byte (src/scene/image.cpp:23-37)."""
    data = bytes(data)
    w, h = C.c_int(), C.c_int()
    _check(lib().phs_image_decode(data, len(data), C.byref(w), C.byref(h), None, 0))
    out = np.zeros((h.value, w.value), np.uint32)
    _check(lib().phs_image_decode(data, len(data), C.byref(w), C.byref(h), vptr(out), out.size))
    return out


def save_png(path, rgba):
    a = np.ascontiguousarray(rgba, np.uint32)
    _check(lib().phs_image_save_png(str(path).encode(), a.shape[1], a.shape[0], vptr(a)))


def save_pfm(path, rgb):
    a = np.ascontiguousarray(rgb, np.float32)
    _check(lib().phs_image_save_pfm(str(path).encode(), a.shape[1], a.shape[0], vptr(a)))


def load_pfm(path):
    w, h = C.c_int(), C.c_int()
    _check(lib().phs_image_load_pfm(str(path).encode(), C.byref(w), C.byref(h), None, 0))
    out = np.zeros((h.value, w.value, 3), np.float32)
    _check(lib().phs_image_load_pfm(str(path).encode(), C.byref(w), C.byref(h), vptr(out), out.size))
    return out


def json_canonical(text):
    """Parses JSON text with the library's reader and returns its canonical form."""
    b = text.encode() if isinstance(text, str) else bytes(text)
    n = _check(lib().phs_json_canonical(b, None, 0))
    buf = C.create_string_buffer(n + 1)
    _check(lib().phs_json_canonical(b, buf, n + 1))
    return buf.raw[:n].decode()
