"""ctypes mirrors of include/photon_types.h (the plain-data layouts shared by the host producers,
the HIP device library and the CPU oracle). Sizes are asserted at import so a layout drift fails
loudly instead of corrupting memory."""
import ctypes as C

import numpy as np

F3 = C.c_float * 3
F4 = C.c_float * 4

PH_INF = 1e30

TONEMAP_LINEAR, TONEMAP_ACES, TONEMAP_REINHARD = 0, 1, 2
DISPLAY_COLOR, DISPLAY_ALBEDO, DISPLAY_NORMAL, DISPLAY_DEPTH, DISPLAY_VARIANCE, DISPLAY_HEATMAP = range(6)
ALPHA_OPAQUE, ALPHA_MASK, ALPHA_BLEND = 0, 1, 2
LIGHT_POINT, LIGHT_SPOT, LIGHT_DIRECTIONAL = 0, 1, 2


class BvhNode(C.Structure):
    _fields_ = [("lmin", F3), ("left", C.c_int32), ("lmax", F3), ("lcount", C.c_int32),
                ("rmin", F3), ("right", C.c_int32), ("rmax", F3), ("rcount", C.c_int32)]


class TriGeom(C.Structure):
    _fields_ = [("v0", F3), ("material", C.c_int32), ("e1", F3), ("flags", C.c_int32),
                ("e2", F3), ("pad", C.c_int32)]


class TextureInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("offset", C.c_uint32),
                ("wrap_repeat", C.c_int32), ("srgb", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("base_color", F4), ("emission", F3), ("metallic", C.c_float), ("roughness", C.c_float),
                ("ior", C.c_float), ("transmission", C.c_float), ("normal_scale", C.c_float),
                ("alpha_cutoff", C.c_float), ("alpha_mode", C.c_int32), ("double_sided", C.c_int32),
                ("base_color_tex", C.c_int32), ("metallic_roughness_tex", C.c_int32), ("normal_tex", C.c_int32),
                ("emissive_tex", C.c_int32), ("occlusion_tex", C.c_int32)]

    @classmethod
    def default(cls):
        """glTF defaults as in the reference Material (src/render/scene_view.h:46-62)."""
        m = cls()
        m.base_color[:] = (1.0, 1.0, 1.0, 1.0)
        m.metallic, m.roughness, m.ior, m.transmission = 1.0, 1.0, 1.5, 0.0
        m.normal_scale, m.alpha_cutoff, m.alpha_mode, m.double_sided = 1.0, 0.5, ALPHA_OPAQUE, 0
        m.base_color_tex = m.metallic_roughness_tex = m.normal_tex = m.emissive_tex = m.occlusion_tex = -1
        return m


class EmissiveTri(C.Structure):
    _fields_ = [("tri", C.c_int32), ("area", C.c_float)]


class AliasEntry(C.Structure):
    _fields_ = [("prob", C.c_float), ("alias", C.c_int32), ("pmf", C.c_float), ("pad", C.c_int32)]


class PunctualLight(C.Structure):
    _fields_ = [("position", F3), ("type", C.c_int32), ("direction", F3), ("cos_inner", C.c_float),
                ("intensity", F3), ("cos_outer", C.c_float)]


class CameraView(C.Structure):
    _fields_ = [("position", F3), ("tan_half_fov", C.c_float), ("forward", F3), ("aspect", C.c_float),
                ("right", F3), ("aperture", C.c_float), ("up", F3), ("focus_distance", C.c_float)]


class Ray(C.Structure):
    _fields_ = [("o", F3), ("d", F3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("tri", C.c_int32), ("u", C.c_float), ("v", C.c_float)]


class PathOptions(C.Structure):
    _fields_ = [("max_depth", C.c_int32), ("rr_depth", C.c_int32), ("clamp", C.c_float), ("nee", C.c_int32)]


class RenderSettings(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("path", PathOptions), ("exposure", C.c_float),
                ("tonemap", C.c_int32), ("seed", C.c_uint32)]

    @classmethod
    def make(cls, width=960, height=540, max_depth=8, rr_depth=3, clamp=0.0, nee=1, exposure=1.0,
             tonemap=TONEMAP_ACES, seed=0):
        """Defaults of the reference RenderSettings/PathOptions (pixel.h:10-16, integrator.h:21-27)."""
        return cls(width, height, PathOptions(max_depth, rr_depth, clamp, nee), exposure, tonemap, seed)


class SvgfSettings(C.Structure):
    _fields_ = [("atrous_iterations", C.c_int32), ("alpha_color", C.c_float), ("alpha_moments", C.c_float),
                ("phi_color", C.c_float), ("phi_normal", C.c_float), ("phi_depth", C.c_float)]

    @classmethod
    def make(cls, atrous_iterations=5, alpha_color=0.2, alpha_moments=0.2, phi_color=4.0, phi_normal=128.0,
             phi_depth=1.0):
        """Defaults of the reference SvgfSettings (gpu_renderer.h:18-25)."""
        return cls(atrous_iterations, alpha_color, alpha_moments, phi_color, phi_normal, phi_depth)


class SceneDesc(C.Structure):
    _fields_ = [
        ("nodes", C.POINTER(BvhNode)), ("num_nodes", C.c_int32),
        ("tris", C.POINTER(TriGeom)), ("num_tris", C.c_int32),
        ("indices", C.POINTER(C.c_int32)), ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
        ("tangents", C.POINTER(C.c_float)), ("num_vertices", C.c_int32),
        ("materials", C.POINTER(Material)), ("num_materials", C.c_int32),
        ("textures", C.POINTER(TextureInfo)), ("num_textures", C.c_int32),
        ("texels", C.POINTER(C.c_uint32)), ("num_texels", C.c_int64), ("srgb_lut", C.POINTER(C.c_float)),
        ("emissive", C.POINTER(EmissiveTri)), ("tri_emissive", C.POINTER(C.c_int32)),
        ("emissive_alias", C.POINTER(AliasEntry)), ("num_emissive", C.c_int32), ("emissive_power", C.c_float),
        ("punctual", C.POINTER(PunctualLight)), ("num_punctual", C.c_int32),
        ("env_texels", C.POINTER(C.c_float)), ("env_marginal", C.POINTER(C.c_float)),
        ("env_conditional", C.POINTER(C.c_float)), ("env_width", C.c_int32), ("env_height", C.c_int32),
        ("env_intensity", C.c_float), ("env_rotation", C.c_float), ("env_color", F3),
        ("env_marginal_integral", C.c_float),
        ("p_env", C.c_float), ("p_emissive", C.c_float), ("p_punctual", C.c_float),
        ("bounds_min", F3), ("bounds_max", F3),
    ]


class RawScene(C.Structure):
    _fields_ = [
        ("positions", C.POINTER(C.c_float)), ("num_vertices", C.c_int32),
        ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
        ("indices", C.POINTER(C.c_int32)), ("num_tris", C.c_int32), ("tri_material", C.POINTER(C.c_int32)),
        ("materials", C.POINTER(Material)), ("num_materials", C.c_int32), ("num_textures", C.c_int32),
        ("tex_width", C.POINTER(C.c_int32)), ("tex_height", C.POINTER(C.c_int32)),
        ("tex_texels", C.POINTER(C.POINTER(C.c_uint32))), ("tex_repeat", C.POINTER(C.c_int32)),
        ("tex_srgb", C.POINTER(C.c_int32)), ("punctual", C.POINTER(PunctualLight)), ("num_punctual", C.c_int32),
        ("env_texels", C.POINTER(C.c_float)), ("env_width", C.c_int32), ("env_height", C.c_int32),
        ("env_color", F3), ("env_intensity", C.c_float), ("env_rotation", C.c_float),
        ("cam_position", F3), ("cam_target", F3), ("cam_up", F3),
        ("cam_vfov", C.c_float), ("cam_aperture", C.c_float), ("cam_focus", C.c_float),
        ("tangents", C.POINTER(C.c_float)),
    ]


class BvhStats(C.Structure):
    _fields_ = [("nodes", C.c_int32), ("leaves", C.c_int32), ("max_depth", C.c_int32), ("sah_cost", C.c_float),
                ("build_ms", C.c_double)]


for _t, _n in ((BvhNode, 64), (TriGeom, 48), (TextureInfo, 20), (Material, 80), (EmissiveTri, 8), (AliasEntry, 16),
               (PunctualLight, 48), (CameraView, 64), (Ray, 28), (Hit, 16), (PathOptions, 16), (RenderSettings, 36),
               (SvgfSettings, 24)):
    assert C.sizeof(_t) == _n, (_t.__name__, C.sizeof(_t), _n)

# numpy views of the record types (for zero-copy array handling)
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("tri", "<i4"), ("u", "<f4"), ("v", "<f4")])
NODE_DTYPE = np.dtype([("lmin", "<f4", 3), ("left", "<i4"), ("lmax", "<f4", 3), ("lcount", "<i4"),
                       ("rmin", "<f4", 3), ("right", "<i4"), ("rmax", "<f4", 3), ("rcount", "<i4")])
assert RAY_DTYPE.itemsize == 28 and HIT_DTYPE.itemsize == 16 and NODE_DTYPE.itemsize == 64


def fptr(a):
    """float* of a contiguous float32 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def uptr(a):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def vptr(a):
    """void* of a contiguous numpy array (record arrays included)."""
    if a is None:
        return None
    assert a.flags.c_contiguous
    return C.c_void_p(a.ctypes.data)
