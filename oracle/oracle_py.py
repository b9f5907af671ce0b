"""ORACLE — test infrastructure only (see oracle/__init__.py).

ctypes bindings of liboracle.so (the CPU restatement) and of _ref/libphoton_ref.so (the reference's
own CPU build, present only where the reference checkout was available at build time)."""
import ctypes as C
import os
import subprocess

import numpy as np

from photon_amd.types import (AliasEntry, BvhStats, CameraView, Hit, Ray, RawScene, RenderSettings, SceneDesc,
                              SvgfSettings, HIT_DTYPE, RAY_DTYPE, NODE_DTYPE, fptr, iptr, uptr, vptr)

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libphoton_ref.so")

_libs = {}
P = C.c_void_p
FP = C.POINTER(C.c_float)


def build(ref=False):
    """Compile the oracle (and the reference CPU build when asked and available)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def oracle():
    L = _libs.get("oracle")
    if L is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        L = C.CDLL(ORACLE_LIB)
        D, CV, RS = C.POINTER(SceneDesc), C.POINTER(CameraView), C.POINTER(RenderSettings)
        L.or_render.argtypes = [D, CV, RS, C.c_int, C.c_uint32, C.c_int, FP, FP, FP]
        L.or_render_samples.argtypes = [D, CV, RS, C.c_int, P, P, P, FP, P]
        L.or_render_gbuffer.argtypes = [D, CV, RS, C.c_uint32, C.c_int, FP, FP, FP, FP, FP]
        L.or_trace.argtypes = [D, P, C.c_int, C.c_int, P, P, C.c_int]
        L.or_heatmap.argtypes = [D, CV, C.c_int, C.c_int, FP]
        L.or_display_color.argtypes = [FP, C.c_int, C.c_int, C.c_float, C.c_int, P]
        L.or_display.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, FP, FP, FP, FP, FP, FP, P]
        L.or_rng_u32.argtypes = [C.c_uint64, C.c_uint64, C.c_int, P]
        L.or_rng_float.argtypes = [C.c_uint64, C.c_uint64, C.c_int, FP]
        L.or_hash_u32.argtypes = [P, C.c_int, P]
        L.or_warp.argtypes = [C.c_int, C.c_int, FP, FP]
        L.or_frame.argtypes = [C.c_int, FP, FP, FP, FP, FP]
        L.or_offset_ray.argtypes = [C.c_int, FP, FP, FP]
        L.or_bsdf_sample.argtypes = [FP, C.c_int, FP, FP, FP, FP, FP, P]
        L.or_bsdf_eval.argtypes = [FP, C.c_int, FP, FP, FP, FP]
        L.or_bsdf_info.argtypes = [FP, FP]
        L.or_sample_light.argtypes = [D, C.c_int, FP, FP, FP]
        L.or_env_eval.argtypes = [D, C.c_int, FP, FP, FP]
        L.or_env_sample.argtypes = [D, C.c_int, FP, FP, FP, P]
        L.or_emissive_pdf.argtypes = [D, C.c_int, P, FP, FP, FP]
        L.or_hit_emission.argtypes = [D, C.c_int, P, P, FP]
        L.or_tonemap.argtypes = [C.c_int, FP, C.c_float, C.c_int, FP, P]
        L.or_srgb.argtypes = [C.c_int, FP, C.c_int, FP]
        L.or_svgf_create.argtypes = [C.c_int, C.c_int, C.POINTER(P)]
        L.or_svgf_destroy.argtypes = [P]
        L.or_svgf_destroy.restype = None
        L.or_svgf_reset.argtypes = [P]
        L.or_svgf_reset.restype = None
        L.or_svgf_run.argtypes = [P, FP, FP, FP, FP, FP, CV, C.POINTER(SvgfSettings), FP]
        L.or_svgf_plane.argtypes = [P, C.c_int]
        L.or_svgf_atrous.argtypes = [P, FP, FP, C.c_int, C.c_float, C.c_float, C.c_float, FP, FP]
        L.or_svgf_plane.restype = FP
        L.or_f32_to_f16.argtypes = [C.c_float]
        L.or_f32_to_f16.restype = C.c_uint32
        L.or_f16_to_f32.argtypes = [C.c_uint32]
        L.or_f16_to_f32.restype = C.c_float
        L.or_oct_encode.argtypes = [FP]
        L.or_oct_encode.restype = C.c_uint32
        L.or_oct_decode.argtypes = [C.c_uint32, FP]
        L.or_oct_decode.restype = None
        L.or_gbuffer_compact.argtypes = [C.c_int, FP, FP, FP, FP, FP, FP, FP, FP, FP, P, P]
        L.or_svgf_run_compact.argtypes = [P, FP, FP, FP, FP, P, P, CV, CV, C.POINTER(SvgfSettings), FP]
        _libs["oracle"] = L
    return L


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    L = _libs.get("ref")
    if L is None:
        if not ref_available():
            raise FileNotFoundError(REF_LIB)
        L = C.CDLL(REF_LIB)
        D, CV, RS = C.POINTER(SceneDesc), C.POINTER(CameraView), C.POINTER(RenderSettings)
        L.ref_scene_builtin.argtypes = [C.c_char_p, C.POINTER(P)]
        L.ref_scene_from_raw.argtypes = [C.POINTER(RawScene), C.POINTER(P)]
        L.ref_scene_destroy.argtypes = [P]
        L.ref_scene_destroy.restype = None
        L.ref_scene_clear_normals.argtypes = [P]
        L.ref_scene_build.argtypes = [P, C.c_int]
        L.ref_scene_desc.argtypes = [P, D]
        L.ref_scene_order.argtypes = [P]
        L.ref_scene_order.restype = C.POINTER(C.c_int32)
        L.ref_scene_stats.argtypes = [P, C.POINTER(BvhStats)]
        L.ref_camera_view.argtypes = [P, C.c_float, CV]
        L.ref_render.argtypes = [P, CV, RS, C.c_int, C.c_int, FP]
        L.ref_render_gbuffer.argtypes = [P, CV, RS, C.c_uint32, FP, FP, FP, FP, FP]
        L.ref_trace.argtypes = [P, P, C.c_int, C.c_int, P]
        L.ref_alias_table.argtypes = [FP, C.c_int, P]
        _libs["ref"] = L
    return L


# ---- oracle helpers ------------------------------------------------------------------------------

def render(desc, cam, rs, spp, frame0=0, threads=0, gbuffer=False):
    """Average radiance (h, w, 3) of `spp` samples, as CpuRenderer::image() after one render call."""
    w, h = rs.width, rs.height
    s = np.zeros(w * h * 3, np.float32)
    alb = np.zeros(w * h * 4, np.float32) if gbuffer else None
    nd = np.zeros(w * h * 4, np.float32) if gbuffer else None
    rc = oracle().or_render(C.byref(desc), C.byref(cam), C.byref(rs), spp, frame0, threads, fptr(s), fptr(alb), fptr(nd))
    assert rc == 0
    inv = np.float32(1.0) / np.float32(spp) if spp > 0 else np.float32(0)
    img = (s * inv).reshape(h, w, 3)
    if gbuffer:
        return img, alb.reshape(h, w, 4), nd.reshape(h, w, 4)
    return img


def render_samples(desc, cam, rs, xs, ys, frames):
    """Radiance (n, 3) of single samples at pixels (xs, ys) with the given frame indices, and the rays
    each path traced."""
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    fr = np.ascontiguousarray(frames, np.uint32)
    n = len(xs)
    rgb = np.zeros(n * 3, np.float32)
    rays = np.zeros(n, np.uint32)
    rc = oracle().or_render_samples(C.byref(desc), C.byref(cam), C.byref(rs), n, vptr(xs), vptr(ys), vptr(fr), fptr(rgb), vptr(rays))
    assert rc == 0
    return rgb.reshape(n, 3), rays


def render_sum(desc, cam, rs, spp, frame0=0, threads=0):
    s = np.zeros(rs.width * rs.height * 3, np.float32)
    assert oracle().or_render(C.byref(desc), C.byref(cam), C.byref(rs), spp, frame0, threads, fptr(s), None, None) == 0
    return s.reshape(rs.height, rs.width, 3)


def display(mode, n, exposure=1.0, tonemap=0, samples=1, max_depth=1.0, heat_scale=1.0 / 160.0, denoised=False,
            accum=None, denoised_rgb=None, albedo=None, nd=None, variance=None, heat=None):
    """display_kernel's packed RGBA8 pixels for `n` pixels of float4 planes (absent planes read as zero)."""
    def plane(a, width):
        if a is None:
            return np.zeros(n * width, np.float32)
        a = np.ascontiguousarray(a, np.float32).reshape(-1)
        assert a.size == n * width
        return a
    planes = [plane(accum, 4), plane(denoised_rgb, 4), plane(albedo, 4), plane(nd, 4), plane(variance, 4), plane(heat, 1)]
    out = np.zeros(n, np.uint32)
    rc = oracle().or_display(n, mode, exposure, tonemap, samples, max_depth, heat_scale, int(denoised), *[fptr(p) for p in planes],
                             vptr(out))
    assert rc == 0
    return out


def render_gbuffer(desc, cam, rs, frame, threads=0):
    n = rs.width * rs.height
    planes = [np.zeros(n * 4, np.float32) for _ in range(5)]
    assert oracle().or_render_gbuffer(C.byref(desc), C.byref(cam), C.byref(rs), frame, threads, *[fptr(p) for p in planes]) == 0
    return [p.reshape(rs.height, rs.width, 4) for p in planes]


def trace(desc, rays, anyhit=False, counters=False, threads=0):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(rays), HIT_DTYPE)
    cnt = np.zeros((len(rays), 3), np.int64) if counters else None
    rc = oracle().or_trace(C.byref(desc), vptr(rays), len(rays), int(anyhit), vptr(hits), vptr(cnt), threads)
    assert rc == 0
    return (hits, cnt) if counters else hits


def f32_to_f16(x):
    return int(oracle().or_f32_to_f16(float(x)))


def f16_to_f32(h):
    return float(oracle().or_f16_to_f32(int(h)))


def oct_encode(n):
    v = np.ascontiguousarray(n, np.float32)
    return int(oracle().or_oct_encode(fptr(v)))


def oct_decode(code):
    out = np.zeros(3, np.float32)
    oracle().or_oct_decode(int(code), fptr(out))
    return out


def compact_gbuffer(color, albedo, normal_depth, position, emission):
    """The compact G-buffer (3-float colour/albedo/emission, depth, octahedral fp16 normal, material)
    of the reference's five float4 planes."""
    planes = [np.ascontiguousarray(a, np.float32) for a in (color, albedo, normal_depth, position, emission)]
    h, w = planes[0].shape[:2]
    n = w * h
    out = dict(color=np.zeros((h, w, 3), np.float32), albedo=np.zeros((h, w, 3), np.float32), emission=np.zeros((h, w, 3), np.float32),
               depth=np.zeros((h, w), np.float32), normal=np.zeros((h, w), np.uint32), material=np.zeros((h, w), np.int32))
    rc = oracle().or_gbuffer_compact(n, *[fptr(p) for p in planes], fptr(out["color"]), fptr(out["albedo"]), fptr(out["emission"]),
                                     fptr(out["depth"]), out["normal"].ctypes.data, out["material"].ctypes.data)
    assert rc == 0
    return out


class SvgfOracle:
    def __init__(self, width, height):
        self.w, self.h = width, height
        self._h = C.c_void_p()
        assert oracle().or_svgf_create(width, height, C.byref(self._h)) == 0

    def __del__(self):
        if getattr(self, "_h", None):
            oracle().or_svgf_destroy(self._h)
            self._h = None

    def reset(self):
        oracle().or_svgf_reset(self._h)

    def run(self, color, albedo, normal_depth, position, emission, prev_cam, settings):
        planes = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in (color, albedo, normal_depth, position, emission)]
        out = np.zeros(self.w * self.h * 4, np.float32)
        rc = oracle().or_svgf_run(self._h, *[fptr(p) for p in planes], C.byref(prev_cam), C.byref(settings), fptr(out))
        assert rc == 0
        return out.reshape(self.h, self.w, 4)

    def run_compact(self, g, cam, prev_cam, settings):
        """The denoiser on a compact G-buffer (a dict as compact_gbuffer returns) traced from `cam`."""
        c = {k: np.ascontiguousarray(g[k]).reshape(-1) for k in ("color", "albedo", "emission", "depth", "normal", "material")}
        out = np.zeros(self.w * self.h * 4, np.float32)
        rc = oracle().or_svgf_run_compact(self._h, fptr(c["color"]), fptr(c["albedo"]), fptr(c["emission"]), fptr(c["depth"]),
                                          c["normal"].ctypes.data, c["material"].ctypes.data, C.byref(cam), C.byref(prev_cam),
                                          C.byref(settings), fptr(out))
        assert rc == 0
        return out.reshape(self.h, self.w, 4)

    def atrous(self, normal_depth, position, src, step, phi_color, phi_normal, phi_depth):
        """One a-trous pass alone over (h, w, 4) planes."""
        nd, pos, sp = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in (normal_depth, position, src)]
        out = np.zeros(self.w * self.h * 4, np.float32)
        assert oracle().or_svgf_atrous(self._h, fptr(nd), fptr(pos), step, phi_color, phi_normal, phi_depth, fptr(sp), fptr(out)) == 0
        return out.reshape(self.h, self.w, 4)

    def plane(self, which):
        p = oracle().or_svgf_plane(self._h, which)
        return np.ctypeslib.as_array(p, (self.h, self.w, 4)).copy()


# ---- reference build helpers -------------------------------------------------------------------

class RefScene:
    """The reference's own Scene, built from the same authored arrays (or a built-in name)."""

    def __init__(self, raw=None, builtin=None):
        self._h = C.c_void_p()
        if builtin is not None:
            assert ref().ref_scene_builtin(builtin.encode(), C.byref(self._h)) == 0
        else:
            assert ref().ref_scene_from_raw(C.byref(raw), C.byref(self._h)) == 0

    def __del__(self):
        if getattr(self, "_h", None):
            ref().ref_scene_destroy(self._h)
            self._h = None

    def clear_normals(self):
        ref().ref_scene_clear_normals(self._h)

    def build(self, threads=0):
        ref().ref_scene_build(self._h, threads)
        return self

    def desc(self):
        d = SceneDesc()
        ref().ref_scene_desc(self._h, C.byref(d))
        return d

    def order(self):
        d = self.desc()
        return np.ctypeslib.as_array(ref().ref_scene_order(self._h), (d.num_tris,)).copy()

    def stats(self):
        s = BvhStats()
        ref().ref_scene_stats(self._h, C.byref(s))
        return s

    def camera_view(self, aspect):
        v = CameraView()
        ref().ref_camera_view(self._h, aspect, C.byref(v))
        return v

    def render(self, cam, rs, spp, threads=0):
        out = np.zeros(rs.width * rs.height * 3, np.float32)
        ref().ref_render(self._h, C.byref(cam), C.byref(rs), spp, threads, fptr(out))
        return out.reshape(rs.height, rs.width, 3)

    def render_gbuffer(self, cam, rs, frame):
        n = rs.width * rs.height
        planes = [np.zeros(n * 4, np.float32) for _ in range(5)]
        ref().ref_render_gbuffer(self._h, C.byref(cam), C.byref(rs), frame, *[fptr(p) for p in planes])
        return [p.reshape(rs.height, rs.width, 4) for p in planes]

    def trace(self, rays, anyhit=False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        ref().ref_trace(self._h, vptr(rays), len(rays), int(anyhit), vptr(hits))
        return hits


def desc_arrays(desc):
    """Copies of the flat arrays of a scene descriptor, as numpy arrays keyed by name."""
    def arr(ptr, n, dtype, width=1):
        if not ptr or n == 0:
            return np.zeros(0, dtype)
        nbytes = n * width * np.dtype(dtype).itemsize
        buf = (C.c_uint8 * nbytes).from_address(C.cast(ptr, C.c_void_p).value)
        return np.frombuffer(buf, dtype=dtype).copy()
    d = desc
    out = {
        "nodes": arr(d.nodes, d.num_nodes, NODE_DTYPE),
        "tris": arr(d.tris, d.num_tris, np.uint8, 48).reshape(-1, 48),
        "indices": arr(d.indices, d.num_tris * 3, np.int32),
        "materials": arr(d.materials, d.num_materials, np.uint8, 80).reshape(-1, 80),
        "emissive": arr(d.emissive, d.num_emissive, np.uint8, 8).reshape(-1, 8),
        "tri_emissive": arr(d.tri_emissive, d.num_tris, np.int32),
        "alias": arr(d.emissive_alias, d.num_emissive, np.uint8, 16).reshape(-1, 16),
        "srgb_lut": arr(d.srgb_lut, 256, np.float32),
        "texels": arr(d.texels, d.num_texels, np.uint32),
        "textures": arr(d.textures, d.num_textures, np.uint8, 20).reshape(-1, 20),
        "env_marginal": arr(d.env_marginal, d.env_height + 1 if d.env_marginal else 0, np.float32),
        "env_conditional": arr(d.env_conditional, d.env_height * (d.env_width + 1) if d.env_conditional else 0, np.float32),
        "scalars": np.array([d.emissive_power, d.env_marginal_integral, d.p_env, d.p_emissive, d.p_punctual,
                             *d.bounds_min, *d.bounds_max], np.float32),
    }
    if d.tangents:
        out["tangents"] = arr(d.tangents, d.num_vertices * 4, np.float32)
    if d.normals:
        out["normals"] = arr(d.normals, d.num_vertices * 3, np.float32)
    if d.uvs:
        out["uvs"] = arr(d.uvs, d.num_vertices * 2, np.float32)
    if d.punctual and d.num_punctual:
        out["punctual"] = arr(d.punctual, d.num_punctual, np.uint8, 48).reshape(-1, 48)
    return out
