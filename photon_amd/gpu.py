"""Python mirror of the reference's GPU renderer API over the HIP device library (include/photon_hip.h).

`GpuRenderer` follows the reference GpuRenderer (src/render/gpu_renderer.h:36-75) method for method:
upload returns (ok, error) like `bool upload(const Scene&, std::string& error)`, the other methods raise
`RuntimeError` where the reference throws. `Svgf` follows the reference Svgf (src/denoise/svgf.h:32-59).
The native library is required: a missing library raises, there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _lib
from .types import (CameraView, PathOptions, RenderSettings, SceneDesc, SvgfSettings, DISPLAY_COLOR)

ENGINE_MEGAKERNEL, ENGINE_WAVEFRONT = 0, 1
BUF_ACCUM, BUF_ALBEDO, BUF_NORMAL_DEPTH, BUF_COLOR, BUF_POSITION, BUF_EMISSION, BUF_DENOISED, BUF_HEAT, BUF_VARIANCE = range(9)


class GpuInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("memory", C.c_int64), ("compute_units", C.c_int32),
                ("wavefront_size", C.c_int32), ("lds_per_block", C.c_int32), ("clock_khz", C.c_int32),
                ("memory_clock_khz", C.c_int32), ("memory_bus_width", C.c_int32)]


class WavefrontConfig(C.Structure):
    _fields_ = [("block", C.c_int32), ("stack_depth", C.c_int32), ("path_slots", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("sort_materials", C.c_int32), ("use_graph", C.c_int32), ("graph_iterations", C.c_int32),
                ("refill_lanes", C.c_int32), ("lds_nodes", C.c_int32), ("ray_bins", C.c_int32), ("xcd_parts", C.c_int32),
                ("sample_split", C.c_int32), ("finish_percent", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RenderStats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_extend", C.c_double), ("ms_shadow", C.c_double), ("ms_shade", C.c_double),
                ("ms_other", C.c_double), ("extend_launches", C.c_int64), ("shadow_launches", C.c_int64),
                ("shade_launches", C.c_int64), ("iterations", C.c_int64), ("rays_extend", C.c_int64),
                ("rays_shadow", C.c_int64), ("samples", C.c_int64), ("node_visits", C.c_int64), ("tri_tests", C.c_int64),
                ("alpha_tests", C.c_int64), ("shadow_node_visits", C.c_int64), ("shadow_tri_tests", C.c_int64),
                ("shadow_alpha_tests", C.c_int64), ("finish_paths", C.c_int64), ("rays_finish", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_int64), ("path_slots", C.c_int32), ("iterations_per_step", C.c_int32),
                ("max_jobs", C.c_int32), ("pass_samples", C.c_int64), ("schedule", C.c_int32), ("min_items", C.c_int32),
                ("express_samples", C.c_int64), ("express_slots", C.c_int32), ("express_iterations", C.c_int32),
                ("deadline_stretch", C.c_double)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("camera", CameraView), ("path", PathOptions), ("seed", C.c_uint32), ("frame0", C.c_uint32)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("pixels", C.c_int32), ("samples_done", C.c_int64), ("samples", C.c_int64),
                ("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_RELEASED = range(4)
SCHEDULE_SRPT, SCHEDULE_FIFO, SCHEDULE_HRRN, SCHEDULE_SHARE, SCHEDULE_DEADLINE = 0, 1, 2, 3, 4


class SvgfInput(C.Structure):
    _fields_ = [("color_dev", C.c_void_p), ("albedo_dev", C.c_void_p), ("normal_depth_dev", C.c_void_p),
                ("position_dev", C.c_void_p), ("emission_dev", C.c_void_p)]


class BvhBuildStats(C.Structure):
    _fields_ = [("nodes", C.c_int32), ("leaves", C.c_int32), ("max_depth", C.c_int32), ("sah_cost", C.c_float),
                ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3), ("build_ms", C.c_double)]


class CompactGBufferDesc(C.Structure):
    _fields_ = [("color_dev", C.c_void_p), ("albedo_dev", C.c_void_p), ("emission_dev", C.c_void_p),
                ("depth_dev", C.c_void_p), ("normal_dev", C.c_void_p), ("material_dev", C.c_void_p)]


_sig_done = False


def lib():
    global _sig_done
    L = _lib.load(_lib.HIP_LIB)
    if not _sig_done:
        P, R = C.c_void_p, C.c_void_p
        CV, RS = C.POINTER(CameraView), C.POINTER(RenderSettings)
        L.ph_last_error.restype = C.c_char_p
        L.ph_device_count.restype = C.c_int
        L.ph_device_info.argtypes = [C.c_int, C.POINTER(GpuInfo)]
        L.ph_renderer_create.argtypes = [C.POINTER(R)]
        L.ph_renderer_destroy.argtypes = [R]
        L.ph_renderer_destroy.restype = None
        L.ph_renderer_upload.argtypes = [R, C.POINTER(SceneDesc)]
        L.ph_renderer_update_lights.argtypes = [R, C.POINTER(SceneDesc)]
        L.ph_renderer_resize.argtypes = [R, C.c_int, C.c_int]
        L.ph_renderer_reset.argtypes = [R]
        L.ph_renderer_render.argtypes = [R, CV, RS, C.c_int]
        L.ph_renderer_render_denoised.argtypes = [R, CV, CV, RS, C.POINTER(SvgfSettings)]
        L.ph_renderer_display.argtypes = [R, RS, C.c_int, P]
        L.ph_renderer_download_image.argtypes = [R, P]
        L.ph_renderer_download_display.argtypes = [R, P]
        for f in ("width", "height", "samples"):
            getattr(L, "ph_renderer_" + f).argtypes = [R]
        L.ph_renderer_last_frame_ms.argtypes = [R]
        L.ph_renderer_last_frame_ms.restype = C.c_float
        L.ph_renderer_display_buffer.argtypes = [R]
        L.ph_renderer_display_buffer.restype = C.c_void_p
        L.ph_renderer_set_engine.argtypes = [R, C.c_int]
        L.ph_renderer_get_engine.argtypes = [R]
        L.ph_renderer_set_wavefront_config.argtypes = [R, C.POINTER(WavefrontConfig)]
        L.ph_renderer_get_wavefront_config.argtypes = [R, C.POINTER(WavefrontConfig)]
        L.ph_renderer_autotune.argtypes = [R, CV, RS, C.c_int, C.POINTER(WavefrontConfig)]
        L.ph_renderer_set_counting.argtypes = [R, C.c_int]
        L.ph_renderer_stats.argtypes = [R, C.POINTER(RenderStats)]
        L.ph_renderer_buffer.argtypes = [R, C.c_int]
        L.ph_renderer_buffer.restype = C.c_void_p
        L.ph_trace.argtypes = [R, P, C.c_int, C.c_int, P, P]
        L.ph_trace_config.argtypes = [R, P, C.c_int, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]
        L.ph_svgf_create.argtypes = [C.POINTER(P)]
        L.ph_svgf_destroy.argtypes = [P]
        L.ph_svgf_destroy.restype = None
        L.ph_svgf_resize.argtypes = [P, C.c_int, C.c_int]
        L.ph_svgf_reset.argtypes = [P]
        L.ph_svgf_run.argtypes = [P, C.POINTER(SvgfInput), CV, C.POINTER(SvgfSettings), P]
        L.ph_bvh_build.argtypes = [P, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_float, P, C.c_int, P, C.POINTER(BvhBuildStats)]
        L.ph_svgf_run_compact.argtypes = [P, C.POINTER(CompactGBufferDesc), CV, CV, C.POINTER(SvgfSettings), P]
        L.ph_renderer_sample_compact.argtypes = [R, CV, RS, C.POINTER(CompactGBufferDesc)]
        L.ph_svgf_variance_buffer.argtypes = [P]
        L.ph_svgf_variance_buffer.restype = C.c_void_p
        L.ph_svgf_last_ms.argtypes = [P]
        L.ph_svgf_last_ms.restype = C.c_float
        L.ph_device_alloc.argtypes = [C.c_int64, C.POINTER(P)]
        L.ph_device_free.argtypes = [P]
        L.ph_host_alloc.argtypes = [C.c_int64, C.POINTER(P)]
        L.ph_host_free.argtypes = [P]
        L.ph_pointer_kind.argtypes = [P, C.POINTER(C.c_int)]
        L.ph_copy_to_device.argtypes = [P, P, C.c_int64]
        L.ph_copy_to_host.argtypes = [P, P, C.c_int64]
        L.ph_device_memset.argtypes = [P, C.c_int, C.c_int64]
        L.ph_copy_device.argtypes = [P, P, C.c_int64]
        L.ph_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(P)]
        L.ph_server_destroy.argtypes = [P]
        L.ph_server_destroy.restype = None
        L.ph_server_set_wavefront_config.argtypes = [P, C.POINTER(WavefrontConfig)]
        L.ph_server_add_scene.argtypes = [P, C.POINTER(SceneDesc)]
        L.ph_server_submit.argtypes = [P, C.POINTER(JobDesc)]
        L.ph_server_step.argtypes = [P]
        L.ph_server_status.argtypes = [P, C.c_int, C.POINTER(JobStatus)]
        L.ph_server_result.argtypes = [P, C.c_int, P]
        L.ph_server_release.argtypes = [P, C.c_int]
        L.ph_server_pending.argtypes = [P]
        L.ph_server_now_ms.argtypes = [P]
        L.ph_server_now_ms.restype = C.c_double
        _sig_done = True
    return L


POINTER_UNUSABLE, POINTER_DEVICE, POINTER_PINNED_HOST, POINTER_MANAGED = 0, 1, 2, 3


def pointer_kind(ptr):
    """What memory `ptr` is as kernels see it (POINTER_*); pageable host memory is POINTER_UNUSABLE."""
    k = C.c_int(0)
    _check(lib().ph_pointer_kind(ptr, C.byref(k)))
    return k.value


class PinnedArray:
    """Pinned host memory allocated by the library (ph_host_alloc): kernels write it directly, and numpy
    reads it in place."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _check(lib().ph_host_alloc(self.nbytes, C.byref(self.ptr)))

    def numpy(self, dtype, shape=None):
        dtype = np.dtype(dtype)
        buf = (C.c_uint8 * self.nbytes).from_address(self.ptr.value)
        out = np.frombuffer(buf, dtype=dtype, count=self.nbytes // dtype.itemsize)
        return out if shape is None else out.reshape(shape)

    def free(self):
        if self.ptr:
            lib().ph_host_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """A device allocation made by the library (no torch needed), with numpy round trips."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _check(lib().ph_device_alloc(self.nbytes, C.byref(self.ptr)))

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _check(lib().ph_copy_to_device(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def numpy(self, dtype, shape=None):
        dtype = np.dtype(dtype)
        out = np.empty(self.nbytes // dtype.itemsize, dtype)
        _check(lib().ph_copy_to_host(out.ctypes.data, self.ptr, out.nbytes))
        return out if shape is None else out.reshape(shape)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().ph_copy_to_device(self.ptr, a.ctypes.data, a.nbytes))

    def zero(self):
        _check(lib().ph_device_memset(self.ptr, 0, self.nbytes))

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            lib().ph_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_device_to_numpy(ptr, nbytes, dtype, shape=None):
    dtype = np.dtype(dtype)
    out = np.empty(int(nbytes) // dtype.itemsize, dtype)
    _check(lib().ph_copy_to_host(out.ctypes.data, C.c_void_p(ptr), out.nbytes))
    return out if shape is None else out.reshape(shape)


def _check(rc):
    if rc < 0:
        raise RuntimeError(lib().ph_last_error().decode())
    return rc


def gpu_available(info=None):
    """cuda_available(GpuInfo*) of the reference (gpu_renderer.h:34): True when a device is present."""
    n = lib().ph_device_count()
    if n <= 0:
        return False
    if info is not None:
        _check(lib().ph_device_info(0, C.byref(info)))
    return True


def device_info(device=0):
    info = GpuInfo()
    _check(lib().ph_device_info(device, C.byref(info)))
    return info


class GpuRenderer:
    def __init__(self):
        self._h = C.c_void_p()
        _check(lib().ph_renderer_create(C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None):
            lib().ph_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def upload(self, scene):
        """Returns (ok, error) like the reference's bool upload(const Scene&, std::string&)."""
        d = scene.desc() if hasattr(scene, "desc") else scene
        if lib().ph_renderer_upload(self._h, C.byref(d)) < 0:
            return False, lib().ph_last_error().decode()
        return True, ""

    def update_lights(self, scene):
        d = scene.desc() if hasattr(scene, "desc") else scene
        if lib().ph_renderer_update_lights(self._h, C.byref(d)) < 0:
            return False, lib().ph_last_error().decode()
        return True, ""

    def resize(self, width, height):
        _check(lib().ph_renderer_resize(self._h, width, height))

    def reset(self):
        _check(lib().ph_renderer_reset(self._h))

    def render(self, camera, settings, spp):
        _check(lib().ph_renderer_render(self._h, C.byref(camera), C.byref(settings), spp))

    def render_denoised(self, camera, previous, settings, svgf=None):
        svgf = svgf or SvgfSettings.make()
        _check(lib().ph_renderer_render_denoised(self._h, C.byref(camera), C.byref(previous), C.byref(settings), C.byref(svgf)))

    def sample_compact(self, camera, settings, gbuffer):
        """One sample per pixel into a CompactGBuffer (ph_renderer_sample_compact)."""
        _check(lib().ph_renderer_sample_compact(self._h, C.byref(camera), C.byref(settings), C.byref(gbuffer.desc())))

    def display(self, settings, mode=DISPLAY_COLOR, target=None):
        _check(lib().ph_renderer_display(self._h, C.byref(settings), mode, target))

    def download_image(self):
        out = np.zeros((self.height, self.width, 3), np.float32)
        _check(lib().ph_renderer_download_image(self._h, out.ctypes.data))
        return out

    def download_display(self):
        out = np.zeros((self.height, self.width), np.uint32)
        _check(lib().ph_renderer_download_display(self._h, out.ctypes.data))
        return out

    @property
    def width(self):
        return lib().ph_renderer_width(self._h)

    @property
    def height(self):
        return lib().ph_renderer_height(self._h)

    @property
    def samples(self):
        return lib().ph_renderer_samples(self._h)

    def last_frame_ms(self):
        return lib().ph_renderer_last_frame_ms(self._h)

    def display_buffer(self):
        return lib().ph_renderer_display_buffer(self._h)

    # -- engine, tuning, statistics ---------------------------------------------------------------
    def set_engine(self, engine):
        _check(lib().ph_renderer_set_engine(self._h, engine))

    @property
    def engine(self):
        return lib().ph_renderer_get_engine(self._h)

    def wavefront_config(self):
        c = WavefrontConfig()
        _check(lib().ph_renderer_get_wavefront_config(self._h, C.byref(c)))
        return c

    def set_wavefront_config(self, cfg=None, **kw):
        c = cfg if cfg is not None else self.wavefront_config()
        for k, v in kw.items():
            setattr(c, k, v)
        _check(lib().ph_renderer_set_wavefront_config(self._h, C.byref(c)))
        return c

    def autotune(self, camera, settings, spp=1):
        c = WavefrontConfig()
        _check(lib().ph_renderer_autotune(self._h, C.byref(camera), C.byref(settings), spp, C.byref(c)))
        return c

    def set_counting(self, enable):
        _check(lib().ph_renderer_set_counting(self._h, int(enable)))

    def stats(self):
        s = RenderStats()
        _check(lib().ph_renderer_stats(self._h, C.byref(s)))
        return s

    def buffer_ptr(self, which):
        return lib().ph_renderer_buffer(self._h, which)

    # -- ray buffers --------------------------------------------------------------------------------
    def trace(self, rays_dev, n, hits_dev, anyhit=False, counts_dev=None):
        _check(lib().ph_trace(self._h, rays_dev, n, int(anyhit), hits_dev, counts_dev))

    def trace_config(self, rays_dev, n, hits_dev, anyhit=False, block=128, stack_depth=16, persistent=True, lds_nodes=0):
        ms = C.c_float()
        _check(lib().ph_trace_config(self._h, rays_dev, n, int(anyhit), hits_dev, block, stack_depth, int(persistent), lds_nodes, C.byref(ms)))
        return ms.value


def bvh_build(positions, indices, bins=0, max_leaf=0, traversal_cost=0.0):
    """Binned-SAH BVH of triangles (ph_bvh_build, on the device): the host builder's nodes, triangle
    order and statistics. positions: (V, 3) float32; indices: (T, 3) int32."""
    from .types import NODE_DTYPE
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
    n = len(i)
    cap = max(1, 2 * n - 1)
    nodes = np.zeros(cap, NODE_DTYPE)
    order = np.zeros(n, np.int32)
    st = BvhBuildStats()
    count = _check(lib().ph_bvh_build(p.ctypes.data if len(p) else None, len(p), i.ctypes.data if n else None, n, bins, max_leaf,
                                      traversal_cost, nodes.ctypes.data, cap, order.ctypes.data if n else None, C.byref(st)))
    return nodes[:count], order, st


class CompactGBuffer:
    """Device planes of the compact G-buffer (ph_compact_gbuffer): 3-float colour, albedo and emission,
    float depth, uint32 octahedral fp16 normal, int32 material."""

    PLANES = (("color", np.float32, 3), ("albedo", np.float32, 3), ("emission", np.float32, 3),
              ("depth", np.float32, 1), ("normal", np.uint32, 1), ("material", np.int32, 1))

    def __init__(self, width, height):
        self.width, self.height = width, height
        self.planes = {name: DeviceArray(width * height * k * np.dtype(dt).itemsize) for name, dt, k in self.PLANES}

    @classmethod
    def from_numpy(cls, planes, width, height):
        g = cls(width, height)
        for name, dt, k in cls.PLANES:
            g.planes[name].upload(np.asarray(planes[name], dt))
        return g

    def desc(self):
        p = self.planes
        return CompactGBufferDesc(p["color"].ptr, p["albedo"].ptr, p["emission"].ptr, p["depth"].ptr, p["normal"].ptr, p["material"].ptr)

    def numpy(self):
        out = {}
        for name, dt, k in self.PLANES:
            shape = (self.height, self.width, k) if k > 1 else (self.height, self.width)
            out[name] = self.planes[name].numpy(dt, shape)
        return out


class Svgf:
    def __init__(self):
        self._h = C.c_void_p()
        _check(lib().ph_svgf_create(C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ph_svgf_destroy(self._h)
            self._h = None

    def resize(self, width, height):
        _check(lib().ph_svgf_resize(self._h, width, height))

    def reset(self):
        _check(lib().ph_svgf_reset(self._h))

    def run(self, color_dev, albedo_dev, normal_depth_dev, position_dev, emission_dev, previous_camera, settings, output_dev):
        inp = SvgfInput(color_dev, albedo_dev, normal_depth_dev, position_dev, emission_dev)
        _check(lib().ph_svgf_run(self._h, C.byref(inp), C.byref(previous_camera), C.byref(settings), output_dev))

    def run_compact(self, gbuffer, camera, previous_camera, settings, output_dev):
        """The denoiser on a CompactGBuffer traced from `camera` (ph_svgf_run_compact)."""
        _check(lib().ph_svgf_run_compact(self._h, C.byref(gbuffer.desc()), C.byref(camera), C.byref(previous_camera), C.byref(settings),
                                         output_dev))

    def variance_buffer(self):
        return lib().ph_svgf_variance_buffer(self._h)

    def last_ms(self):
        return lib().ph_svgf_last_ms(self._h)


class Server:
    """The serving loop (ph_server_*): resident scenes, jobs admitted by HBM footprint, continuous batching."""

    def __init__(self, memory_budget=0, path_slots=0, iterations_per_step=0, max_jobs=0, pass_samples=0, schedule=SCHEDULE_SRPT, min_items=0,
                 express_samples=0, express_slots=0, express_iterations=0, deadline_stretch=0.0):
        self._h = C.c_void_p()
        cfg = ServerConfig(memory_budget, path_slots, iterations_per_step, max_jobs, pass_samples, schedule, min_items,
                           express_samples, express_slots, express_iterations, deadline_stretch)
        _check(lib().ph_server_create(C.byref(cfg), C.byref(self._h)))
        self._sizes = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().ph_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_wavefront_config(self, cfg):
        _check(lib().ph_server_set_wavefront_config(self._h, C.byref(cfg)))

    def add_scene(self, scene):
        d = scene.desc() if hasattr(scene, "desc") else scene
        return _check(lib().ph_server_add_scene(self._h, C.byref(d)))

    def submit(self, scene, width, height, spp, camera, max_depth=8, rr_depth=3, clamp=0.0, nee=1, seed=0, frame0=0):
        job = JobDesc(scene, width, height, spp, camera, PathOptions(max_depth, rr_depth, clamp, nee), seed, frame0)
        jid = _check(lib().ph_server_submit(self._h, C.byref(job)))
        self._sizes[jid] = (height, width)
        return jid

    def step(self):
        return _check(lib().ph_server_step(self._h))

    def status(self, job):
        st = JobStatus()
        _check(lib().ph_server_status(self._h, job, C.byref(st)))
        return st

    def result(self, job):
        h, w = self._sizes[job]
        out = np.zeros((h, w, 3), np.float32)
        _check(lib().ph_server_result(self._h, job, out.ctypes.data))
        return out

    def release(self, job):
        _check(lib().ph_server_release(self._h, job))

    @property
    def pending(self):
        return lib().ph_server_pending(self._h)

    def now_ms(self):
        return lib().ph_server_now_ms(self._h)

    def run_until_idle(self, max_steps=1 << 20):
        for _ in range(max_steps):
            if self.pending == 0:
                return
            self.step()
        raise RuntimeError("server did not drain")
