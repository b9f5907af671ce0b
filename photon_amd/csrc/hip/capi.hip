// C ABI of the device library (include/photon_hip.h): the renderer object that owns the device copy of
// a scene and the frame buffers, with the reference GpuRenderer's semantics
// (src/render/gpu_renderer.cu:243-452): upload/update_lights copy the flat scene arrays, render adds
// samples to the accumulation, render_denoised runs one sample plus SVGF, display tone-maps into RGBA8,
// and each call blocks until the device is done. All work is issued on one HIP stream per renderer.
#include <chrono>
#include <cstring>
#include <memory>

#include "device_scene.h"
#include "internal.h"
#include "svgf_state.h"
#include "wavefront.h"

using namespace phd;
using namespace phk;

namespace {
thread_local std::string g_error;
int fail(const std::string& msg) { return phk::set_error(msg); }
}  // namespace

int phk::set_error(const std::string& msg) {
    g_error = msg;
    return -1;
}

struct ph_svgf {
    SvgfState state;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0;
    ph_svgf() {
        PH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        PH_CHECK(hipEventCreate(&ev0));
        PH_CHECK(hipEventCreate(&ev1));
    }
    ~ph_svgf() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct ph_renderer {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_mid = nullptr;

    // The scene in HBM, its device view and its material -> shading-queue table.
    DeviceScene scene;
    DScene view{};
    DevBuf<int> mat_class;

    void refresh_view() {
        view = scene.view();
        std::vector<int> cls = material_classes(view, stream);
        mat_class.upload(cls.data(), cls.size(), stream);
        PH_CHECK(hipStreamSynchronize(stream));
    }
    float scene_extent = 1;
    bool uploaded = false;

    // Frame buffers.
    int width = 0, height = 0, samples = 0;
    uint32_t frame = 0;
    float last_ms = 0;
    bool denoised = false, heat_valid = false;
    DevBuf<float4> accum, color, albedo, nd, position, emission, denoised_buf;
    DevBuf<float> heat;
    DevBuf<uint32_t> display;
    ph_camera_view last_camera{};
    SvgfState svgf;

    int engine = PH_ENGINE_WAVEFRONT;
    std::unique_ptr<WavefrontEngine> wavefront;
    ph_render_stats stats{};
    bool counting = false;

    ph_renderer() {
        PH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        PH_CHECK(hipEventCreate(&ev0));
        PH_CHECK(hipEventCreate(&ev1));
        PH_CHECK(hipEventCreate(&ev_mid));
        wavefront = std::make_unique<WavefrontEngine>();
    }
    ~ph_renderer() {
        wavefront.reset();
        if (stream) (void)hipStreamSynchronize(stream);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ev_mid) (void)hipEventDestroy(ev_mid);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void reset() {
        samples = 0;
        frame = 0;
        denoised = false;
        heat_valid = false;
        accum.zero(stream);
        svgf.has_history = false;
        PH_CHECK(hipStreamSynchronize(stream));
    }
};

namespace {

PathOpts opts_of(const ph_render_settings& rs) { return PathOpts{rs.path.max_depth, rs.path.rr_depth, rs.path.clamp, rs.path.nee}; }

}  // namespace

extern "C" {

const char* ph_last_error(void) { return g_error.c_str(); }

int ph_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int ph_device_info(int device, ph_gpu_info* out) {
    if (!out) return fail("null argument");
    return guarded([&] {
        hipDeviceProp_t p{};
        PH_CHECK(hipGetDeviceProperties(&p, device));
        std::memset(out, 0, sizeof(*out));
        std::strncpy(out->name, p.name, sizeof(out->name) - 1);
        std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
        out->memory = int64_t(p.totalGlobalMem);
        out->compute_units = p.multiProcessorCount;
        out->wavefront_size = p.warpSize;
        out->lds_per_block = int(p.sharedMemPerBlock);
        out->clock_khz = p.clockRate;
        out->memory_clock_khz = p.memoryClockRate;
        out->memory_bus_width = p.memoryBusWidth;
        return 0;
    });
}

int ph_device_alloc(int64_t bytes, void** out) {
    if (!out || bytes < 0) return fail("invalid argument");
    return guarded([&] {
        *out = nullptr;
        if (bytes > 0) PH_CHECK(hipMalloc(out, size_t(bytes)));
        return 0;
    });
}

int ph_device_free(void* p) {
    return guarded([&] {
        if (p) PH_CHECK(hipFree(p));
        return 0;
    });
}

int ph_host_alloc(int64_t bytes, void** out) {
    if (!out || bytes < 0) return fail("invalid argument");
    return guarded([&] {
        *out = nullptr;
        if (bytes > 0) PH_CHECK(hipHostMalloc(out, size_t(bytes), hipHostMallocDefault));
        return 0;
    });
}

int ph_pointer_kind(const void* p, int* kind) {
    if (!kind) return fail("null argument");
    *kind = PH_POINTER_UNUSABLE;
    if (!p) return 0;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // an unknown (pageable) pointer: clear the sticky error it leaves
        return 0;
    }
    if (a.type == hipMemoryTypeDevice) *kind = PH_POINTER_DEVICE;
    else if (a.type == hipMemoryTypeHost && a.devicePointer) *kind = PH_POINTER_PINNED_HOST;
    else if (a.type == hipMemoryTypeManaged) *kind = PH_POINTER_MANAGED;
    return 0;
}

namespace {
// A caller's target of `bytes` bytes that kernels may write: device, pinned host or managed memory,
// and for device memory, an allocation that holds all of it. Anything else (a pageable host pointer)
// is refused before a launch could fault on it.
void check_target(const void* p, size_t bytes, const char* what) {
    int kind = PH_POINTER_UNUSABLE;
    if (ph_pointer_kind(p, &kind) != 0 || kind == PH_POINTER_UNUSABLE)
        throw std::runtime_error(std::string(what) + " must be device memory or pinned host memory (ph_host_alloc)");
    if (kind == PH_POINTER_DEVICE) {
        hipDeviceptr_t base = nullptr;
        size_t size = 0;
        if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) == hipSuccess) {
            const size_t off = size_t(static_cast<const char*>(p) - static_cast<const char*>(base));
            if (off > size || size - off < bytes) throw std::runtime_error(std::string(what) + " is smaller than the frame");
        } else {
            (void)hipGetLastError();
        }
    }
}
}  // namespace

int ph_host_free(void* p) {
    return guarded([&] {
        if (p) PH_CHECK(hipHostFree(p));
        return 0;
    });
}

int ph_copy_to_device(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return fail("invalid argument");
    return guarded([&] {
        if (bytes) PH_CHECK(hipMemcpy(dst, src, size_t(bytes), hipMemcpyHostToDevice));
        return 0;
    });
}

int ph_copy_to_host(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return fail("invalid argument");
    return guarded([&] {
        if (bytes) PH_CHECK(hipMemcpy(dst, src, size_t(bytes), hipMemcpyDeviceToHost));
        return 0;
    });
}

int ph_copy_device(void* dst_dev, const void* src_dev, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst_dev || !src_dev))) return fail("invalid argument");
    return guarded([&] {
        if (bytes) PH_CHECK(hipMemcpy(dst_dev, src_dev, size_t(bytes), hipMemcpyDeviceToDevice));
        return 0;
    });
}

int ph_device_memset(void* dev, int value, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && !dev)) return fail("invalid argument");
    return guarded([&] {
        if (bytes) PH_CHECK(hipMemset(dev, value, size_t(bytes)));
        return 0;
    });
}

int ph_renderer_create(ph_renderer** out) {
    if (!out) return fail("null argument");
    return guarded([&] {
        *out = new ph_renderer();
        return 0;
    });
}

void ph_renderer_destroy(ph_renderer* r) { delete r; }

int ph_renderer_upload(ph_renderer* r, const ph_scene_desc* d) {
    if (!r) return fail("null renderer");
    return guarded([&] {
        DeviceScene::validate(d);
        ph_renderer& m = *r;
        m.uploaded = false;
        m.scene.upload(*d, m.stream);
        m.refresh_view();
        m.scene_extent = m.scene.extent();
        m.uploaded = true;
        m.reset();
        return 0;
    });
}

int ph_renderer_update_lights(ph_renderer* r, const ph_scene_desc* d) {
    if (!r) return fail("null renderer");
    if (!r->uploaded) return ph_renderer_upload(r, d);
    return guarded([&] {
        DeviceScene::validate(d);
        r->scene.upload_lights(*d, r->stream);
        PH_CHECK(hipStreamSynchronize(r->stream));
        r->refresh_view();
        r->reset();
        return 0;
    });
}

int ph_renderer_resize(ph_renderer* r, int width, int height) {
    if (!r) return fail("null renderer");
    if (width < 0 || height < 0) return fail("negative size");
    return guarded([&] {
        ph_renderer& m = *r;
        if (width == m.width && height == m.height && m.accum.count) return 0;
        m.width = width;
        m.height = height;
        const size_t n = size_t(width) * size_t(height);
        for (DevBuf<float4>* b : {&m.accum, &m.color, &m.albedo, &m.nd, &m.position, &m.emission, &m.denoised_buf}) {
            b->allocate(n);
            b->zero(m.stream);
        }
        m.heat.allocate(n);
        m.display.allocate(n);
        m.display.zero(m.stream);
        svgf_resize(m.svgf, width, height);
        m.reset();
        return 0;
    });
}

int ph_renderer_reset(ph_renderer* r) {
    if (!r) return fail("null renderer");
    return guarded([&] {
        r->reset();
        return 0;
    });
}

int ph_renderer_render(ph_renderer* r, const ph_camera_view* cam, const ph_render_settings* rs, int spp) {
    if (!r || !cam || !rs) return fail("null argument");
    if (spp < 0) return fail("negative sample count");
    return guarded([&] {
        ph_renderer& m = *r;
        if (!m.uploaded || m.width == 0 || m.height == 0 || spp == 0) return 0;
        if (m.denoised) m.reset();
        auto t0 = std::chrono::steady_clock::now();
        PathOpts opt = opts_of(*rs);
        if (m.engine == PH_ENGINE_MEGAKERNEL) {
            PH_CHECK(hipEventRecord(m.ev0, m.stream));
            launch_render(m.stream, m.view, *cam, m.width, m.height, uint32_t(m.samples), spp, rs->seed, opt, m.accum.ptr, m.albedo.ptr, m.nd.ptr);
            PH_CHECK(hipEventRecord(m.ev1, m.stream));
            PH_CHECK(hipEventSynchronize(m.ev1));
            PH_CHECK(hipEventElapsedTime(&m.last_ms, m.ev0, m.ev1));
            m.stats = ph_render_stats{};
            m.stats.samples = int64_t(spp) * m.width * m.height;
        } else {
            WavefrontJob job{};
            job.scene = m.view;
            job.camera = *cam;
            job.width = m.width;
            job.height = m.height;
            job.spp = spp;
            job.frame0 = uint32_t(m.samples);
            job.seed = rs->seed;
            job.opt = opt;
            job.accum = m.accum.ptr;
            job.albedo = m.albedo.ptr;
            job.nd = m.nd.ptr;
            job.mat_class = m.mat_class.ptr;
            m.wavefront->render(m.stream, job, m.counting, m.stats);
            m.last_ms = float(m.stats.ms_total);
        }
        m.stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        m.samples += spp;
        m.last_camera = *cam;
        m.heat_valid = false;
        return 0;
    });
}

int ph_renderer_render_denoised(ph_renderer* r, const ph_camera_view* cam, const ph_camera_view* prev, const ph_render_settings* rs,
                                const ph_svgf_settings* sv) {
    if (!r || !cam || !prev || !rs || !sv) return fail("null argument");
    return guarded([&] {
        ph_renderer& m = *r;
        if (!m.uploaded || m.width == 0 || m.height == 0) return 0;
        if (!m.denoised) {
            m.samples = 0;
            m.svgf.has_history = false;
        }
        m.denoised = true;
        PH_CHECK(hipEventRecord(m.ev0, m.stream));
        launch_sample(m.stream, m.view, *cam, m.width, m.height, m.frame, rs->seed, opts_of(*rs), m.color.ptr, m.albedo.ptr, m.nd.ptr, m.position.ptr,
                      m.emission.ptr);
        PH_CHECK(hipEventRecord(m.ev_mid, m.stream));
        ph_svgf_input in{&m.color.ptr->x, &m.albedo.ptr->x, &m.nd.ptr->x, &m.position.ptr->x, &m.emission.ptr->x};
        svgf_run(m.svgf, m.stream, in, *prev, *sv, m.denoised_buf.ptr);
        PH_CHECK(hipEventRecord(m.ev1, m.stream));
        PH_CHECK(hipEventSynchronize(m.ev1));
        PH_CHECK(hipEventElapsedTime(&m.last_ms, m.ev0, m.ev1));
        float trace_ms = 0, denoise_ms = 0;
        PH_CHECK(hipEventElapsedTime(&trace_ms, m.ev0, m.ev_mid));
        PH_CHECK(hipEventElapsedTime(&denoise_ms, m.ev_mid, m.ev1));
        m.stats = ph_render_stats{};
        m.stats.ms_total = m.last_ms;
        m.stats.ms_shade = trace_ms;
        m.stats.ms_other = denoise_ms;
        m.stats.samples = int64_t(m.width) * m.height;
        m.frame++;
        m.samples = 1;
        m.last_camera = *cam;
        m.heat_valid = false;
        return 0;
    });
}

int ph_renderer_sample_compact(ph_renderer* r, const ph_camera_view* cam, const ph_render_settings* rs, const ph_compact_gbuffer* out) {
    if (!r || !cam || !rs || !out) return fail("null argument");
    if (!out->color_dev || !out->albedo_dev || !out->emission_dev || !out->depth_dev || !out->normal_dev || !out->material_dev)
        return fail("missing G-buffer plane");
    return guarded([&] {
        ph_renderer& m = *r;
        if (!m.uploaded || m.width == 0 || m.height == 0) return 0;
        PH_CHECK(hipEventRecord(m.ev0, m.stream));
        launch_sample_compact(m.stream, m.view, *cam, m.width, m.height, m.frame, rs->seed, opts_of(*rs), *out);
        PH_CHECK(hipEventRecord(m.ev1, m.stream));
        PH_CHECK(hipEventSynchronize(m.ev1));
        PH_CHECK(hipEventElapsedTime(&m.last_ms, m.ev0, m.ev1));
        m.frame++;
        m.last_camera = *cam;
        return 0;
    });
}

int ph_renderer_display(ph_renderer* r, const ph_render_settings* rs, int mode, uint32_t* target) {
    if (!r || !rs) return fail("null argument");
    if (mode < PH_DISPLAY_COLOR || mode > PH_DISPLAY_HEATMAP) return fail("unknown display mode");
    return guarded([&] {
        ph_renderer& m = *r;
        if (m.width == 0 || m.height == 0) return 0;
        if (target) check_target(target, size_t(m.width) * size_t(m.height) * sizeof(uint32_t), "display target");
        if (mode == PH_DISPLAY_HEATMAP && !m.heat_valid && m.uploaded) {
            launch_heatmap(m.stream, m.view, m.last_camera, m.width, m.height, m.heat.ptr);
            m.heat_valid = true;
        }
        DisplayParams p;
        p.w = m.width;
        p.h = m.height;
        p.mode = mode;
        p.exposure = rs->exposure;
        p.tonemap = rs->tonemap;
        p.samples = float(m.samples);
        p.inv_samples = m.samples > 0 ? 1.0f / float(m.samples) : 0.0f;
        p.max_depth = m.scene_extent;
        p.heat_scale = 1.0f / 160.0f;
        p.denoised = m.denoised ? 1 : 0;
        launch_display(m.stream, p, m.accum.ptr, m.denoised_buf.ptr, m.albedo.ptr, m.nd.ptr, m.svgf.plane[6].ptr, m.heat.ptr,
                       target ? target : m.display.ptr);
        PH_CHECK(hipStreamSynchronize(m.stream));
        return 0;
    });
}

int ph_renderer_download_image(ph_renderer* r, float* rgb) {
    if (!r || !rgb) return fail("null argument");
    return guarded([&] {
        ph_renderer& m = *r;
        const size_t n = size_t(m.width) * size_t(m.height);
        if (n == 0) return 0;
        std::vector<float4> raw(n);
        PH_CHECK(hipMemcpyAsync(raw.data(), m.denoised ? m.denoised_buf.ptr : m.accum.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, m.stream));
        PH_CHECK(hipStreamSynchronize(m.stream));
        const float inv = m.denoised ? 1.0f : (m.samples > 0 ? 1.0f / float(m.samples) : 0.0f);
        for (size_t i = 0; i < n; i++) {
            rgb[3 * i] = raw[i].x * inv;
            rgb[3 * i + 1] = raw[i].y * inv;
            rgb[3 * i + 2] = raw[i].z * inv;
        }
        return 0;
    });
}

int ph_renderer_download_display(ph_renderer* r, uint32_t* out) {
    if (!r || !out) return fail("null argument");
    return guarded([&] {
        const size_t n = size_t(r->width) * size_t(r->height);
        if (n) {
            PH_CHECK(hipMemcpyAsync(out, r->display.ptr, n * 4, hipMemcpyDeviceToHost, r->stream));
            PH_CHECK(hipStreamSynchronize(r->stream));
        }
        return 0;
    });
}

int ph_renderer_width(const ph_renderer* r) { return r ? r->width : fail("null renderer"); }
int ph_renderer_height(const ph_renderer* r) { return r ? r->height : fail("null renderer"); }
int ph_renderer_samples(const ph_renderer* r) { return r ? r->samples : fail("null renderer"); }
float ph_renderer_last_frame_ms(const ph_renderer* r) { return r ? r->last_ms : 0.0f; }
uint32_t* ph_renderer_display_buffer(const ph_renderer* r) { return r ? r->display.ptr : nullptr; }

int ph_renderer_set_engine(ph_renderer* r, int engine) {
    if (!r) return fail("null renderer");
    if (engine != PH_ENGINE_MEGAKERNEL && engine != PH_ENGINE_WAVEFRONT) return fail("unknown engine");
    r->engine = engine;
    return 0;
}

int ph_renderer_get_engine(const ph_renderer* r) { return r ? r->engine : fail("null renderer"); }

int ph_renderer_set_wavefront_config(ph_renderer* r, const ph_wavefront_config* cfg) {
    if (!r || !cfg) return fail("null argument");
    return guarded([&] {
        r->wavefront->set_config(*cfg);
        return 0;
    });
}

int ph_renderer_get_wavefront_config(const ph_renderer* r, ph_wavefront_config* cfg) {
    if (!r || !cfg) return fail("null argument");
    *cfg = r->wavefront->config();
    return 0;
}

int ph_renderer_autotune(ph_renderer* r, const ph_camera_view* cam, const ph_render_settings* rs, int spp, ph_wavefront_config* chosen) {
    if (!r || !cam || !rs) return fail("null argument");
    return guarded([&] {
        ph_renderer& m = *r;
        if (!m.uploaded) return fail("no scene uploaded");
        WavefrontJob job{};
        job.scene = m.view;
        job.camera = *cam;
        job.width = m.width;
        job.height = m.height;
        job.spp = spp > 0 ? spp : 1;
        job.seed = rs->seed;
        job.opt = opts_of(*rs);
        // Tuning renders go to scratch planes so the caller's accumulation is untouched.
        job.accum = m.color.ptr;
        job.albedo = m.position.ptr;
        job.nd = m.emission.ptr;
        job.mat_class = m.mat_class.ptr;
        ph_wavefront_config best = m.wavefront->autotune(m.stream, job);
        if (chosen) *chosen = best;
        return 0;
    });
}

int ph_renderer_set_counting(ph_renderer* r, int enable) {
    if (!r) return fail("null renderer");
    r->counting = enable != 0;
    return 0;
}

int ph_renderer_stats(const ph_renderer* r, ph_render_stats* out) {
    if (!r || !out) return fail("null argument");
    *out = r->stats;
    return 0;
}

const float* ph_renderer_buffer(const ph_renderer* r, int which) {
    if (!r) return nullptr;
    if (which == 7) return r->heat.ptr;
    if (which == 8) return r->svgf.plane[6].ptr ? &r->svgf.plane[6].ptr->x : nullptr;
    const DevBuf<float4>* b[] = {&r->accum, &r->albedo, &r->nd, &r->color, &r->position, &r->emission, &r->denoised_buf};
    if (which < 0 || which > 6) return nullptr;
    return b[which]->ptr ? &b[which]->ptr->x : nullptr;
}

int ph_trace_config(ph_renderer* r, const ph_ray* rays, int n, int anyhit, ph_hit* hits, int block, int depth, int persistent, int lds_nodes,
                    float* ms) {
    if (!r || (n > 0 && (!rays || !hits))) return fail("null argument");
    if (n < 0) return fail("negative ray count");
    if (!r->uploaded) return fail("no scene uploaded");
    if (lds_nodes < 0 || size_t(block) * size_t(depth) * 4 + size_t(lds_nodes) * 64 > 65536) return fail("lds_nodes does not fit the block's LDS");
    return guarded([&] {
        ph_renderer& m = *r;
        int grid = 0;
        if (persistent) {
            int dev = 0, cus = 0;
            PH_CHECK(hipGetDevice(&dev));
            PH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            grid = cus * (2048 / block);
        }
        PH_CHECK(hipEventRecord(m.ev0, m.stream));
        if (!launch_trace(m.stream, m.view, rays, n, anyhit != 0, hits, nullptr, block, depth, persistent != 0, lds_nodes, grid))
            return fail("unsupported traversal launch shape");
        PH_CHECK(hipEventRecord(m.ev1, m.stream));
        PH_CHECK(hipEventSynchronize(m.ev1));
        float t = 0;
        PH_CHECK(hipEventElapsedTime(&t, m.ev0, m.ev1));
        if (ms) *ms = t;
        return 0;
    });
}

int ph_trace(ph_renderer* r, const ph_ray* rays, int n, int anyhit, ph_hit* hits, uint32_t* counts) {
    if (!r || (n > 0 && (!rays || !hits))) return fail("null argument");
    if (n < 0) return fail("negative ray count");
    if (!r->uploaded) return fail("no scene uploaded");
    return guarded([&] {
        ph_renderer& m = *r;
        if (!launch_trace(m.stream, m.view, rays, n, anyhit != 0, hits, counts, 128, 32, false, 0, 0))
            return fail("unsupported traversal launch shape");
        PH_CHECK(hipStreamSynchronize(m.stream));
        return 0;
    });
}

int ph_bvh_build(const float* positions, int num_vertices, const int32_t* indices, int num_tris, int bins, int max_leaf, float traversal_cost,
                 ph_bvh_node* nodes, int node_capacity, int32_t* order, ph_bvh_build_stats* stats) {
    if (num_tris < 0 || num_vertices < 0) return fail("negative size");
    if (num_tris > 0 && (!positions || !indices || !order)) return fail("null argument");
    if (!nodes || !stats) return fail("null argument");
    if (bins < 0 || bins > 64 || bins == 1) return fail("bins must be 0 (32) or in [2, 64]");
    if (max_leaf < 0) return fail("negative leaf size");
    if (node_capacity < std::max(1, 2 * num_tris - 1)) return fail("node array too small: it needs room for max(1, 2 num_tris - 1) nodes");
    if (num_tris > PH_BVH_BUILD_MAX_TRIS)
        return fail("device BVH build is limited to PH_BVH_BUILD_MAX_TRIS (262144) triangles: a 1M-triangle build with an earlier subtree "
                    "kernel made a GPU memory fault whose cause was not found, so larger builds stay switched off; build larger scenes on the host");
    // Every index is read on the device: reject any outside the vertex array here.
    for (int64_t i = 0; i < int64_t(num_tris) * 3; i++)
        if (indices[i] < 0 || indices[i] >= num_vertices) return fail("vertex index out of range");
    return guarded([&] {
        const int64_t count = bvh_build_device(positions, num_vertices, indices, num_tris, bins > 0 ? bins : 32, max_leaf > 0 ? max_leaf : 8,
                                               traversal_cost > 0 ? traversal_cost : 1.0f, nodes, order, *stats);
        return int(count);
    });
}

int ph_svgf_create(ph_svgf** out) {
    if (!out) return fail("null argument");
    return guarded([&] {
        *out = new ph_svgf();
        return 0;
    });
}

void ph_svgf_destroy(ph_svgf* f) { delete f; }

int ph_svgf_resize(ph_svgf* f, int width, int height) {
    if (!f) return fail("null argument");
    if (width <= 0 || height <= 0) return fail("invalid size");
    return guarded([&] {
        svgf_resize(f->state, width, height);
        return 0;
    });
}

int ph_svgf_reset(ph_svgf* f) {
    if (!f) return fail("null argument");
    f->state.has_history = false;
    return 0;
}

int ph_svgf_run(ph_svgf* f, const ph_svgf_input* in, const ph_camera_view* prev, const ph_svgf_settings* s, float* out) {
    if (!f || !in || !prev || !s || !out) return fail("null argument");
    if (!in->color_dev || !in->albedo_dev || !in->normal_depth_dev || !in->position_dev || !in->emission_dev) return fail("missing G-buffer plane");
    if (f->state.w <= 0) return fail("the denoiser is not sized");
    if (s->atrous_iterations < 0 || s->atrous_iterations > 16) return fail("a-trous iterations out of range");
    return guarded([&] {
        PH_CHECK(hipEventRecord(f->ev0, f->stream));
        svgf_run(f->state, f->stream, *in, *prev, *s, reinterpret_cast<float4*>(out));
        PH_CHECK(hipEventRecord(f->ev1, f->stream));
        PH_CHECK(hipEventSynchronize(f->ev1));
        PH_CHECK(hipEventElapsedTime(&f->last_ms, f->ev0, f->ev1));
        return 0;
    });
}

int ph_svgf_run_static(const ph_svgf_input* in, int width, int height, const ph_camera_view* prev, const ph_svgf_settings* s, float* history,
                       int has_history, float* out) {
    if (!in || !prev || !s || !history || !out) return fail("null argument");
    if (!in->color_dev || !in->albedo_dev || !in->normal_depth_dev || !in->position_dev || !in->emission_dev) return fail("missing G-buffer plane");
    if (width <= 0 || height <= 0) return fail("frame size must be positive");
    if (int64_t(width) * int64_t(height) > int64_t(PH_SVGF_STATIC_MAX_PIXELS))
        return fail("frame of " + std::to_string(int64_t(width) * height) + " pixels exceeds the static SVGF scratch of " +
                    std::to_string(PH_SVGF_STATIC_MAX_PIXELS) + " pixels");
    if (s->atrous_iterations < 0 || s->atrous_iterations > 16) return fail("a-trous iterations out of range");
    return guarded([&] {
        svgf_run_static(nullptr, width, height, *in, *prev, *s, reinterpret_cast<float4*>(history), has_history != 0, reinterpret_cast<float4*>(out));
        PH_CHECK(hipStreamSynchronize(nullptr));
        return 0;
    });
}

int ph_svgf_run_compact(ph_svgf* f, const ph_compact_gbuffer* in, const ph_camera_view* cam, const ph_camera_view* prev, const ph_svgf_settings* s,
                        float* out) {
    if (!f || !in || !cam || !prev || !s || !out) return fail("null argument");
    if (!in->color_dev || !in->albedo_dev || !in->emission_dev || !in->depth_dev || !in->normal_dev || !in->material_dev)
        return fail("missing G-buffer plane");
    if (f->state.w <= 0) return fail("the denoiser is not sized");
    if (s->atrous_iterations < 0 || s->atrous_iterations > 16) return fail("a-trous iterations out of range");
    return guarded([&] {
        PH_CHECK(hipEventRecord(f->ev0, f->stream));
        svgf_run_compact(f->state, f->stream, *in, *cam, *prev, *s, reinterpret_cast<float4*>(out));
        PH_CHECK(hipEventRecord(f->ev1, f->stream));
        PH_CHECK(hipEventSynchronize(f->ev1));
        PH_CHECK(hipEventElapsedTime(&f->last_ms, f->ev0, f->ev1));
        return 0;
    });
}

const float* ph_svgf_variance_buffer(const ph_svgf* f) { return f && f->state.plane[6].ptr ? &f->state.plane[6].ptr->x : nullptr; }
float ph_svgf_last_ms(const ph_svgf* f) { return f ? f->last_ms : 0.0f; }

}  // extern "C"
