// C ABI of the device library (include/photon_hip.h): the renderer object that owns the device copy of
// a scene and the frame buffers, with the reference GpuRenderer's semantics
// (src/render/gpu_renderer.cu:243-452): upload/update_lights copy the flat scene arrays, render adds
// samples to the accumulation, render_denoised runs one sample plus SVGF, display tone-maps into RGBA8,
// and each call blocks until the device is done. All work is issued on one HIP stream per renderer.
#include <chrono>
#include <cstring>
#include <memory>

#include "internal.h"
#include "svgf_state.h"
#include "wavefront.h"

using namespace phd;
using namespace phk;

namespace {

thread_local std::string g_error;

int fail(const std::string& msg) {
    g_error = msg;
    return -1;
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail("out of host memory");
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

}  // namespace

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
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // Scene arrays in HBM.
    DevBuf<float4> nodes, tris, materials;
    DevBuf<int> indices, tri_emissive;
    DevBuf<float> normals, uvs, tangents, srgb_lut, env_texels, env_marginal, env_conditional;
    DevBuf<ph_texture_info> textures;
    DevBuf<uint32_t> texels;
    DevBuf<ph_emissive_tri> emissive;
    DevBuf<ph_alias_entry> alias;
    DevBuf<ph_punctual_light> punctual;
    DScene view{};
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
    DevBuf<int> work_counter;
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
        work_counter.allocate(1);
        wavefront = std::make_unique<WavefrontEngine>();
    }
    ~ph_renderer() {
        wavefront.reset();
        if (stream) (void)hipStreamSynchronize(stream);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
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

void validate(const ph_scene_desc* d) {
    if (!d) throw std::invalid_argument("null scene descriptor");
    if (!d->nodes || d->num_nodes <= 0) throw std::invalid_argument("the scene is not built (no BVH nodes)");
    if (d->num_tris < 0 || (d->num_tris > 0 && (!d->tris || !d->indices))) throw std::invalid_argument("missing triangle arrays");
    if (!d->materials || d->num_materials <= 0) throw std::invalid_argument("the scene has no materials");
    if (!d->srgb_lut) throw std::invalid_argument("missing sRGB table");
    if (d->num_emissive > 0 && (!d->emissive || !d->emissive_alias || !d->tri_emissive)) throw std::invalid_argument("missing light tables");
    if (d->env_texels && (d->env_width <= 0 || d->env_height <= 0 || !d->env_marginal || !d->env_conditional))
        throw std::invalid_argument("incomplete environment image");
    if (d->num_textures > 0 && (!d->textures || !d->texels)) throw std::invalid_argument("missing texture pool");
}

void upload_lights(ph_renderer& m, const ph_scene_desc& d) {
    hipStream_t st = m.stream;
    m.tris.upload(d.tris, size_t(d.num_tris) * 3, st);
    m.materials.upload(d.materials, size_t(d.num_materials) * 5, st);
    m.emissive.upload(d.emissive, size_t(d.num_emissive), st);
    m.tri_emissive.upload(d.tri_emissive, d.tri_emissive ? size_t(d.num_tris) : 0, st);
    m.alias.upload(d.emissive_alias, size_t(d.num_emissive), st);
    const size_t env_n = d.env_texels ? size_t(d.env_width) * size_t(d.env_height) : 0;
    m.env_texels.upload(d.env_texels, env_n * 3, st);
    m.env_marginal.upload(d.env_marginal, env_n ? size_t(d.env_height) + 1 : 0, st);
    m.env_conditional.upload(d.env_conditional, env_n ? size_t(d.env_height) * size_t(d.env_width + 1) : 0, st);
    DScene& v = m.view;
    v.tris = or_null(m.tris);
    v.materials = or_null(m.materials);
    v.emissive = or_null(m.emissive);
    v.tri_emissive = or_null(m.tri_emissive);
    v.alias = or_null(m.alias);
    v.num_emissive = d.num_emissive;
    v.emissive_power = d.emissive_power;
    v.env_texels = or_null(m.env_texels);
    v.env_marginal = or_null(m.env_marginal);
    v.env_conditional = or_null(m.env_conditional);
    v.env_w = d.env_width;
    v.env_h = d.env_height;
    v.env_intensity = d.env_intensity;
    v.env_rotation = d.env_rotation;
    v.env_integral = d.env_marginal_integral;
    std::memcpy(v.env_color, d.env_color, 12);
    v.p_env = d.p_env;
    v.p_emissive = d.p_emissive;
    v.p_punctual = d.p_punctual;
    v.num_materials = d.num_materials;
}

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
        validate(d);
        ph_renderer& m = *r;
        hipStream_t st = m.stream;
        m.uploaded = false;
        m.nodes.upload(d->nodes, size_t(d->num_nodes) * 4, st);
        m.indices.upload(d->indices, size_t(d->num_tris) * 3, st);
        m.normals.upload(d->normals, d->normals ? size_t(d->num_vertices) * 3 : 0, st);
        m.uvs.upload(d->uvs, d->uvs ? size_t(d->num_vertices) * 2 : 0, st);
        m.tangents.upload(d->tangents, d->tangents ? size_t(d->num_vertices) * 4 : 0, st);
        m.textures.upload(d->textures, size_t(d->num_textures), st);
        m.texels.upload(d->texels, size_t(d->num_texels), st);
        m.srgb_lut.upload(d->srgb_lut, 256, st);
        m.punctual.upload(d->punctual, size_t(d->num_punctual), st);
        upload_lights(m, *d);
        DScene& v = m.view;
        v.nodes = or_null(m.nodes);
        v.indices = or_null(m.indices);
        v.normals = or_null(m.normals);
        v.uvs = or_null(m.uvs);
        v.tangents = or_null(m.tangents);
        v.textures = or_null(m.textures);
        v.texels = or_null(m.texels);
        v.srgb_lut = or_null(m.srgb_lut);
        v.punctual = or_null(m.punctual);
        v.num_tris = d->num_tris;
        v.num_nodes = d->num_nodes;
        v.num_punctual = d->num_punctual;
        v.num_textures = d->num_textures;
        float dx = d->bounds_max[0] - d->bounds_min[0], dy = d->bounds_max[1] - d->bounds_min[1], dz = d->bounds_max[2] - d->bounds_min[2];
        m.scene_extent = fmaxf(sqrtf(dx * dx + dy * dy + dz * dz), 1e-3f);
        PH_CHECK(hipStreamSynchronize(st));
        m.uploaded = true;
        m.reset();
        return 0;
    });
}

int ph_renderer_update_lights(ph_renderer* r, const ph_scene_desc* d) {
    if (!r) return fail("null renderer");
    if (!r->uploaded) return ph_renderer_upload(r, d);
    return guarded([&] {
        validate(d);
        upload_lights(*r, *d);
        PH_CHECK(hipStreamSynchronize(r->stream));
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
        ph_svgf_input in{&m.color.ptr->x, &m.albedo.ptr->x, &m.nd.ptr->x, &m.position.ptr->x, &m.emission.ptr->x};
        svgf_run(m.svgf, m.stream, in, *prev, *sv, m.denoised_buf.ptr);
        PH_CHECK(hipEventRecord(m.ev1, m.stream));
        PH_CHECK(hipEventSynchronize(m.ev1));
        PH_CHECK(hipEventElapsedTime(&m.last_ms, m.ev0, m.ev1));
        m.frame++;
        m.samples = 1;
        m.last_camera = *cam;
        m.heat_valid = false;
        return 0;
    });
}

int ph_renderer_display(ph_renderer* r, const ph_render_settings* rs, int mode, uint32_t* target) {
    if (!r || !rs) return fail("null argument");
    if (mode < PH_DISPLAY_COLOR || mode > PH_DISPLAY_HEATMAP) return fail("unknown display mode");
    return guarded([&] {
        ph_renderer& m = *r;
        if (m.width == 0 || m.height == 0) return 0;
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
    const DevBuf<float4>* b[] = {&r->accum, &r->albedo, &r->nd, &r->color, &r->position, &r->emission, &r->denoised_buf};
    if (which < 0 || which > 6) return nullptr;
    return b[which]->ptr ? &b[which]->ptr->x : nullptr;
}

int ph_trace_config(ph_renderer* r, const ph_ray* rays, int n, int anyhit, ph_hit* hits, int block, int depth, int persistent, float* ms) {
    if (!r || (n > 0 && (!rays || !hits))) return fail("null argument");
    if (n < 0) return fail("negative ray count");
    if (!r->uploaded) return fail("no scene uploaded");
    return guarded([&] {
        ph_renderer& m = *r;
        int grid = 0;
        if (persistent) {
            int dev = 0, cus = 0;
            PH_CHECK(hipGetDevice(&dev));
            PH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            grid = cus * (2048 / block);
            PH_CHECK(hipMemsetAsync(m.work_counter.ptr, 0, sizeof(int), m.stream));
        }
        PH_CHECK(hipEventRecord(m.ev0, m.stream));
        if (!launch_trace(m.stream, m.view, rays, n, anyhit != 0, hits, nullptr, block, depth, persistent != 0, m.work_counter.ptr, grid))
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
        if (!launch_trace(m.stream, m.view, rays, n, anyhit != 0, hits, counts, 128, 32, false, m.work_counter.ptr, 0))
            return fail("unsupported traversal launch shape");
        PH_CHECK(hipStreamSynchronize(m.stream));
        return 0;
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

const float* ph_svgf_variance_buffer(const ph_svgf* f) { return f && f->state.plane[6].ptr ? &f->state.plane[6].ptr->x : nullptr; }
float ph_svgf_last_ms(const ph_svgf* f) { return f ? f->last_ms : 0.0f; }

}  // extern "C"
