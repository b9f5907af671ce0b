// The reference's GPU renderer API (src/render/gpu_renderer.h:16-75 and src/denoise/svgf.h:24-59),
// kept as a header-only C++ wrapper over the MI355X device library's C ABI (photon_hip.h).
//
// Callers of the reference (the CLI's render and bench commands, the viewer's backend, the GPU tests)
// keep their code: class and method names, argument meaning, blocking behaviour and error behaviour
// are the reference's. upload/update_lights return false with a message; the other methods throw
// std::runtime_error on a device error; render/render_denoised silently do nothing before an upload or
// at size 0. The scene is passed as the flat descriptor (ph_scene_desc) the reference's Scene::view()
// maps onto field by field, and the camera/settings types are layout-identical to the reference's.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "photon_hip.h"

namespace photon {

enum class DisplayMode : int { Color = 0, Albedo = 1, Normal = 2, Depth = 3, Variance = 4, Heatmap = 5 };

struct SvgfSettings {
    int atrous_iterations = 5;
    float alpha_color = 0.2f;
    float alpha_moments = 0.2f;
    float phi_color = 4.0f;
    float phi_normal = 128.0f;
    float phi_depth = 1.0f;
    ph_svgf_settings c() const { return {atrous_iterations, alpha_color, alpha_moments, phi_color, phi_normal, phi_depth}; }
};

struct GpuInfo {
    std::string name;
    std::string arch;
    size_t memory = 0;
    int multiprocessors = 0;
};

inline bool cuda_available(GpuInfo* info = nullptr) {
    if (ph_device_count() <= 0) return false;
    if (info) {
        ph_gpu_info g{};
        if (ph_device_info(0, &g) != 0) return false;
        info->name = g.name;
        info->arch = g.arch;
        info->memory = size_t(g.memory);
        info->multiprocessors = g.compute_units;
    }
    return true;
}

namespace detail {
inline void check(int rc) {
    if (rc != 0) throw std::runtime_error(ph_last_error());
}
}  // namespace detail

class GpuRenderer {
public:
    GpuRenderer() { detail::check(ph_renderer_create(&r_)); }
    ~GpuRenderer() { ph_renderer_destroy(r_); }
    GpuRenderer(const GpuRenderer&) = delete;
    GpuRenderer& operator=(const GpuRenderer&) = delete;

    bool upload(const ph_scene_desc& scene, std::string& error) {
        if (ph_renderer_upload(r_, &scene) != 0) {
            error = ph_last_error();
            return false;
        }
        return true;
    }
    bool update_lights(const ph_scene_desc& scene, std::string& error) {
        if (ph_renderer_update_lights(r_, &scene) != 0) {
            error = ph_last_error();
            return false;
        }
        return true;
    }
    void resize(int width, int height) { detail::check(ph_renderer_resize(r_, width, height)); }
    void reset() { detail::check(ph_renderer_reset(r_)); }
    void render(const ph_camera_view& camera, const ph_render_settings& settings, int spp) {
        detail::check(ph_renderer_render(r_, &camera, &settings, spp));
    }
    void render_denoised(const ph_camera_view& camera, const ph_camera_view& previous, const ph_render_settings& settings,
                         const SvgfSettings& svgf) {
        ph_svgf_settings s = svgf.c();
        detail::check(ph_renderer_render_denoised(r_, &camera, &previous, &settings, &s));
    }
    void display(const ph_render_settings& settings, DisplayMode mode, uint32_t* target = nullptr) {
        detail::check(ph_renderer_display(r_, &settings, int(mode), target));
    }
    // Average radiance as 3 floats per pixel (the reference returns std::vector<vec3>).
    std::vector<float> download_image() {
        std::vector<float> out(size_t(width()) * size_t(height()) * 3);
        detail::check(ph_renderer_download_image(r_, out.data()));
        return out;
    }
    std::vector<uint32_t> download_display() {
        std::vector<uint32_t> out(size_t(width()) * size_t(height()));
        detail::check(ph_renderer_download_display(r_, out.data()));
        return out;
    }
    int width() const { return ph_renderer_width(r_); }
    int height() const { return ph_renderer_height(r_); }
    int samples() const { return ph_renderer_samples(r_); }
    float last_frame_ms() const { return ph_renderer_last_frame_ms(r_); }
    uint32_t* display_buffer() const { return ph_renderer_display_buffer(r_); }

    // Beyond the reference API: engine choice, tuning and statistics.
    void set_engine(int engine) { detail::check(ph_renderer_set_engine(r_, engine)); }
    ph_wavefront_config autotune(const ph_camera_view& camera, const ph_render_settings& settings, int spp = 1) {
        ph_wavefront_config c{};
        detail::check(ph_renderer_autotune(r_, &camera, &settings, spp, &c));
        return c;
    }
    ph_render_stats stats() const {
        ph_render_stats s{};
        detail::check(ph_renderer_stats(r_, &s));
        return s;
    }
    ph_renderer* handle() const { return r_; }

private:
    ph_renderer* r_ = nullptr;
};

// SvgfInput of the reference: five device planes of 4 floats per pixel.
struct SvgfInput {
    const float* color;
    const float* albedo;
    const float* normal_depth;
    const float* position;
    const float* emission;
};

class Svgf {
public:
    Svgf() { detail::check(ph_svgf_create(&f_)); }
    ~Svgf() { ph_svgf_destroy(f_); }
    Svgf(const Svgf&) = delete;
    Svgf& operator=(const Svgf&) = delete;

    void resize(int width, int height) { detail::check(ph_svgf_resize(f_, width, height)); }
    void reset() { detail::check(ph_svgf_reset(f_)); }
    void run(const SvgfInput& in, const ph_camera_view& previous_camera, const SvgfSettings& settings, float* output) {
        ph_svgf_input i{in.color, in.albedo, in.normal_depth, in.position, in.emission};
        ph_svgf_settings s = settings.c();
        detail::check(ph_svgf_run(f_, &i, &previous_camera, &s, output));
    }
    const float* variance_buffer() const { return ph_svgf_variance_buffer(f_); }

private:
    ph_svgf* f_ = nullptr;
};

}  // namespace photon
