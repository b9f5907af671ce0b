// Reference-side adapter: photon::Svgf (src/denoise/svgf.h:32-59) over the device library's ph_svgf
// (include/photon_hip.h). In the reference build it replaces src/denoise/svgf.cu; svgf.h stays as it is.
// The header has no slot for an implementation handle, so the handles live in a table keyed by the
// object; variance_buffer() (inline in the header, returning filter_[0]) reads the device library's
// variance plane, which run() stores in filter_[0].
#include <mutex>
#include <stdexcept>
#include <unordered_map>

#include "denoise/svgf.h"
#include "photon_hip.h"
#include "render/gpu_renderer.h"

namespace photon {

namespace {

std::mutex g_lock;
std::unordered_map<const Svgf*, ph_svgf*> g_handles;

ph_svgf* handle_of(const Svgf* f) {
    std::lock_guard<std::mutex> hold(g_lock);
    ph_svgf*& h = g_handles[f];
    if (!h && ph_svgf_create(&h) != 0) throw std::runtime_error(ph_last_error());
    return h;
}

void check(int rc) {
    if (rc != 0) throw std::runtime_error(ph_last_error());
}

}  // namespace

Svgf::~Svgf() { free_buffers(); }

void Svgf::free_buffers() {
    std::lock_guard<std::mutex> hold(g_lock);
    const auto it = g_handles.find(this);
    if (it == g_handles.end()) return;
    ph_svgf_destroy(it->second);
    g_handles.erase(it);
    filter_[0] = nullptr;
}

void Svgf::resize(int width, int height) {
    check(ph_svgf_resize(handle_of(this), width, height));
    width_ = width;
    height_ = height;
    has_history_ = false;
}

void Svgf::reset() {
    check(ph_svgf_reset(handle_of(this)));
    has_history_ = false;
}

void Svgf::run(const SvgfInput& in, const CameraView& previous_camera, const SvgfSettings& s, vec4* output) {
    ph_svgf* h = handle_of(this);
    const ph_svgf_input planes{reinterpret_cast<const float*>(in.color), reinterpret_cast<const float*>(in.albedo),
                               reinterpret_cast<const float*>(in.normal_depth), reinterpret_cast<const float*>(in.position),
                               reinterpret_cast<const float*>(in.emission)};
    const ph_svgf_settings g{s.atrous_iterations, s.alpha_color, s.alpha_moments, s.phi_color, s.phi_normal, s.phi_depth};
    check(ph_svgf_run(h, &planes, reinterpret_cast<const ph_camera_view*>(&previous_camera), &g, reinterpret_cast<float*>(output)));
    filter_[0] = reinterpret_cast<vec4*>(const_cast<float*>(ph_svgf_variance_buffer(h)));
    has_history_ = true;
}

}  // namespace photon
