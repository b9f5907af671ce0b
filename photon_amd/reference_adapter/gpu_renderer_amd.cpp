// Reference-side adapter: the implementation of photon::GpuRenderer (src/render/gpu_renderer.h:36-75) over
// the MI355X device library's C ABI (include/photon_hip.h). In the reference build it replaces
// src/render/gpu_renderer.cu; the header and every caller (render_cli.cpp, viewer.cpp, tests) stay as they
// are. It needs no HIP headers: photon_hip.h and photon_types.h are plain C.
//
// Error behaviour is the reference's: upload and update_lights return false with the message, every other
// method throws std::runtime_error. The inline accessors of the header (width, height, samples,
// last_frame_ms) read the members kept here in step with the device library.
#include <cstring>
#include <stdexcept>

#include "photon_hip.h"
#include "render/gpu_renderer.h"
#include "scene/scene.h"

namespace photon {

static_assert(sizeof(BvhNode) == sizeof(ph_bvh_node) && sizeof(TriGeom) == sizeof(ph_tri_geom), "node and triangle layouts");
static_assert(sizeof(Material) == sizeof(ph_material) && sizeof(CameraView) == sizeof(ph_camera_view), "material and camera layouts");
static_assert(sizeof(AliasEntry) == sizeof(ph_alias_entry) && sizeof(EmissiveTri) == sizeof(ph_emissive_tri), "light table layouts");
static_assert(sizeof(PunctualLight) == sizeof(ph_punctual_light) && sizeof(TextureInfo) == sizeof(ph_texture_info), "light and texture layouts");

namespace {

void check(int rc) {
    if (rc != 0) throw std::runtime_error(ph_last_error());
}

// The flat descriptor of a built scene: SceneView (src/render/scene_view.h:117-150) plus array lengths.
ph_scene_desc describe(const Scene& s) {
    const SceneView v = s.view();
    ph_scene_desc d{};
    d.nodes = reinterpret_cast<const ph_bvh_node*>(v.nodes);
    d.num_nodes = v.num_nodes;
    d.tris = reinterpret_cast<const ph_tri_geom*>(v.tris);
    d.num_tris = v.num_tris;
    d.indices = v.indices;
    d.normals = reinterpret_cast<const float*>(v.normals);
    d.uvs = reinterpret_cast<const float*>(v.uvs);
    d.tangents = reinterpret_cast<const float*>(v.tangents);
    d.num_vertices = int(s.positions.size());
    d.materials = reinterpret_cast<const ph_material*>(v.materials);
    d.num_materials = int(s.materials.size());
    d.textures = reinterpret_cast<const ph_texture_info*>(v.textures);
    d.num_textures = int(s.texture_infos.size());
    d.texels = v.texels;
    d.num_texels = int64_t(s.texel_pool.size());
    d.srgb_lut = v.srgb_lut;
    d.emissive = reinterpret_cast<const ph_emissive_tri*>(v.emissive);
    d.tri_emissive = v.tri_emissive;
    d.emissive_alias = reinterpret_cast<const ph_alias_entry*>(v.emissive_alias);
    d.num_emissive = v.num_emissive;
    d.emissive_power = v.emissive_power;
    d.punctual = reinterpret_cast<const ph_punctual_light*>(v.punctual);
    d.num_punctual = v.num_punctual;
    d.env_texels = reinterpret_cast<const float*>(v.env.texels);
    d.env_marginal = v.env.marginal_cdf;
    d.env_conditional = v.env.conditional_cdf;
    d.env_width = v.env.width;
    d.env_height = v.env.height;
    d.env_intensity = v.env.intensity;
    d.env_rotation = v.env.rotation;
    std::memcpy(d.env_color, &v.env.color, sizeof(d.env_color));
    d.env_marginal_integral = v.env.marginal_integral;
    d.p_env = v.p_env;
    d.p_emissive = v.p_emissive;
    d.p_punctual = v.p_punctual;
    std::memcpy(d.bounds_min, &v.bounds_min, sizeof(d.bounds_min));
    std::memcpy(d.bounds_max, &v.bounds_max, sizeof(d.bounds_max));
    return d;
}

// PathOptions without its host-only ray counter (integrator.h:21-27).
ph_render_settings settings_of(const RenderSettings& s) {
    ph_render_settings c{};
    c.width = s.width;
    c.height = s.height;
    c.path.max_depth = s.path.max_depth;
    c.path.rr_depth = s.path.rr_depth;
    c.path.clamp = s.path.clamp;
    c.path.nee = s.path.nee ? 1 : 0;
    c.exposure = s.exposure;
    c.tonemap = int(s.tonemap);
    c.seed = s.seed;
    return c;
}

const ph_camera_view* view_of(const CameraView& c) { return reinterpret_cast<const ph_camera_view*>(&c); }

}  // namespace

bool cuda_available(GpuInfo* info) {
    if (ph_device_count() <= 0) return false;
    if (info) {
        ph_gpu_info g{};
        if (ph_device_info(0, &g) != 0) return false;
        info->name = g.name;
        info->memory = size_t(g.memory);
        info->multiprocessors = g.compute_units;
    }
    return true;
}

struct GpuRenderer::Impl {
    ph_renderer* r = nullptr;
};

GpuRenderer::GpuRenderer() : impl_(std::make_unique<Impl>()) { check(ph_renderer_create(&impl_->r)); }

GpuRenderer::~GpuRenderer() { ph_renderer_destroy(impl_->r); }

bool GpuRenderer::upload(const Scene& scene, std::string& error) {
    const ph_scene_desc d = describe(scene);
    if (ph_renderer_upload(impl_->r, &d) == 0) {
        samples_ = 0;
        return true;
    }
    error = ph_last_error();
    return false;
}

bool GpuRenderer::update_lights(const Scene& scene, std::string& error) {
    const ph_scene_desc d = describe(scene);
    if (ph_renderer_update_lights(impl_->r, &d) == 0) {
        samples_ = 0;
        return true;
    }
    error = ph_last_error();
    return false;
}

void GpuRenderer::resize(int width, int height) {
    check(ph_renderer_resize(impl_->r, width, height));
    width_ = width;
    height_ = height;
    samples_ = ph_renderer_samples(impl_->r);
}

void GpuRenderer::reset() {
    check(ph_renderer_reset(impl_->r));
    samples_ = 0;
    frame_ = 0;
    denoised_ = false;
}

void GpuRenderer::render(const CameraView& camera, const RenderSettings& settings, int spp) {
    const ph_render_settings c = settings_of(settings);
    check(ph_renderer_render(impl_->r, view_of(camera), &c, spp));
    samples_ = ph_renderer_samples(impl_->r);
    last_ms_ = ph_renderer_last_frame_ms(impl_->r);
    denoised_ = false;
}

void GpuRenderer::render_denoised(const CameraView& camera, const CameraView& previous, const RenderSettings& settings, const SvgfSettings& svgf) {
    const ph_render_settings c = settings_of(settings);
    const ph_svgf_settings g{svgf.atrous_iterations, svgf.alpha_color, svgf.alpha_moments, svgf.phi_color, svgf.phi_normal, svgf.phi_depth};
    check(ph_renderer_render_denoised(impl_->r, view_of(camera), view_of(previous), &c, &g));
    samples_ = ph_renderer_samples(impl_->r);
    last_ms_ = ph_renderer_last_frame_ms(impl_->r);
    denoised_ = true;
}

void GpuRenderer::display(const RenderSettings& settings, DisplayMode mode, uint32_t* target) {
    const ph_render_settings c = settings_of(settings);
    check(ph_renderer_display(impl_->r, &c, int(mode), target));
}

std::vector<vec3> GpuRenderer::download_image() {
    std::vector<vec3> out(size_t(width_) * size_t(height_));
    check(ph_renderer_download_image(impl_->r, reinterpret_cast<float*>(out.data())));
    return out;
}

std::vector<uint32_t> GpuRenderer::download_display() {
    std::vector<uint32_t> out(size_t(width_) * size_t(height_));
    check(ph_renderer_download_display(impl_->r, out.data()));
    return out;
}

uint32_t* GpuRenderer::display_buffer() const { return ph_renderer_display_buffer(impl_->r); }

}  // namespace photon
