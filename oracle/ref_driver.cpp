// ORACLE — test infrastructure only. Compiled together with the reference's own CPU sources (read in
// place from the reference checkout, never copied) into oracle/_ref/libphoton_ref.so by
// oracle/Makefile. It exposes the reference's Scene::build, CpuRenderer and traversal through a C ABI
// so that tests and the fixture script can compare this repository's producers and oracle with the
// reference bit for bit.
#include <cstring>
#include <string>
#include <vector>

#include "photon_host.h"
#include "photon_types.h"
#include "render/cpu_renderer.h"
#include "render/pixel.h"
#include "render/traverse.h"
#include "scene/builtin.h"
#include "scene/scene.h"

using namespace photon;

static_assert(sizeof(BvhNode) == sizeof(ph_bvh_node), "node layout");
static_assert(sizeof(TriGeom) == sizeof(ph_tri_geom), "triangle layout");
static_assert(sizeof(Material) == sizeof(ph_material), "material layout");
static_assert(sizeof(AliasEntry) == sizeof(ph_alias_entry), "alias layout");
static_assert(sizeof(EmissiveTri) == sizeof(ph_emissive_tri), "emissive layout");
static_assert(sizeof(TextureInfo) == sizeof(ph_texture_info), "texture layout");
static_assert(sizeof(PunctualLight) == sizeof(ph_punctual_light), "light layout");
static_assert(sizeof(CameraView) == sizeof(ph_camera_view), "camera layout");

struct ref_scene {
    Scene scene;
};

extern "C" {

typedef phs_raw_scene ref_raw_scene;

int ref_scene_builtin(const char* name, ref_scene** out) {
    auto* r = new ref_scene();
    if (!make_builtin_scene(name, r->scene)) {
        delete r;
        return -1;
    }
    *out = r;
    return 0;
}

int ref_scene_from_raw(const ref_raw_scene* raw, ref_scene** out) {
    auto* r = new ref_scene();
    Scene& s = r->scene;
    const int nv = raw->num_vertices;
    s.positions.resize(size_t(nv));
    std::memcpy(s.positions.data(), raw->positions, size_t(nv) * 12);
    if (raw->normals) {
        s.normals.resize(size_t(nv));
        std::memcpy(s.normals.data(), raw->normals, size_t(nv) * 12);
    }
    if (raw->uvs) {
        s.uvs.resize(size_t(nv));
        std::memcpy(s.uvs.data(), raw->uvs, size_t(nv) * 8);
    }
    s.indices.assign(raw->indices, raw->indices + 3 * size_t(raw->num_tris));
    s.tri_material.assign(raw->tri_material, raw->tri_material + raw->num_tris);
    s.materials.resize(size_t(raw->num_materials));
    std::memcpy(s.materials.data(), raw->materials, size_t(raw->num_materials) * sizeof(Material));
    for (int t = 0; t < raw->num_textures; t++) {
        Texture tex;
        tex.width = raw->tex_width[t];
        tex.height = raw->tex_height[t];
        tex.texels.assign(raw->tex_texels[t], raw->tex_texels[t] + size_t(tex.width) * size_t(tex.height));
        tex.repeat = raw->tex_repeat[t] != 0;
        tex.srgb = raw->tex_srgb[t] != 0;
        s.textures.push_back(tex);
    }
    s.punctual.resize(size_t(raw->num_punctual));
    if (raw->num_punctual) std::memcpy(s.punctual.data(), raw->punctual, size_t(raw->num_punctual) * sizeof(PunctualLight));
    if (raw->env_texels) {
        s.env.width = raw->env_width;
        s.env.height = raw->env_height;
        s.env.texels.resize(size_t(raw->env_width) * size_t(raw->env_height));
        std::memcpy(s.env.texels.data(), raw->env_texels, s.env.texels.size() * 12);
    }
    s.env.color = vec3(raw->env_color[0], raw->env_color[1], raw->env_color[2]);
    s.env.intensity = raw->env_intensity;
    s.env.rotation = raw->env_rotation;
    s.camera.position = vec3(raw->cam_position[0], raw->cam_position[1], raw->cam_position[2]);
    s.camera.target = vec3(raw->cam_target[0], raw->cam_target[1], raw->cam_target[2]);
    s.camera.up = vec3(raw->cam_up[0], raw->cam_up[1], raw->cam_up[2]);
    s.camera.vfov = raw->cam_vfov;
    s.camera.aperture = raw->cam_aperture;
    s.camera.focus_distance = raw->cam_focus;
    *out = r;
    return 0;
}

void ref_scene_destroy(ref_scene* r) { delete r; }

int ref_scene_clear_normals(ref_scene* r) {
    r->scene.normals.clear();
    return 0;
}

int ref_scene_build(ref_scene* r, int threads) {
    BvhBuildOptions o;
    o.threads = threads;
    r->scene.build(o);
    return 0;
}

int ref_scene_desc(const ref_scene* r, ph_scene_desc* d) {
    SceneView v = r->scene.view();
    std::memset(d, 0, sizeof(*d));
    d->nodes = reinterpret_cast<const ph_bvh_node*>(v.nodes);
    d->num_nodes = v.num_nodes;
    d->tris = reinterpret_cast<const ph_tri_geom*>(v.tris);
    d->num_tris = v.num_tris;
    d->indices = v.indices;
    d->normals = v.normals ? &v.normals[0].x : nullptr;
    d->uvs = v.uvs ? &v.uvs[0].x : nullptr;
    d->tangents = v.tangents ? &v.tangents[0].x : nullptr;
    d->num_vertices = int(r->scene.positions.size());
    d->materials = reinterpret_cast<const ph_material*>(v.materials);
    d->num_materials = int(r->scene.materials.size());
    d->textures = reinterpret_cast<const ph_texture_info*>(v.textures);
    d->num_textures = int(r->scene.texture_infos.size());
    d->texels = v.texels;
    d->num_texels = int64_t(r->scene.texel_pool.size());
    d->srgb_lut = v.srgb_lut;
    d->emissive = reinterpret_cast<const ph_emissive_tri*>(v.emissive);
    d->tri_emissive = v.tri_emissive;
    d->emissive_alias = reinterpret_cast<const ph_alias_entry*>(v.emissive_alias);
    d->num_emissive = v.num_emissive;
    d->emissive_power = v.emissive_power;
    d->punctual = reinterpret_cast<const ph_punctual_light*>(v.punctual);
    d->num_punctual = v.num_punctual;
    d->env_texels = v.env.texels ? &v.env.texels[0].x : nullptr;
    d->env_marginal = v.env.marginal_cdf;
    d->env_conditional = v.env.conditional_cdf;
    d->env_width = v.env.width;
    d->env_height = v.env.height;
    d->env_intensity = v.env.intensity;
    d->env_rotation = v.env.rotation;
    d->env_color[0] = v.env.color.x, d->env_color[1] = v.env.color.y, d->env_color[2] = v.env.color.z;
    d->env_marginal_integral = v.env.marginal_integral;
    d->p_env = v.p_env;
    d->p_emissive = v.p_emissive;
    d->p_punctual = v.p_punctual;
    d->bounds_min[0] = v.bounds_min.x, d->bounds_min[1] = v.bounds_min.y, d->bounds_min[2] = v.bounds_min.z;
    d->bounds_max[0] = v.bounds_max.x, d->bounds_max[1] = v.bounds_max.y, d->bounds_max[2] = v.bounds_max.z;
    return 0;
}

const int32_t* ref_scene_order(const ref_scene* r) { return r->scene.bvh.order.data(); }

int ref_scene_stats(const ref_scene* r, phs_bvh_stats* out) {
    const BvhStats& st = r->scene.bvh.stats;
    out->nodes = st.nodes;
    out->leaves = st.leaves;
    out->max_depth = st.max_depth;
    out->sah_cost = st.sah_cost;
    out->build_ms = st.build_ms;
    return 0;
}

int ref_camera_view(const ref_scene* r, float aspect, ph_camera_view* out) {
    CameraView v = r->scene.camera.view(aspect);
    std::memcpy(out, &v, sizeof(v));
    return 0;
}

static CameraView cam_of(const ph_camera_view* c) {
    CameraView v;
    std::memcpy(&v, c, sizeof(v));
    return v;
}

static RenderSettings settings_of(const ph_render_settings* rs) {
    RenderSettings s;
    s.width = rs->width;
    s.height = rs->height;
    s.path.max_depth = rs->path.max_depth;
    s.path.rr_depth = rs->path.rr_depth;
    s.path.clamp = rs->path.clamp;
    s.path.nee = rs->path.nee;
    s.exposure = rs->exposure;
    s.tonemap = ToneMap(rs->tonemap);
    s.seed = rs->seed;
    return s;
}

// CpuRenderer: `spp` samples from a reset accumulation; writes the average radiance (3 floats/pixel).
int ref_render(const ref_scene* r, const ph_camera_view* cam, const ph_render_settings* rs, int spp, int threads, float* out_rgb) {
    CpuRenderer cpu(r->scene);
    cpu.resize(rs->width, rs->height);
    cpu.render(cam_of(cam), settings_of(rs), spp, threads);
    std::vector<vec3> img = cpu.image();
    std::memcpy(out_rgb, img.data(), img.size() * 12);
    return 0;
}

int ref_render_gbuffer(const ref_scene* r, const ph_camera_view* cam, const ph_render_settings* rs, uint32_t frame, float* color4, float* albedo4,
                       float* nd4, float* pos4, float* emis4) {
    SceneView v = r->scene.view();
    CameraView c = cam_of(cam);
    RenderSettings s = settings_of(rs);
    for (int y = 0; y < rs->height; y++)
        for (int x = 0; x < rs->width; x++) {
            PrimaryInfo p;
            vec3 L = render_sample(v, c, x, y, rs->width, rs->height, frame, s.seed, s.path, &p);
            size_t i = size_t(y) * size_t(rs->width) + size_t(x);
            vec4 outs[5] = {vec4(L, 1), vec4(p.albedo, 1), vec4(p.normal, p.depth), vec4(p.position, float(p.material)), vec4(p.emission, 1)};
            float* dst[5] = {color4, albedo4, nd4, pos4, emis4};
            for (int k = 0; k < 5; k++) std::memcpy(dst[k] + 4 * i, &outs[k], 16);
        }
    return 0;
}

int ref_trace(const ref_scene* r, const ph_ray* rays, int n, int anyhit, ph_hit* hits) {
    SceneView v = r->scene.view();
    for (int i = 0; i < n; i++) {
        Ray ray;
        ray.o = vec3(rays[i].o[0], rays[i].o[1], rays[i].o[2]);
        ray.d = vec3(rays[i].d[0], rays[i].d[1], rays[i].d[2]);
        ray.tmax = rays[i].tmax;
        Hit h;
        if (anyhit)
            traverse<true>(v, ray, h, nullptr);
        else
            traverse<false>(v, ray, h, nullptr);
        hits[i].t = h.t;
        hits[i].tri = h.tri;
        hits[i].u = h.u;
        hits[i].v = h.v;
    }
    return 0;
}

int ref_alias_table(const float* w, int n, ph_alias_entry* out) {
    std::vector<float> weights(w, w + n);
    std::vector<AliasEntry> t = build_alias_table(weights);
    std::memcpy(out, t.data(), t.size() * sizeof(AliasEntry));
    return 0;
}

}  // extern "C"
