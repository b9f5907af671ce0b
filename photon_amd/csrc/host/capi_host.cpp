// C ABI over the host scene producers (declared in include/photon_host.h).
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "gltf.hpp"
#include "image_io.hpp"
#include "json.hpp"
#include "photon_host.h"
#include "scene.hpp"

using namespace phh;

struct phs_scene {
    Scene scene;
    // Texture tables handed out by phs_scene_raw.
    std::vector<int32_t> tex_w, tex_h, tex_repeat, tex_srgb;
    std::vector<const uint32_t*> tex_ptr;
};

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

vec3 v3(const float* p) { return vec3(p[0], p[1], p[2]); }

}  // namespace

extern "C" {

const char* phs_last_error(void) { return g_error.c_str(); }

int phs_scene_create(phs_scene** out) {
    if (!out) return fail("null output pointer");
    return guarded([&] {
        *out = new phs_scene();
        return 0;
    });
}

void phs_scene_destroy(phs_scene* s) { delete s; }

int phs_scene_builtin(phs_scene* s, const char* name) {
    if (!s || !name) return fail("null argument");
    return guarded([&] { return make_builtin_scene(name, s->scene) ? 0 : fail(std::string("unknown built-in scene ") + name); });
}

int phs_scene_spheres(phs_scene* s, int count, float box, float alpha_fraction, uint64_t seed) {
    if (!s || count < 0) return fail("invalid argument");
    return guarded([&] {
        make_sphere_soup(s->scene, count, box, alpha_fraction, seed);
        return 0;
    });
}

int phs_scene_field(phs_scene* s, int count, uint64_t seed, int env_width) {
    if (!s || count < 0 || env_width < 2) return fail("invalid argument");
    return guarded([&] {
        make_sphere_field(s->scene, count, seed, env_width);
        return 0;
    });
}

int phs_scene_add_mesh(phs_scene* s, const float* positions, int num_vertices, const int32_t* indices, int num_indices, int material,
                       const float* normals, const float* uvs) {
    if (!s || (num_vertices > 0 && !positions) || (num_indices > 0 && !indices) || num_indices % 3) return fail("invalid mesh");
    for (int i = 0; i < num_indices; i++)
        if (indices[i] < 0 || indices[i] >= num_vertices) return fail("mesh index out of range");
    return guarded([&] {
        std::vector<vec3> pos(static_cast<size_t>(num_vertices)), nrm;
        std::vector<vec2> uv;
        for (int i = 0; i < num_vertices; i++) pos[size_t(i)] = v3(positions + 3 * i);
        if (normals) {
            nrm.resize(size_t(num_vertices));
            for (int i = 0; i < num_vertices; i++) nrm[size_t(i)] = v3(normals + 3 * i);
        }
        if (uvs) {
            uv.resize(size_t(num_vertices));
            for (int i = 0; i < num_vertices; i++) uv[size_t(i)] = vec2(uvs[2 * i], uvs[2 * i + 1]);
        }
        std::vector<int> idx(indices, indices + num_indices);
        append_mesh(s->scene.mesh, pos, idx, material, nrm, uv);
        return 0;
    });
}

int phs_scene_add_material(phs_scene* s, const ph_material* m) {
    if (!s || !m) return fail("null argument");
    return guarded([&] { return s->scene.add_material(*m); });
}

int phs_scene_set_material(phs_scene* s, int index, const ph_material* m) {
    if (!s || !m || index < 0 || index >= int(s->scene.materials.size())) return fail("material index out of range");
    s->scene.materials[size_t(index)] = *m;
    return 0;
}

int phs_scene_get_material(const phs_scene* s, int index, ph_material* out) {
    if (!s || !out || index < 0 || index >= int(s->scene.materials.size())) return fail("material index out of range");
    *out = s->scene.materials[size_t(index)];
    return 0;
}

int phs_scene_add_texture(phs_scene* s, int width, int height, const uint32_t* texels, int repeat, int srgb) {
    if (!s || width <= 0 || height <= 0 || !texels) return fail("invalid texture");
    return guarded([&] {
        Image8 t;
        t.width = width;
        t.height = height;
        t.px.assign(texels, texels + size_t(width) * size_t(height));
        t.tile = repeat != 0;
        t.srgb = srgb != 0;
        return s->scene.add_image(std::move(t));
    });
}

int phs_scene_add_punctual(phs_scene* s, const ph_punctual_light* l) {
    if (!s || !l) return fail("null argument");
    return guarded([&] {
        s->scene.lamps.push_back(*l);
        return int(s->scene.lamps.size()) - 1;
    });
}

int phs_scene_set_env_color(phs_scene* s, const float rgb[3], float intensity) {
    if (!s || !rgb) return fail("null argument");
    s->scene.backdrop = Backdrop();
    s->scene.backdrop.flat = v3(rgb);
    s->scene.backdrop.gain = intensity;
    return 0;
}

int phs_scene_set_env_image(phs_scene* s, int width, int height, const float* rgb, float intensity, float rotation) {
    if (!s || width <= 0 || height <= 0 || !rgb) return fail("invalid environment image");
    return guarded([&] {
        Backdrop& e = s->scene.backdrop;
        e.width = width;
        e.height = height;
        e.rgb.resize(size_t(width) * size_t(height));
        for (size_t i = 0; i < e.rgb.size(); i++) e.rgb[i] = v3(rgb + 3 * i);
        e.gain = intensity;
        e.yaw = rotation;
        return 0;
    });
}

int phs_scene_set_sky(phs_scene* s, const float sun[3], float sun_intensity, int width) {
    if (!s || !sun || width < 2) return fail("invalid argument");
    return guarded([&] {
        make_sky(s->scene.backdrop, v3(sun), sun_intensity, width);
        return 0;
    });
}

int phs_scene_clear_normals(phs_scene* s) {
    if (!s) return fail("null argument");
    s->scene.mesh.nrm.clear();
    return 0;
}

int phs_scene_set_camera(phs_scene* s, const float position[3], const float target[3], const float up[3], float vfov, float aperture,
                         float focus_distance) {
    if (!s || !position || !target || !up) return fail("null argument");
    Viewpoint& c = s->scene.view;
    c.eye = v3(position);
    c.aim = v3(target);
    c.sky = v3(up);
    c.fov_deg = vfov;
    c.lens = aperture;
    c.focus = focus_distance;
    return 0;
}

int phs_scene_get_camera(const phs_scene* s, float position[3], float target[3], float up[3], float* vfov) {
    if (!s || !position || !target || !up || !vfov) return fail("null argument");
    const Viewpoint& c = s->scene.view;
    std::memcpy(position, &c.eye, 12);
    std::memcpy(target, &c.aim, 12);
    std::memcpy(up, &c.sky, 12);
    *vfov = c.fov_deg;
    return 0;
}

int phs_scene_camera_view(const phs_scene* s, float aspect, ph_camera_view* out) {
    if (!s || !out) return fail("null argument");
    *out = camera_frame(s->scene.view, aspect);
    return 0;
}

int phs_scene_build(phs_scene* s, int bins, int max_leaf, float traversal_cost, int threads) {
    if (!s) return fail("null argument");
    if (bins > 64) return fail("at most 64 bins");
    return guarded([&] {
        BvhSettings o;
        if (bins > 0) o.bin_count = bins;
        if (max_leaf > 0) o.leaf_limit = max_leaf;
        if (traversal_cost > 0) o.visit_cost = traversal_cost;
        o.workers = threads > 0 ? threads : 0;
        s->scene.compile(o);
        return 0;
    });
}

int phs_scene_rebuild_lights(phs_scene* s) {
    if (!s || !s->scene.out.done) return fail("the scene is not built");
    return guarded([&] {
        s->scene.relight();
        return 0;
    });
}

int phs_scene_desc(const phs_scene* s, ph_scene_desc* out) {
    if (!s || !out) return fail("null argument");
    if (!s->scene.out.done) return fail("the scene is not built");
    *out = s->scene.describe();
    return 0;
}

int phs_scene_num_tris(const phs_scene* s) { return s ? s->scene.mesh.triangles() : fail("null argument"); }
int phs_scene_num_materials(const phs_scene* s) { return s ? int(s->scene.materials.size()) : fail("null argument"); }
const int32_t* phs_scene_order(const phs_scene* s) { return s && s->scene.out.done ? s->scene.out.tree.order.data() : nullptr; }

int phs_scene_stats(const phs_scene* s, phs_bvh_stats* out) {
    if (!s || !out) return fail("null argument");
    const BvhReport& st = s->scene.out.tree.report;
    out->nodes = st.nodes;
    out->leaves = st.leaves;
    out->max_depth = st.depth;
    out->sah_cost = st.sah;
    out->build_ms = st.ms;
    return 0;
}

int64_t phs_scene_device_bytes(const phs_scene* s) {
    if (!s || !s->scene.out.done) return fail("the scene is not built");
    const ph_scene_desc d = s->scene.describe();
    const int64_t nv = d.num_vertices, tris = d.num_tris;
    int64_t b = int64_t(d.num_nodes) * 64 + tris * (48 + 12);
    b += (d.normals ? nv * 12 : 0) + (d.uvs ? nv * 8 : 0) + (d.tangents ? nv * 16 : 0);
    b += int64_t(d.num_materials) * 80 + int64_t(d.num_textures) * 20 + d.num_texels * 4 + 256 * 4;
    b += int64_t(d.num_emissive) * (8 + 16) + (d.tri_emissive ? tris * 4 : 0) + int64_t(d.num_punctual) * 48;
    if (d.env_texels) b += int64_t(d.env_width) * d.env_height * 12 + (int64_t(d.env_height) + 1 + int64_t(d.env_height) * (d.env_width + 1)) * 4;
    return b;
}

int phs_scene_raw(phs_scene* s, phs_raw_scene* out) {
    if (!s || !out) return fail("null argument");
    const Scene& c = s->scene;
    const Mesh& m = c.mesh;
    *out = phs_raw_scene{};
    out->positions = m.pos.empty() ? nullptr : &m.pos[0].x;
    out->num_vertices = int(m.pos.size());
    out->normals = m.nrm.empty() ? nullptr : &m.nrm[0].x;
    out->uvs = m.uv.empty() ? nullptr : &m.uv[0].x;
    out->tangents = m.tan.empty() ? nullptr : &m.tan[0].x;
    out->indices = m.corner.data();
    out->num_tris = m.triangles();
    out->tri_material = m.mat.data();
    out->materials = c.materials.data();
    out->num_materials = int(c.materials.size());
    s->tex_w.clear(), s->tex_h.clear(), s->tex_repeat.clear(), s->tex_srgb.clear(), s->tex_ptr.clear();
    for (const Image8& t : c.images) {
        s->tex_w.push_back(t.width);
        s->tex_h.push_back(t.height);
        s->tex_repeat.push_back(t.tile ? 1 : 0);
        s->tex_srgb.push_back(t.srgb ? 1 : 0);
        s->tex_ptr.push_back(t.px.data());
    }
    out->num_textures = int(c.images.size());
    out->tex_width = s->tex_w.data();
    out->tex_height = s->tex_h.data();
    out->tex_texels = s->tex_ptr.data();
    out->tex_repeat = s->tex_repeat.data();
    out->tex_srgb = s->tex_srgb.data();
    out->punctual = c.lamps.empty() ? nullptr : c.lamps.data();
    out->num_punctual = int(c.lamps.size());
    const Backdrop& b = c.backdrop;
    out->env_texels = b.rgb.empty() ? nullptr : &b.rgb[0].x;
    out->env_width = b.width;
    out->env_height = b.height;
    std::memcpy(out->env_color, &b.flat, 12);
    out->env_intensity = b.gain;
    out->env_rotation = b.yaw;
    const Viewpoint& v = c.view;
    std::memcpy(out->cam_position, &v.eye, 12);
    std::memcpy(out->cam_target, &v.aim, 12);
    std::memcpy(out->cam_up, &v.sky, 12);
    out->cam_vfov = v.fov_deg;
    out->cam_aperture = v.lens;
    out->cam_focus = v.focus;
    return 0;
}

int phs_scene_load_gltf(phs_scene* s, const char* path) {
    if (!s || !path) return fail("null argument");
    return guarded([&] {
        std::string error;
        return load_gltf(path, s->scene, error) ? 0 : fail(error);
    });
}

int phs_scene_save_gltf(const phs_scene* s, const char* path) {
    if (!s || !path) return fail("null argument");
    return guarded([&] {
        std::string error;
        return save_gltf(path, s->scene, error) ? 0 : fail(error);
    });
}

int phs_scene_num_cameras(const phs_scene* s) { return s ? int(s->scene.file_views.size()) : fail("null argument"); }

int phs_scene_camera_at(const phs_scene* s, int index, float position[3], float target[3], float up[3], float* vfov) {
    if (!s || !position || !target || !up || !vfov) return fail("null argument");
    if (index < 0 || size_t(index) >= s->scene.file_views.size()) return fail("camera index out of range");
    const Viewpoint& c = s->scene.file_views[size_t(index)];
    std::memcpy(position, &c.eye, 12);
    std::memcpy(target, &c.aim, 12);
    std::memcpy(up, &c.sky, 12);
    *vfov = c.fov_deg;
    return 0;
}

int phs_scene_load_environment(phs_scene* s, const char* path, float intensity, float rotation) {
    if (!s || !path) return fail("null argument");
    return guarded([&] {
        Backdrop env;
        std::string error;
        if (!load_environment(path, env, error)) return fail(error);
        env.gain = intensity;
        env.yaw = rotation;
        s->scene.backdrop = std::move(env);
        return 0;
    });
}

int phs_image_decode(const uint8_t* data, size_t size, int* width, int* height, uint32_t* rgba, size_t capacity) {
    if (!data || !width || !height) return fail("null argument");
    return guarded([&] {
        Image8 t;
        std::string error;
        if (!decode_image(data, size, t, error)) return fail(error);
        *width = t.width;
        *height = t.height;
        if (rgba) {
            if (capacity < t.px.size()) return fail("output buffer too small");
            std::memcpy(rgba, t.px.data(), t.px.size() * 4);
        }
        return 0;
    });
}

int phs_image_save_png(const char* path, int width, int height, const uint32_t* rgba) {
    if (!path || !rgba || width <= 0 || height <= 0) return fail("invalid argument");
    return guarded([&] { return save_png(path, width, height, rgba) ? 0 : fail(std::string("cannot write ") + path); });
}

int phs_image_save_pfm(const char* path, int width, int height, const float* rgb) {
    if (!path || !rgb || width <= 0 || height <= 0) return fail("invalid argument");
    return guarded([&] { return save_pfm(path, width, height, rgb) ? 0 : fail(std::string("cannot write ") + path); });
}

int phs_image_load_pfm(const char* path, int* width, int* height, float* rgb, size_t capacity) {
    if (!path || !width || !height) return fail("null argument");
    return guarded([&] {
        std::vector<float> px;
        if (!load_pfm(path, *width, *height, px)) return fail(std::string("cannot read PFM file ") + path);
        if (rgb) {
            if (capacity < px.size()) return fail("output buffer too small");
            std::memcpy(rgb, px.data(), px.size() * 4);
        }
        return 0;
    });
}

int64_t phs_json_canonical(const char* text, char* out, size_t capacity) {
    if (!text) return fail("null argument");
    try {
        const json::Document doc = json::Document::read(text);
        const std::string s = json::canonical(doc.root());
        if (out && capacity > s.size()) std::memcpy(out, s.c_str(), s.size() + 1);
        return int64_t(s.size());
    } catch (const json::SyntaxError& e) {
        return fail(e.what());
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

int phs_camera_rays(const ph_camera_view* cam, int width, int height, ph_ray* out) {
    if (!cam || !out || width <= 0 || height <= 0) return fail("invalid argument");
    vec3 fwd = v3(cam->forward), right = v3(cam->right), up = v3(cam->up), pos = v3(cam->position);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            float px = float(x) + 0.5f, py = float(y) + 0.5f;
            float sx = (2.0f * px / float(width) - 1.0f) * cam->tan_half_fov * cam->aspect;
            float sy = (1.0f - 2.0f * py / float(height)) * cam->tan_half_fov;
            vec3 d = normalize(fwd + right * sx + up * sy);
            ph_ray& r = out[size_t(y) * size_t(width) + size_t(x)];
            std::memcpy(r.o, &pos, 12);
            std::memcpy(r.d, &d, 12);
            r.tmax = kInf;
        }
    return 0;
}

int phs_sphere_shadow_rays(const phs_scene* s, int count, uint64_t seed, ph_ray* out) {
    if (!s || !out || count < 0) return fail("invalid argument");
    const std::vector<vec4>& sp = s->scene.sphere_meta;
    if (sp.size() < 2) return fail("the scene has fewer than two procedural spheres");
    Pcg32 rng(seed, 0x5AD0u);
    const int n = int(sp.size());
    auto on_sphere = [&](vec4 c, float scale) {
        // Uniform direction from two draws (z = 1 - 2u, phi = 2 pi v), in double for portability.
        double z = 1.0 - 2.0 * double(rng.next());
        double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        double phi = 2.0 * 3.14159265358979323846 * double(rng.next());
        vec3 d(float(r * std::cos(phi)), float(r * std::sin(phi)), float(z));
        return vec3(c.x, c.y, c.z) + d * (c.w * scale);
    };
    for (int i = 0; i < count; i++) {
        int a = std::min(n - 1, int(rng.next() * float(n)));
        int b = std::min(n - 1, int(rng.next() * float(n)));
        if (b == a) b = (a + 1) % n;
        vec3 o = on_sphere(sp[size_t(a)], 1.06f);
        vec3 t = on_sphere(sp[size_t(b)], 1.0f);
        vec3 d = t - o;
        float dist = length(d);
        vec3 dir = d / dist;
        ph_ray& r = out[size_t(i)];
        std::memcpy(r.o, &o, 12);
        std::memcpy(r.d, &dir, 12);
        r.tmax = dist * (1.0f - 1e-4f);
    }
    return 0;
}

}  // extern "C"
