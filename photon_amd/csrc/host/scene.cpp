// Scene authoring helpers and the compile pipeline: tangents -> BVH -> BVH-order triangle records ->
// texel pool -> light tables -> backdrop tables, then the flat descriptor the device library uploads.
#include "scene.hpp"

#include <algorithm>
#include <cmath>

namespace phh {

namespace {

void store3(float* dst, vec3 v) {
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
}

// Pads `stream` with zeros up to `have` entries and then either appends `src` or `count` more zeros; a
// stream no part of the mesh has stays empty.
template <class T>
void extend_stream(std::vector<T>& stream, const std::vector<T>& src, size_t have, size_t count) {
    if (stream.empty() && src.empty()) return;
    stream.resize(have, T{});
    if (src.empty())
        stream.resize(have + count, T{});
    else
        stream.insert(stream.end(), src.begin(), src.end());
}

int alpha_flag(const ph_material& m) { return m.alpha_mode != PH_ALPHA_OPAQUE ? PH_TRI_ALPHA_TESTED : 0; }

}  // namespace

ph_material default_material() {
    ph_material m{};
    std::fill(m.base_color, m.base_color + 4, 1.0f);
    m.metallic = m.roughness = 1.0f;
    m.ior = 1.5f;
    m.normal_scale = 1.0f;
    m.alpha_cutoff = 0.5f;
    m.alpha_mode = PH_ALPHA_OPAQUE;
    m.base_color_tex = m.metallic_roughness_tex = m.normal_tex = m.emissive_tex = m.occlusion_tex = -1;
    return m;
}

ph_camera_view camera_frame(const Viewpoint& v, float aspect) {
    const vec3 to_aim = v.aim - v.eye;
    const vec3 fwd = normalize(to_aim);
    vec3 right = normalize(cross(fwd, v.sky));
    if (length_squared(right) < 1e-12f) right = vec3(1, 0, 0);
    ph_camera_view c{};
    store3(c.position, v.eye);
    store3(c.forward, fwd);
    store3(c.right, right);
    store3(c.up, cross(right, fwd));
    c.tan_half_fov = std::tan(v.fov_deg * 0.5f * kPi / 180.0f);
    c.aspect = aspect;
    c.aperture = v.lens;
    c.focus_distance = v.focus > 0 ? v.focus : length(to_aim);
    return c;
}

void append_mesh(Mesh& m, const std::vector<vec3>& p, const std::vector<int>& idx, int material, const std::vector<vec3>& n,
                 const std::vector<vec2>& t) {
    const size_t first = m.pos.size();
    extend_stream(m.nrm, n, first, p.size());
    extend_stream(m.uv, t, first, p.size());
    m.pos.insert(m.pos.end(), p.begin(), p.end());
    // resize keeps the vector's geometric growth; reserving the exact new size on every call would
    // reallocate and copy the whole corner list once per appended mesh.
    const size_t at = m.corner.size();
    m.corner.resize(at + idx.size());
    for (size_t k = 0; k < idx.size(); k++) m.corner[at + k] = int(first) + idx[k];
    m.mat.insert(m.mat.end(), idx.size() / 3, material);
}

int Scene::add_material(const ph_material& m) {
    materials.push_back(m);
    return int(materials.size() - 1);
}

int Scene::add_image(Image8 img) {
    images.push_back(std::move(img));
    return int(images.size() - 1);
}

void Scene::compile(const BvhSettings& settings) {
    fill_tangents(mesh, materials);
    if (materials.empty()) materials.push_back(default_material());
    const int tris = mesh.triangles();
    const int last_material = int(materials.size()) - 1;

    // The tree over the triangles' corner positions.
    std::vector<vec3> corners(size_t(tris) * 3);
    for (size_t k = 0; k < corners.size(); k++) corners[k] = mesh.pos[size_t(mesh.corner[k])];
    out.tree = build_tree(corners, settings);

    // Triangle records and corner indices in BVH order.
    out.tri.assign(size_t(tris), ph_tri_geom{});
    out.corner.resize(size_t(tris) * 3);
    for (int slot = 0; slot < tris; slot++) {
        const size_t src = size_t(out.tree.order[size_t(slot)]);
        const vec3 a = corners[3 * src], b = corners[3 * src + 1], c = corners[3 * src + 2];
        ph_tri_geom& g = out.tri[size_t(slot)];
        store3(g.v0, a);
        store3(g.e1, b - a);
        store3(g.e2, c - a);
        g.material = std::clamp(mesh.mat[src], 0, last_material);
        g.flags = alpha_flag(materials[size_t(g.material)]);
        std::copy_n(mesh.corner.begin() + long(3 * src), 3, out.corner.begin() + long(3 * size_t(slot)));
    }

    // One texel pool; each image keeps its own row of the texture table.
    out.texel_pool.clear();
    out.texture_table.clear();
    for (const Image8& img : images) {
        ph_texture_info info{};
        info.width = img.width;
        info.height = img.height;
        info.offset = uint32_t(out.texel_pool.size());
        info.wrap_repeat = img.tile;
        info.srgb = img.srgb;
        out.texture_table.push_back(info);
        out.texel_pool.insert(out.texel_pool.end(), img.px.begin(), img.px.end());
    }
    out.srgb_lut.resize(256);
    for (int code = 0; code < 256; code++) out.srgb_lut[size_t(code)] = srgb_to_linear(float(code) / 255.0f);

    out.lights = light_tables(*this);
    out.backdrop = backdrop_tables(backdrop);
    out.done = true;
}

void Scene::relight() {
    for (ph_tri_geom& g : out.tri) g.flags = alpha_flag(materials[size_t(g.material)]);
    out.lights = light_tables(*this);
    out.backdrop = backdrop_tables(backdrop);
}

ph_scene_desc Scene::describe() const {
    auto or_null = [](const auto& v) { return v.empty() ? nullptr : v.data(); };
    ph_scene_desc d{};
    d.nodes = or_null(out.tree.nodes);
    d.num_nodes = int(out.tree.nodes.size());
    d.tris = or_null(out.tri);
    d.num_tris = int(out.tri.size());
    d.indices = or_null(out.corner);
    d.num_vertices = int(mesh.pos.size());
    d.normals = mesh.nrm.empty() ? nullptr : &mesh.nrm[0].x;
    d.uvs = mesh.uv.empty() ? nullptr : &mesh.uv[0].x;
    d.tangents = mesh.tan.empty() ? nullptr : &mesh.tan[0].x;
    d.materials = or_null(materials);
    d.num_materials = int(materials.size());
    d.textures = or_null(out.texture_table);
    d.num_textures = int(out.texture_table.size());
    d.texels = or_null(out.texel_pool);
    d.num_texels = int64_t(out.texel_pool.size());
    d.srgb_lut = or_null(out.srgb_lut);
    const LightTables& L = out.lights;
    d.emissive = or_null(L.emitters);
    d.num_emissive = int(L.emitters.size());
    d.tri_emissive = or_null(L.slot_of_tri);
    d.emissive_alias = or_null(L.alias);
    d.emissive_power = L.emitted_power;
    d.punctual = or_null(lamps);
    d.num_punctual = int(lamps.size());
    d.p_env = L.p_backdrop;
    d.p_emissive = L.p_emitters;
    d.p_punctual = L.p_punctual;
    d.env_texels = backdrop.rgb.empty() ? nullptr : &backdrop.rgb[0].x;
    d.env_width = backdrop.width;
    d.env_height = backdrop.height;
    d.env_intensity = backdrop.gain;
    d.env_rotation = backdrop.yaw;
    store3(d.env_color, backdrop.flat);
    d.env_marginal = or_null(out.backdrop.marginal);
    d.env_conditional = or_null(out.backdrop.conditional);
    d.env_marginal_integral = out.backdrop.mean;
    store3(d.bounds_min, out.tree.bounds.low);
    store3(d.bounds_max, out.tree.bounds.high);
    return d;
}

}  // namespace phh
