// Host producers of the device inputs. Each function restates the reference's host-side construction
// with the same float/double evaluation order, because its outputs are compared bit for bit:
//   Camera::view            <- src/scene/scene.cpp:13-25
//   Scene::add_mesh         <- src/scene/scene.cpp:27-42
//   compute_tangents        <- src/scene/scene.cpp:50-80 (Lengyel's per-vertex tangents)
//   build_alias_table       <- src/scene/scene.cpp:82-111 (Vose's alias method in double)
//   build_lights            <- src/scene/scene.cpp:113-176 (emitter power, group probabilities, 0.05 floor)
//   build_environment       <- src/scene/scene.cpp:180-210 (luminance * sin(theta) CDFs)
//   Scene::build            <- src/scene/scene.cpp:212-255
#include "scene.hpp"

#include <algorithm>
#include <cmath>

namespace phh {

ph_material default_material() {
    ph_material m{};
    m.base_color[0] = m.base_color[1] = m.base_color[2] = m.base_color[3] = 1.0f;
    m.metallic = 1.0f;
    m.roughness = 1.0f;
    m.ior = 1.5f;
    m.transmission = 0.0f;
    m.normal_scale = 1.0f;
    m.alpha_cutoff = 0.5f;
    m.alpha_mode = PH_ALPHA_OPAQUE;
    m.double_sided = 0;
    m.base_color_tex = m.metallic_roughness_tex = m.normal_tex = m.emissive_tex = m.occlusion_tex = -1;
    return m;
}

static void put3(float* dst, vec3 v) {
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
}

ph_camera_view Camera::view(float aspect) const {
    ph_camera_view v{};
    vec3 fwd = normalize(target - position);
    vec3 right = normalize(cross(fwd, up));
    if (length_squared(right) < 1e-12f) right = vec3(1, 0, 0);
    vec3 upv = cross(right, fwd);
    put3(v.position, position);
    put3(v.forward, fwd);
    put3(v.right, right);
    put3(v.up, upv);
    v.tan_half_fov = std::tan(vfov * 0.5f * kPi / 180.0f);
    v.aspect = aspect;
    v.aperture = aperture;
    v.focus_distance = focus_distance > 0 ? focus_distance : length(target - position);
    return v;
}

void Scene::add_mesh(const std::vector<vec3>& pos, const std::vector<int>& idx, int material, const std::vector<vec3>& nrm,
                     const std::vector<vec2>& uv) {
    const size_t base = positions.size();
    // A per-vertex attribute array is either empty or complete: pad it when one mesh has the attribute.
    auto append = [&](auto& arr, const auto& src, auto zero) {
        if (src.empty() && arr.empty()) return;
        arr.resize(base, zero);
        if (src.empty())
            arr.resize(base + pos.size(), zero);
        else
            arr.insert(arr.end(), src.begin(), src.end());
    };
    append(normals, nrm, vec3(0));
    append(uvs, uv, vec2(0, 0));
    positions.insert(positions.end(), pos.begin(), pos.end());
    for (int i : idx) indices.push_back(int(base) + i);
    for (size_t t = 0; t < idx.size() / 3; t++) tri_material.push_back(material);
}

int Scene::add_material(const ph_material& m) {
    materials.push_back(m);
    return int(materials.size()) - 1;
}

void Scene::compute_tangents() {
    if (uvs.empty() || !tangents.empty()) return;
    bool needed = false;
    for (const ph_material& m : materials) needed = needed || m.normal_tex >= 0;
    if (!needed) return;
    const size_t nv = positions.size();
    std::vector<vec3> sdir_sum(nv, vec3(0)), tdir_sum(nv, vec3(0));
    for (size_t t = 0; t + 2 < indices.size(); t += 3) {
        int a = indices[t], b = indices[t + 1], c = indices[t + 2];
        vec3 e1 = positions[size_t(b)] - positions[size_t(a)], e2 = positions[size_t(c)] - positions[size_t(a)];
        vec2 d1 = uvs[size_t(b)] - uvs[size_t(a)], d2 = uvs[size_t(c)] - uvs[size_t(a)];
        float det = d1.x * d2.y - d2.x * d1.y;
        if (std::fabs(det) < 1e-20f) continue;
        float r = 1.0f / det;
        vec3 sdir = (e1 * d2.y - e2 * d1.y) * r;
        vec3 tdir = (e2 * d1.x - e1 * d2.x) * r;
        for (int v : {a, b, c}) {
            sdir_sum[size_t(v)] += sdir;
            tdir_sum[size_t(v)] += tdir;
        }
    }
    tangents.resize(nv);
    for (size_t v = 0; v < nv; v++) {
        vec3 n = normals.empty() ? vec3(0, 0, 1) : normals[v];
        vec3 t = sdir_sum[v] - n * dot(n, sdir_sum[v]);
        if (length_squared(t) < 1e-20f) t = basis_tangent(normalize(n));
        t = normalize(t);
        float w = dot(cross(n, t), tdir_sum[v]) < 0 ? -1.0f : 1.0f;
        tangents[v] = vec4(t, w);
    }
}

std::vector<ph_alias_entry> build_alias_table(const std::vector<float>& weights) {
    const size_t n = weights.size();
    std::vector<ph_alias_entry> table(n);
    double total = 0;
    for (float w : weights) total += std::max(0.0f, w);
    if (n == 0) return table;
    std::vector<double> scaled(n);
    std::vector<size_t> under, over;
    for (size_t i = 0; i < n; i++) {
        double p = total > 0 ? std::max(0.0f, weights[i]) / total : 1.0 / double(n);
        table[i].pmf = float(p);
        table[i].alias = int(i);
        table[i].pad = 0;
        scaled[i] = p * double(n);
        (scaled[i] < 1.0 ? under : over).push_back(i);
    }
    while (!under.empty() && !over.empty()) {
        size_t s = under.back(), l = over.back();
        under.pop_back();
        table[s].prob = float(scaled[s]);
        table[s].alias = int(l);
        scaled[l] = scaled[l] + scaled[s] - 1.0;
        if (scaled[l] < 1.0) {
            over.pop_back();
            under.push_back(l);
        }
    }
    for (size_t i : under) table[i].prob = 1.0f;
    for (size_t i : over) table[i].prob = 1.0f;
    return table;
}

namespace {

// Bilinear RGBA8 lookup used for textured emitters (restates src/render/texture.h:9-44).
vec4 texel_rgba(uint32_t p, bool srgb, const std::vector<float>& lut) {
    if (srgb) return {lut[p & 0xFF], lut[(p >> 8) & 0xFF], lut[(p >> 16) & 0xFF], float(p >> 24) * (1.0f / 255.0f)};
    return vec4(float(p & 0xFF), float((p >> 8) & 0xFF), float((p >> 16) & 0xFF), float(p >> 24)) * (1.0f / 255.0f);
}

int wrap(int i, int n, bool repeat) {
    if (repeat) {
        i %= n;
        return i < 0 ? i + n : i;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

vec4 lookup(const std::vector<ph_texture_info>& infos, const std::vector<uint32_t>& pool, const std::vector<float>& lut, int tex, vec2 uv) {
    const ph_texture_info& t = infos[size_t(tex)];
    float x = uv.x * float(t.width) - 0.5f, y = uv.y * float(t.height) - 0.5f;
    float fx = std::floor(x), fy = std::floor(y);
    if (!(std::fabs(fx) < 1e8f) || !(std::fabs(fy) < 1e8f)) return vec4(0, 0, 0, 1);
    int x0 = int(fx), y0 = int(fy);
    float tx = x - fx, ty = y - fy;
    bool rep = t.wrap_repeat != 0;
    int xa = wrap(x0, t.width, rep), xb = wrap(x0 + 1, t.width, rep);
    int ya = wrap(y0, t.height, rep), yb = wrap(y0 + 1, t.height, rep);
    const uint32_t* base = pool.data() + t.offset;
    bool srgb = t.srgb != 0;
    vec4 a = texel_rgba(base[ya * t.width + xa], srgb, lut);
    vec4 b = texel_rgba(base[ya * t.width + xb], srgb, lut);
    vec4 c = texel_rgba(base[yb * t.width + xa], srgb, lut);
    vec4 d = texel_rgba(base[yb * t.width + xb], srgb, lut);
    vec4 top = a * (1 - tx) + b * tx;
    vec4 bottom = c * (1 - tx) + d * tx;
    return top * (1 - ty) + bottom * ty;
}

}  // namespace

void Scene::build_lights() {
    emissive.clear();
    tri_emissive.assign(tris.size(), -1);
    std::vector<float> power;
    emissive_power = 0;
    static const float probe[7][2] = {{1.f / 3, 1.f / 3}, {0.1f, 0.1f}, {0.8f, 0.1f}, {0.1f, 0.8f}, {0.45f, 0.1f}, {0.45f, 0.45f}, {0.1f, 0.45f}};
    for (size_t i = 0; i < tris.size(); i++) {
        const ph_material& m = materials[size_t(tris[i].material)];
        vec3 e(m.emission[0], m.emission[1], m.emission[2]);
        float lum = luminance(e);
        if (lum <= 0) continue;
        if (m.emissive_tex >= 0 && !uvs.empty()) {
            // Textured emitter: mean of seven lookups; dark triangles are left to BSDF sampling.
            float sum = 0;
            const int* idx = &bvh_indices[3 * i];
            for (const auto& b : probe) {
                float u = b[0], v = b[1];
                vec2 uv = uvs[size_t(idx[0])] * (1 - u - v) + uvs[size_t(idx[1])] * u + uvs[size_t(idx[2])] * v;
                vec3 tc = lookup(texture_infos, texel_pool, srgb_lut, m.emissive_tex, uv).xyz();
                sum += luminance(e * tc);
            }
            lum = sum / 7.0f;
            if (lum <= 1e-4f) continue;
        }
        vec3 e1(tris[i].e1[0], tris[i].e1[1], tris[i].e1[2]), e2(tris[i].e2[0], tris[i].e2[1], tris[i].e2[2]);
        float area = 0.5f * length(cross(e1, e2));
        if (area <= 0) continue;
        tri_emissive[i] = int(emissive.size());
        emissive.push_back({int(i), area});
        float p = area * lum * kPi * (m.double_sided ? 2.0f : 1.0f);
        power.push_back(p);
        emissive_power += p;
    }
    emissive_alias = build_alias_table(power);

    const Box& b = bvh.bounds;
    float radius = b.empty() ? 1.0f : 0.5f * length(b.hi - b.lo);
    float env_power = 0;
    if (!env.texels.empty()) {
        double sum = 0;
        for (const vec3& t : env.texels) sum += luminance(t);
        env_power = float(sum / double(env.texels.size())) * env.intensity * kPi * kPi * radius * radius * 4.0f;
    } else {
        env_power = luminance(env.color) * env.intensity * kPi * kPi * radius * radius * 4.0f;
    }
    float punctual_power = 0;
    for (const ph_punctual_light& l : punctual) {
        float lum = luminance(vec3(l.intensity[0], l.intensity[1], l.intensity[2]));
        punctual_power += l.type == PH_LIGHT_DIRECTIONAL ? lum * kPi * radius * radius : lum * 4.0f * kPi;
    }
    float total = env_power + emissive_power + punctual_power;
    p_env = p_emissive = p_punctual = 0;
    if (total > 0) {
        p_env = env_power / total;
        p_emissive = emissive_power / total;
        p_punctual = punctual_power / total;
        // Every present group keeps a 5% floor: the power estimates ignore occlusion.
        if (env_power > 0) p_env = std::max(p_env, 0.05f);
        if (emissive_power > 0) p_emissive = std::max(p_emissive, 0.05f);
        if (punctual_power > 0) p_punctual = std::max(p_punctual, 0.05f);
        float s = p_env + p_emissive + p_punctual;
        p_env /= s;
        p_emissive /= s;
        p_punctual /= s;
    }
}

void Scene::build_environment() {
    env_marginal.clear();
    env_conditional.clear();
    env_integral = 0;
    if (env.texels.empty()) return;
    const int w = env.width, h = env.height;
    env_conditional.resize(size_t(h) * size_t(w + 1));
    env_marginal.resize(size_t(h) + 1);
    std::vector<double> row_sum(static_cast<size_t>(h));
    for (int y = 0; y < h; y++) {
        float sin_theta = std::sin((float(y) + 0.5f) / float(h) * kPi);
        float* cdf = &env_conditional[size_t(y) * size_t(w + 1)];
        double acc = 0;
        cdf[0] = 0;
        for (int x = 0; x < w; x++) {
            acc += double(luminance(env.texels[size_t(y) * size_t(w) + size_t(x)])) * sin_theta;
            cdf[x + 1] = float(acc);
        }
        row_sum[size_t(y)] = acc;
        for (int x = 1; x <= w; x++) cdf[x] = acc > 0 ? float(cdf[x] / acc) : float(x) / float(w);
    }
    double acc = 0;
    env_marginal[0] = 0;
    for (int y = 0; y < h; y++) {
        acc += row_sum[size_t(y)];
        env_marginal[size_t(y) + 1] = float(acc);
    }
    for (int y = 1; y <= h; y++) env_marginal[size_t(y)] = acc > 0 ? float(env_marginal[size_t(y)] / acc) : float(y) / float(h);
    env_integral = float(acc / (double(w) * double(h)));
}

void Scene::build(const BvhOptions& options) {
    compute_tangents();
    if (materials.empty()) materials.push_back(default_material());
    const int n = num_tris();
    std::vector<vec3> verts(size_t(n) * 3);
    for (int t = 0; t < n; t++)
        for (int k = 0; k < 3; k++) verts[size_t(t) * 3 + size_t(k)] = positions[size_t(indices[size_t(t) * 3 + size_t(k)])];
    bvh = build_bvh(verts, options);
    tris.resize(size_t(n));
    bvh_indices.resize(size_t(n) * 3);
    for (int i = 0; i < n; i++) {
        int t = bvh.order[size_t(i)];
        vec3 a = verts[size_t(t) * 3], b = verts[size_t(t) * 3 + 1], c = verts[size_t(t) * 3 + 2];
        ph_tri_geom g{};
        put3(g.v0, a);
        put3(g.e1, b - a);
        put3(g.e2, c - a);
        g.material = std::clamp(tri_material[size_t(t)], 0, int(materials.size()) - 1);
        g.flags = materials[size_t(g.material)].alpha_mode != PH_ALPHA_OPAQUE ? PH_TRI_ALPHA_TESTED : 0;
        tris[size_t(i)] = g;
        for (int k = 0; k < 3; k++) bvh_indices[size_t(i) * 3 + size_t(k)] = indices[size_t(t) * 3 + size_t(k)];
    }
    texel_pool.clear();
    texture_infos.clear();
    for (const Texture& t : textures) {
        ph_texture_info info{};
        info.width = t.width;
        info.height = t.height;
        info.offset = uint32_t(texel_pool.size());
        info.wrap_repeat = t.repeat ? 1 : 0;
        info.srgb = t.srgb ? 1 : 0;
        texel_pool.insert(texel_pool.end(), t.texels.begin(), t.texels.end());
        texture_infos.push_back(info);
    }
    srgb_lut.resize(256);
    for (int i = 0; i < 256; i++) srgb_lut[size_t(i)] = srgb_to_linear(float(i) / 255.0f);
    build_lights();
    build_environment();
    built = true;
}

void Scene::rebuild_lights() {
    for (ph_tri_geom& g : tris) g.flags = materials[size_t(g.material)].alpha_mode != PH_ALPHA_OPAQUE ? PH_TRI_ALPHA_TESTED : 0;
    build_lights();
    build_environment();
}

ph_scene_desc Scene::desc() const {
    ph_scene_desc d{};
    d.nodes = bvh.nodes.data();
    d.num_nodes = int(bvh.nodes.size());
    d.tris = tris.data();
    d.num_tris = int(tris.size());
    d.indices = bvh_indices.data();
    d.normals = normals.empty() ? nullptr : &normals[0].x;
    d.uvs = uvs.empty() ? nullptr : &uvs[0].x;
    d.tangents = tangents.empty() ? nullptr : &tangents[0].x;
    d.num_vertices = int(positions.size());
    d.materials = materials.data();
    d.num_materials = int(materials.size());
    d.textures = texture_infos.empty() ? nullptr : texture_infos.data();
    d.num_textures = int(texture_infos.size());
    d.texels = texel_pool.empty() ? nullptr : texel_pool.data();
    d.num_texels = int64_t(texel_pool.size());
    d.srgb_lut = srgb_lut.data();
    d.emissive = emissive.empty() ? nullptr : emissive.data();
    d.tri_emissive = tri_emissive.data();
    d.emissive_alias = emissive_alias.empty() ? nullptr : emissive_alias.data();
    d.num_emissive = int(emissive.size());
    d.emissive_power = emissive_power;
    d.punctual = punctual.empty() ? nullptr : punctual.data();
    d.num_punctual = int(punctual.size());
    d.env_texels = env.texels.empty() ? nullptr : &env.texels[0].x;
    d.env_marginal = env_marginal.empty() ? nullptr : env_marginal.data();
    d.env_conditional = env_conditional.empty() ? nullptr : env_conditional.data();
    d.env_width = env.width;
    d.env_height = env.height;
    d.env_intensity = env.intensity;
    d.env_rotation = env.rotation;
    put3(d.env_color, env.color);
    d.env_marginal_integral = env_integral;
    d.p_env = p_env;
    d.p_emissive = p_emissive;
    d.p_punctual = p_punctual;
    put3(d.bounds_min, bvh.bounds.lo);
    put3(d.bounds_max, bvh.bounds.hi);
    return d;
}

}  // namespace phh
