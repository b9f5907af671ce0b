// Device-resident scene: the flat arrays of a built scene in HBM, and their accessors.
//
// Layout in HBM (one allocation per array, 256-byte aligned by the allocator):
//   nodes      64 B each, read as four 16-byte loads (dual-box node: both children's bounds at once)
//   tris       48 B each, BVH order (v0|material, e1|flags, e2|pad), three 16-byte loads
//   indices    12 B per triangle; normals 12 B, uvs 8 B, tangents 16 B per vertex
//   materials  80 B each, read as five 16-byte loads
//   texels     RGBA8 pool; srgb_lut 256 floats
//   emissive / tri_emissive / alias: light tables; env texels (12 B) and CDFs
// The scene on the CPU side is identical byte for byte (include/photon_types.h).
#pragma once

#include "dmath.h"
#include "photon_types.h"

namespace phd {

struct DScene {
    const float4* nodes;
    const float4* tris;
    const int* indices;
    const float* normals;
    const float* uvs;
    const float* tangents;
    const float4* materials;
    const ph_texture_info* textures;
    const uint32_t* texels;
    const float* srgb_lut;
    const ph_emissive_tri* emissive;
    const int* tri_emissive;
    const ph_alias_entry* alias;
    const ph_punctual_light* punctual;
    const float* env_texels;
    const float* env_marginal;
    const float* env_conditional;
    // Guide tables of the environment CDFs: entry k of a table with m + 1 entries (m a power of two) is
    // the index the CDF search returns for u = k / m, so a search for u starts between entries
    // floor(u m) and floor(u m) + 1. Null when the scene has no environment map.
    const int* env_guide_marginal;     // env_guide_mh + 1 entries
    const int* env_guide_conditional;  // env_h rows of env_guide_mw + 1 entries
    int env_guide_mh, env_guide_mw;
    int num_tris, num_nodes, num_emissive, num_punctual;
    int top_nodes;  // nodes [0, top_nodes) are the top of the tree in breadth-first order
    int env_w, env_h;
    float env_intensity, env_rotation, env_integral;
    float env_color[3];
    float p_env, p_emissive, p_punctual, emissive_power;
    int num_materials, num_textures;
};

// Material fields as the shading code uses them.
struct Mat {
    float4 base_color;
    f3 emission;
    float metallic, roughness, ior, transmission, normal_scale, alpha_cutoff;
    int alpha_mode, double_sided, base_color_tex, mr_tex, normal_tex, emissive_tex;
};

PH_DEV Mat load_material(const DScene& s, int m) {
    const float4* p = s.materials + 5 * m;
    float4 a = ldg(p), b = ldg(p + 1), c = ldg(p + 2), d = ldg(p + 3), e = ldg(p + 4);
    Mat r;
    r.base_color = a;
    r.emission = mk3(b.x, b.y, b.z);
    r.metallic = b.w;
    r.roughness = c.x;
    r.ior = c.y;
    r.transmission = c.z;
    r.normal_scale = c.w;
    r.alpha_cutoff = d.x;
    r.alpha_mode = __float_as_int(d.y);
    r.double_sided = __float_as_int(d.z);
    r.base_color_tex = __float_as_int(d.w);
    r.mr_tex = __float_as_int(e.x);
    r.normal_tex = __float_as_int(e.y);
    r.emissive_tex = __float_as_int(e.z);
    return r;
}

struct Tri {
    f3 v0, e1, e2;
    int material, flags;
};

PH_DEV Tri load_tri(const DScene& s, int i) {
    const float4* p = s.tris + 3 * i;
    float4 a = ldg(p), b = ldg(p + 1), c = ldg(p + 2);
    return {xyz(a), xyz(b), xyz(c), __float_as_int(a.w), __float_as_int(b.w)};
}

PH_DEV f2 uv_at(const float* uvs, int i) { return {ldg(uvs + 2 * i), ldg(uvs + 2 * i + 1)}; }
PH_DEV f3 n_at(const float* n, int i) { return {ldg(n + 3 * i), ldg(n + 3 * i + 1), ldg(n + 3 * i + 2)}; }
PH_DEV f4 t_at(const float* t, int i) {
    float4 v = ldg(reinterpret_cast<const float4*>(t) + i);
    return {v.x, v.y, v.z, v.w};
}
PH_DEV int index_at(const int* idx, int i) { return ldg(idx + i); }

// ---- textures (texture.h:9-44) ------------------------------------------------------------------------
PH_DEV f4 texel(uint32_t p, bool srgb, const float* lut) {
    float r = float(p & 0xFF), g = float((p >> 8) & 0xFF), b = float((p >> 16) & 0xFF), a = float(p >> 24);
    if (srgb) return {ldg(lut + (p & 0xFF)), ldg(lut + ((p >> 8) & 0xFF)), ldg(lut + ((p >> 16) & 0xFF)), a * (1.0f / 255.0f)};
    return f4{r, g, b, a} * (1.0f / 255.0f);
}

PH_DEV int wrap_index(int i, int n, bool repeat) {
    if (repeat) {
        i %= n;
        return i < 0 ? i + n : i;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

PH_DEV f4 tex_bilinear(const DScene& s, int tex, f2 uv) {
    const ph_texture_info t = ldg(s.textures + tex);
    float x = uv.x * float(t.width) - 0.5f, y = uv.y * float(t.height) - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    if (!(fabsf(fx) < 1e8f) || !(fabsf(fy) < 1e8f)) return {0, 0, 0, 1};
    int x0 = int(fx), y0 = int(fy);
    float tx = x - fx, ty = y - fy;
    bool rep = t.wrap_repeat != 0;
    int xa = wrap_index(x0, t.width, rep), xb = wrap_index(x0 + 1, t.width, rep);
    int ya = wrap_index(y0, t.height, rep), yb = wrap_index(y0 + 1, t.height, rep);
    const uint32_t* base = s.texels + t.offset;
    bool srgb = t.srgb != 0;
    f4 a = texel(ldg(base + ya * t.width + xa), srgb, s.srgb_lut);
    f4 b = texel(ldg(base + ya * t.width + xb), srgb, s.srgb_lut);
    f4 c = texel(ldg(base + yb * t.width + xa), srgb, s.srgb_lut);
    f4 d = texel(ldg(base + yb * t.width + xb), srgb, s.srgb_lut);
    f4 top = a * (1 - tx) + b * tx;
    f4 bottom = c * (1 - tx) + d * tx;
    return top * (1 - ty) + bottom * ty;
}

}  // namespace phd
