// ORACLE — test infrastructure only. Nothing in the product (photon_amd/) links, loads or calls this.
//
// The CPU restatement of the reference's path tracing integrator over the flat scene arrays
// (ph_scene_desc). Each block follows the reference function named beside it, with the same
// evaluation order, the same random-number draw order and glibc's float libm, so that it reproduces
// the reference CpuRenderer bit for bit:
//   texel/texture lookup ........... src/render/texture.h:9-44
//   Moller-Trumbore, slab, traversal src/render/traverse.h:24-147
//   GGX mixture BSDF ............... src/render/bsdf.h:45-281
//   light sampling / densities ..... src/render/lights.h:22-202
//   offset, surface, path, camera .. src/render/integrator.h:61-280
//   render_sample .................. src/render/pixel.h:18-29
#pragma once

#include "oracle_math.hpp"
#include "photon_types.h"

namespace orc {

struct Scene {
    const ph_scene_desc& d;
    explicit Scene(const ph_scene_desc& desc) : d(desc) {}
};

// ---- textures ---------------------------------------------------------------------------------------

inline V4 texel(uint32_t p, bool srgb, const float* lut) {
    float r = float(p & 0xFF), g = float((p >> 8) & 0xFF), b = float((p >> 16) & 0xFF), a = float(p >> 24);
    if (srgb) return v4(lut[p & 0xFF], lut[(p >> 8) & 0xFF], lut[(p >> 16) & 0xFF], a * (1.0f / 255.0f));
    return v4(r, g, b, a) * (1.0f / 255.0f);
}

inline int wrap_index(int i, int n, bool repeat) {
    if (repeat) {
        i %= n;
        return i < 0 ? i + n : i;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

inline V4 tex_bilinear(const ph_scene_desc& s, int tex, V2 uv) {
    const ph_texture_info& t = s.textures[tex];
    float x = uv.x * float(t.width) - 0.5f, y = uv.y * float(t.height) - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    if (!(fabsf(fx) < 1e8f) || !(fabsf(fy) < 1e8f)) return v4(0, 0, 0, 1);
    int x0 = int(fx), y0 = int(fy);
    float tx = x - fx, ty = y - fy;
    bool rep = t.wrap_repeat != 0;
    int xa = wrap_index(x0, t.width, rep), xb = wrap_index(x0 + 1, t.width, rep);
    int ya = wrap_index(y0, t.height, rep), yb = wrap_index(y0 + 1, t.height, rep);
    const uint32_t* base = s.texels + t.offset;
    bool srgb = t.srgb != 0;
    V4 a = texel(base[ya * t.width + xa], srgb, s.srgb_lut);
    V4 b = texel(base[ya * t.width + xb], srgb, s.srgb_lut);
    V4 c = texel(base[yb * t.width + xa], srgb, s.srgb_lut);
    V4 e = texel(base[yb * t.width + xb], srgb, s.srgb_lut);
    V4 top = a * (1 - tx) + b * tx;
    V4 bottom = c * (1 - tx) + e * tx;
    return top * (1 - ty) + bottom * ty;
}

inline V2 uv_at(const float* uvs, int i) { return {uvs[2 * i], uvs[2 * i + 1]}; }
inline V3 n_at(const float* n, int i) { return {n[3 * i], n[3 * i + 1], n[3 * i + 2]}; }
inline V4 t_at(const float* t, int i) { return {t[4 * i], t[4 * i + 1], t[4 * i + 2], t[4 * i + 3]}; }

// ---- traversal ----------------------------------------------------------------------------------------

struct Ray {
    V3 o, d;
    float tmax = kInf;
};

struct Hit {
    float t = kInf;
    int tri = -1;
    float u = 0, v = 0;
};

struct TraceCounters {
    int64_t nodes = 0, tris = 0, alpha = 0;
};

inline bool mt_triangle(const ph_tri_geom& g, V3 o, V3 d, float tmax, float& t, float& u, float& v) {
    V3 e1 = v3p(g.e1), e2 = v3p(g.e2);
    V3 p = cross(d, e2);
    float det = dot(e1, p);
    if (fabsf(det) < 1e-14f) return false;
    float inv = 1.0f / det;
    V3 s = o - v3p(g.v0);
    u = dot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(s, e1);
    v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, q) * inv;
    return t > 0.0f && t < tmax;
}

inline float slab(const float* bmin, const float* bmax, V3 o, V3 inv, float tmax) {
    float tx0 = (bmin[0] - o.x) * inv.x, tx1 = (bmax[0] - o.x) * inv.x;
    float ty0 = (bmin[1] - o.y) * inv.y, ty1 = (bmax[1] - o.y) * inv.y;
    float tz0 = (bmin[2] - o.z) * inv.z, tz1 = (bmax[2] - o.z) * inv.z;
    float tnear = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    float tfar = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    return tnear <= tfar ? tnear : kInf;
}

inline V3 guarded_reciprocal(V3 d) {
    auto r = [](float x) { return fabsf(x) > 1e-20f ? 1.0f / x : copysignf(1e20f, x); };
    return {r(d.x), r(d.y), r(d.z)};
}

inline bool alpha_cut(const ph_scene_desc& s, int tri, float u, float v) {
    const ph_material& m = s.materials[s.tris[tri].material];
    float alpha = m.base_color[3];
    if (m.base_color_tex >= 0 && s.uvs) {
        const int* idx = s.indices + 3 * tri;
        V2 uv = uv_at(s.uvs, idx[0]) * (1 - u - v) + uv_at(s.uvs, idx[1]) * u + uv_at(s.uvs, idx[2]) * v;
        alpha *= tex_bilinear(s, m.base_color_tex, uv).w;
    }
    return alpha < m.alpha_cutoff;
}

constexpr int kStackDepth = 64;

template <bool AnyHit>
inline bool walk(const ph_scene_desc& s, const Ray& ray, Hit& hit, TraceCounters* cnt) {
    V3 inv = guarded_reciprocal(ray.d);
    float tmax = ray.tmax;
    int stack[kStackDepth];
    int sp = 0, node = 0;
    bool found = false;
    for (;;) {
        const ph_bvh_node& n = s.nodes[node];
        if (cnt) cnt->nodes++;
        float tl = slab(n.lmin, n.lmax, ray.o, inv, tmax);
        float tr = slab(n.rmin, n.rmax, ray.o, inv, tmax);
        int next[2];
        float tn[2];
        int nn = 0;
        for (int side = 0; side < 2; side++) {
            float tc = side == 0 ? tl : tr;
            if (tc == kInf) continue;
            int count = side == 0 ? n.lcount : n.rcount;
            int idx = side == 0 ? n.left : n.right;
            if (count > 0) {
                for (int i = idx; i < idx + count; i++) {
                    float t, u, v;
                    if (cnt) cnt->tris++;
                    if (mt_triangle(s.tris[i], ray.o, ray.d, tmax, t, u, v)) {
                        if (s.tris[i].flags & PH_TRI_ALPHA_TESTED) {
                            if (cnt) cnt->alpha++;
                            if (alpha_cut(s, i, u, v)) continue;
                        }
                        if (AnyHit) {
                            hit.t = t;
                            hit.tri = i;
                            return true;
                        }
                        tmax = t;
                        hit.t = t;
                        hit.tri = i;
                        hit.u = u;
                        hit.v = v;
                        found = true;
                    }
                }
            } else if (idx >= 0) {
                next[nn] = idx;
                tn[nn] = tc;
                nn++;
            }
        }
        if (nn == 2) {
            if (tn[0] > tn[1]) {
                int a = next[0];
                next[0] = next[1];
                next[1] = a;
                float b = tn[0];
                tn[0] = tn[1];
                tn[1] = b;
            }
            if (tn[1] < tmax && sp < kStackDepth) stack[sp++] = next[1];
            if (tn[0] < tmax) {
                node = next[0];
                continue;
            }
        } else if (nn == 1 && tn[0] < tmax) {
            node = next[0];
            continue;
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    return found;
}

inline bool closest(const ph_scene_desc& s, const Ray& r, Hit& h, TraceCounters* c = nullptr) { return walk<false>(s, r, h, c); }
inline bool blocked(const ph_scene_desc& s, const Ray& r, TraceCounters* c = nullptr) {
    Hit h;
    return walk<true>(s, r, h, c);
}

// ---- BSDF ---------------------------------------------------------------------------------------------

struct MatParams {
    V3 base = v3s(0.8f);
    float metallic = 0.0f, roughness = 0.5f, transmission = 0.0f, ior = 1.5f;
};

struct Scatter {
    V3 wi, f;
    float pdf = 0;
    bool transmitted = false, valid = false;
};

inline float ndf_ggx(V3 h, float alpha) {
    float a2 = alpha * alpha;
    float c2 = h.z * h.z;
    float d = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * d * d);
}

inline float smith_lambda(V3 w, float alpha) {
    float c2 = w.z * w.z;
    if (c2 >= 1.0f) return 0.0f;
    float t2 = (1.0f - c2) / fmaxf(c2, 1e-12f);
    return 0.5f * (-1.0f + sqrtf(1.0f + alpha * alpha * t2));
}
inline float smith_g1(V3 w, float alpha) { return 1.0f / (1.0f + smith_lambda(w, alpha)); }
inline float smith_g2(V3 wo, V3 wi, float alpha) { return 1.0f / (1.0f + smith_lambda(wo, alpha) + smith_lambda(wi, alpha)); }

inline V3 visible_normal(V3 wo, float alpha, V2 u) {
    V3 vh = unit(v3(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    V3 t1 = lensq > 0 ? v3(-vh.y, vh.x, 0) * (1.0f / sqrtf(lensq)) : v3(1, 0, 0);
    V3 t2 = cross(vh, t1);
    float r = sqrtf(u.x);
    float phi = kTwoPi * u.y;
    float p1 = r * cosf(phi), p2 = r * sinf(phi);
    float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * ssqrt(1.0f - p1 * p1) + s * p2;
    V3 nh = t1 * p1 + t2 * p2 + vh * ssqrt(1.0f - p1 * p1 - p2 * p2);
    return unit(v3(alpha * nh.x, alpha * nh.y, fmaxf(1e-6f, nh.z)));
}

inline float visible_normal_pdf(V3 wo, V3 h, float alpha) {
    if (wo.z <= 0) return 0.0f;
    return smith_g1(wo, alpha) * fmaxf(0.0f, dot(wo, h)) * ndf_ggx(h, alpha) / wo.z;
}

inline V3 schlick(V3 f0, float c) {
    float m = sat(1.0f - c);
    float m5 = m * m * m * m * m;
    return f0 + (v3s(1.0f) - f0) * m5;
}

inline float fresnel_exact(float cos_i, float eta) {
    cos_i = clampf(cos_i, -1.0f, 1.0f);
    if (cos_i < 0) {
        eta = 1.0f / eta;
        cos_i = -cos_i;
    }
    float sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = ssqrt(1.0f - sin2_t);
    float rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (rp * rp + rs * rs);
}

inline float fresnel_mean(float eta) { return (eta - 1.0f) / (4.08567f + 1.00071f * eta); }

struct Mix {
    V3 base;
    float alpha, ior, w_metal, w_opaque, w_glass, f0, diffuse_norm;

    explicit Mix(const MatParams& p) {
        base = p.base;
        float r = clampf(p.roughness, 0.0f, 1.0f);
        alpha = fmaxf(r * r, 0.002f);
        ior = fmaxf(p.ior, 1.0001f);
        float m = sat(p.metallic), t = sat(p.transmission);
        w_metal = m;
        w_glass = (1 - m) * t;
        w_opaque = (1 - m) * (1 - t);
        f0 = sq((ior - 1) / (ior + 1));
        diffuse_norm = 1.0f / (1.0f - fresnel_mean(ior));
    }

    float spec_choice(float cos_o) const {
        if (f0 < 1e-6f) return 0.0f;
        float f = schlick(v3s(f0), cos_o).x;
        float d = (1 - f) * fmaxf(lum(base), 0.01f);
        return clampf(f / (f + d), 0.1f, 0.9f);
    }
    bool smooth() const { return alpha <= 0.002f; }
    bool light_samplable() const { return !smooth() || w_opaque > 0; }
    V3 albedo() const { return base * (w_metal + w_opaque) + v3s(w_glass); }

    V3 dielectric_eval(V3 wo, V3 wi, float* pdf_out) const {
        bool refl = wo.z * wi.z > 0;
        float etap = wo.z > 0 ? ior : 1.0f / ior;
        V3 wm = refl ? wo + wi : wi * etap + wo;
        if (len2(wm) < 1e-20f || wo.z == 0 || wi.z == 0) {
            *pdf_out = 0;
            return v3s(0);
        }
        wm = unit(wm);
        if (wm.z < 0) wm = -wm;
        if (dot(wm, wi) * wi.z < 0 || dot(wm, wo) * wo.z < 0) {
            *pdf_out = 0;
            return v3s(0);
        }
        float F = fresnel_exact(dot(wo, wm), ior);
        float R = F, T = 1 - F;
        V3 wo_up = wo.z > 0 ? wo : -wo;
        float pdf_m = visible_normal_pdf(wo_up, wm, alpha);
        float D = ndf_ggx(wm, alpha), G = smith_g2(wo, wi, alpha);
        if (refl) {
            *pdf_out = pdf_m / (4.0f * fabsf(dot(wo, wm))) * R;
            return v3s(D * G * F / (4.0f * fabsf(wi.z * wo.z)));
        }
        float denom = sq(dot(wi, wm) + dot(wo, wm) / etap);
        float dwm_dwi = fabsf(dot(wi, wm)) / denom;
        *pdf_out = pdf_m * dwm_dwi * T;
        float ft = T * D * G * fabsf(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom)) / (etap * etap);
        return base * ft;
    }

    bool dielectric_sample(V3 wo, V2 u, float u_lobe, V3& wi) const {
        V3 wo_up = wo.z > 0 ? wo : -wo;
        V3 wm = visible_normal(wo_up, alpha, u);
        float F = fresnel_exact(dot(wo, wm), ior);
        if (u_lobe < F) {
            wi = mirror(wo, wm);
            return wi.z * wo.z > 0;
        }
        float etap = wo.z > 0 ? ior : 1.0f / ior;
        float cos_i = dot(wo, wm);
        V3 n = wm;
        if (cos_i < 0) {
            n = -wm;
            cos_i = -cos_i;
        }
        float sin2_t = (1 - cos_i * cos_i) / (etap * etap);
        if (sin2_t >= 1) return false;
        float cos_t = ssqrt(1 - sin2_t);
        wi = -wo / etap + n * (cos_i / etap - cos_t);
        return wi.z * wo.z < 0;
    }

    V3 eval(V3 wo, V3 wi, float* pdf_out) const {
        V3 f = v3s(0);
        float pdf = 0;
        bool same = wo.z * wi.z > 0;
        if (same && wo.z > 0 && (w_metal > 0 || w_opaque > 0)) {
            V3 h = unit(wo + wi);
            float D = ndf_ggx(h, alpha), G = smith_g2(wo, wi, alpha);
            float spec = D * G / (4.0f * wo.z * wi.z);
            float pdf_spec = visible_normal_pdf(wo, h, alpha) / (4.0f * dot(wo, h));
            float voh = dot(wo, h);
            if (w_metal > 0) {
                f = f + schlick(base, voh) * (spec * w_metal);
                pdf += w_metal * pdf_spec;
            }
            if (w_opaque > 0) {
                bool matched = f0 < 1e-6f;
                float F = matched ? 0.0f : fresnel_exact(voh, ior);
                float base_layer = matched ? kInvPi : (1 - fresnel_exact(wo.z, ior)) * (1 - fresnel_exact(wi.z, ior)) * diffuse_norm * kInvPi;
                f = f + (v3s(F * spec) + base * base_layer) * w_opaque;
                float ps = spec_choice(wo.z);
                pdf += w_opaque * (ps * pdf_spec + (1 - ps) * hemi_cosine_pdf(wi.z));
            }
        }
        if (w_glass > 0) {
            float pg;
            f = f + dielectric_eval(wo, wi, &pg) * w_glass;
            pdf += w_glass * pg;
        }
        *pdf_out = pdf;
        return f;
    }

    Scatter sample(V3 wo, float u_lobe, V2 u, float u_sub) const {
        Scatter s;
        V3 wi;
        if (u_lobe < w_glass) {
            if (!dielectric_sample(wo, u, u_sub, wi)) return s;
        } else {
            if (wo.z <= 0) return s;
            float ul = (u_lobe - w_glass) / fmaxf(1 - w_glass, 1e-6f);
            bool spec = true;
            if (ul >= w_metal / fmaxf(w_metal + w_opaque, 1e-6f)) spec = u_sub < spec_choice(wo.z);
            if (spec) {
                V3 h = visible_normal(wo, alpha, u);
                wi = mirror(wo, h);
            } else {
                wi = hemi_cosine(u);
            }
            if (wi.z <= 0) return s;
        }
        s.wi = wi;
        s.f = eval(wo, wi, &s.pdf);
        s.transmitted = wi.z * wo.z < 0;
        s.valid = s.pdf > 0 && finite3(s.f);
        return s;
    }
};

// ---- lights -------------------------------------------------------------------------------------------

struct LightDraw {
    V3 wi;
    float dist = 0;
    V3 radiance;
    float pdf = 0;
    bool delta = false;
    int tri = -1;
    V2 bary;
};

inline int alias_pick(const ph_alias_entry* table, int n, float u, float* pmf) {
    float x = u * float(n);
    int i = int(x);
    if (i >= n) i = n - 1;
    float frac = x - float(i);
    int chosen = frac < table[i].prob ? i : table[i].alias;
    *pmf = table[chosen].pmf;
    return chosen;
}

inline V3 yaw(V3 d, float angle) {
    float c = cosf(angle), s = sinf(angle);
    return {c * d.x + s * d.z, d.y, -s * d.x + c * d.z};
}

inline V3 env_texel(const ph_scene_desc& s, int x, int y) { return n_at(s.env_texels, y * s.env_width + x); }

inline V3 sky_radiance(const ph_scene_desc& s, V3 dir) {
    if (!s.env_texels) return v3p(s.env_color) * s.env_intensity;
    V2 uv = dir_to_latlong(yaw(dir, -s.env_rotation));
    int x = int(uv.x * float(s.env_width));
    int y = int(uv.y * float(s.env_height));
    x = x < 0 ? 0 : (x >= s.env_width ? s.env_width - 1 : x);
    y = y < 0 ? 0 : (y >= s.env_height ? s.env_height - 1 : y);
    return env_texel(s, x, y) * s.env_intensity;
}

inline int cdf_search(const float* cdf, int n, float u) {
    int lo = 0, hi = n;
    while (lo + 1 < hi) {
        int mid = (lo + hi) / 2;
        if (cdf[mid] <= u)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

inline float sky_pdf_dir(const ph_scene_desc& s, V3 dir) {
    if (!s.env_texels) return 1.0f / (4.0f * kPi);
    V3 d = yaw(dir, -s.env_rotation);
    V2 uv = dir_to_latlong(d);
    float sin_theta = ssqrt(1 - d.y * d.y);
    if (sin_theta <= 0 || s.env_marginal_integral <= 0) return 0;
    int x = int(uv.x * float(s.env_width)), y = int(uv.y * float(s.env_height));
    x = x < 0 ? 0 : (x >= s.env_width ? s.env_width - 1 : x);
    y = y < 0 ? 0 : (y >= s.env_height ? s.env_height - 1 : y);
    float row_sin = sinf((float(y) + 0.5f) / float(s.env_height) * kPi);
    float f = lum(env_texel(s, x, y)) * row_sin;
    float pdf_uv = f / s.env_marginal_integral;
    return pdf_uv / (2.0f * kPi * kPi * sin_theta);
}

inline bool sky_sample(const ph_scene_desc& s, V2 u, V3& dir, float& pdf) {
    if (!s.env_texels) {
        dir = sphere_uniform(u);
        pdf = 1.0f / (4.0f * kPi);
        return true;
    }
    const int w = s.env_width, h = s.env_height;
    int y = cdf_search(s.env_marginal, h, u.y);
    float c0 = s.env_marginal[y], c1 = s.env_marginal[y + 1];
    float dv = c1 > c0 ? (u.y - c0) / (c1 - c0) : 0.5f;
    const float* row = s.env_conditional + size_t(y) * size_t(w + 1);
    int x = cdf_search(row, w, u.x);
    float r0 = row[x], r1 = row[x + 1];
    float du = r1 > r0 ? (u.x - r0) / (r1 - r0) : 0.5f;
    V2 uv = v2((float(x) + du) / float(w), (float(y) + dv) / float(h));
    V3 d = latlong_to_dir(uv);
    dir = yaw(d, s.env_rotation);
    float sin_theta = ssqrt(1 - d.y * d.y);
    if (sin_theta <= 1e-6f) return false;
    float row_sin = sinf((float(y) + 0.5f) / float(h) * kPi);
    float f = lum(env_texel(s, x, y)) * row_sin;
    pdf = f / s.env_marginal_integral / (2.0f * kPi * kPi * sin_theta);
    return pdf > 0;
}

inline V3 tri_emission(const ph_scene_desc& s, int tri, float u, float v) {
    const ph_material& m = s.materials[s.tris[tri].material];
    V3 e = v3p(m.emission);
    if (m.emissive_tex >= 0 && s.uvs) {
        const int* idx = s.indices + 3 * tri;
        V2 uv = uv_at(s.uvs, idx[0]) * (1 - u - v) + uv_at(s.uvs, idx[1]) * u + uv_at(s.uvs, idx[2]) * v;
        e = e * tex_bilinear(s, m.emissive_tex, uv).xyz();
    }
    return e;
}

inline bool faces_towards(const ph_scene_desc& s, int tri, V3 dir_from_light) {
    const ph_tri_geom& g = s.tris[tri];
    V3 n = cross(v3p(g.e1), v3p(g.e2));
    return s.materials[g.material].double_sided != 0 || dot(n, dir_from_light) > 0;
}

inline float emitter_pdf(const ph_scene_desc& s, int tri, V3 dir, float dist) {
    if (s.num_emissive == 0 || !s.tri_emissive) return 0;
    int k = s.tri_emissive[tri];
    if (k < 0) return 0;
    const ph_emissive_tri& e = s.emissive[k];
    const ph_tri_geom& g = s.tris[tri];
    V3 n = unit(cross(v3p(g.e1), v3p(g.e2)));
    float cos_l = fabsf(dot(n, dir));
    if (cos_l <= 0) return 0;
    return s.p_emissive * s.emissive_alias[k].pmf * dist * dist / (e.area * cos_l);
}

inline float sky_pdf(const ph_scene_desc& s, V3 dir) { return s.p_env * sky_pdf_dir(s, dir); }

inline LightDraw draw_light(const ph_scene_desc& s, V3 p, float u_group, V2 u, float u_light) {
    LightDraw ls;
    if (u_group < s.p_emissive && s.num_emissive > 0) {
        float pmf;
        int k = alias_pick(s.emissive_alias, s.num_emissive, u_light, &pmf);
        const ph_emissive_tri& e = s.emissive[k];
        const ph_tri_geom& g = s.tris[e.tri];
        V2 b = tri_uniform(u);
        V3 q = v3p(g.v0) + v3p(g.e1) * b.x + v3p(g.e2) * b.y;
        V3 d = q - p;
        float dist2 = len2(d);
        if (dist2 <= 0) return ls;
        float dist = sqrtf(dist2);
        V3 wi = d / dist;
        V3 n = unit(cross(v3p(g.e1), v3p(g.e2)));
        float cos_l = fabsf(dot(n, wi));
        if (cos_l <= 1e-8f || !faces_towards(s, e.tri, -wi)) return ls;
        ls.wi = wi;
        ls.dist = dist;
        ls.radiance = tri_emission(s, e.tri, b.x, b.y);
        ls.pdf = s.p_emissive * pmf * dist2 / (e.area * cos_l);
        ls.tri = e.tri;
        ls.bary = b;
        return ls;
    }
    if (u_group < s.p_emissive + s.p_punctual && s.num_punctual > 0) {
        int k = int(u_light * float(s.num_punctual));
        if (k >= s.num_punctual) k = s.num_punctual - 1;
        const ph_punctual_light& l = s.punctual[k];
        float pick = s.p_punctual / float(s.num_punctual);
        ls.delta = true;
        ls.pdf = pick;
        if (l.type == PH_LIGHT_DIRECTIONAL) {
            ls.wi = -v3p(l.direction);
            ls.dist = kInf;
            ls.radiance = v3p(l.intensity);
            return ls;
        }
        V3 d = v3p(l.position) - p;
        float dist2 = len2(d);
        if (dist2 <= 0) return LightDraw{};
        float dist = sqrtf(dist2);
        ls.wi = d / dist;
        ls.dist = dist;
        V3 li = v3p(l.intensity) / dist2;
        if (l.type == PH_LIGHT_SPOT) {
            float cd = dot(-ls.wi, v3p(l.direction));
            float t = sat((cd - l.cos_outer) / fmaxf(l.cos_inner - l.cos_outer, 1e-4f));
            li = li * (t * t);
        }
        ls.radiance = li;
        return ls;
    }
    if (s.p_env > 0) {
        V3 dir;
        float pdf;
        if (!sky_sample(s, u, dir, pdf)) return ls;
        ls.wi = dir;
        ls.dist = kInf;
        ls.radiance = sky_radiance(s, dir);
        ls.pdf = s.p_env * pdf;
    }
    return ls;
}

// ---- integrator ---------------------------------------------------------------------------------------

struct PathOpts {
    int max_depth = 8, rr_depth = 3;
    float clamp = 0.0f;
    int nee = 1;
    uint32_t* ray_count = nullptr;  // host-side instrumentation
};

struct Primary {
    V3 position;
    float depth = 0;
    V3 normal;
    int material = -1;
    V3 albedo = v3s(1);
    V3 emission;
};

inline V3 nudge(V3 p, V3 n) {
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    auto off = [&](float pc, float nc) {
        int of = int(int_scale * nc);
        float pi = bitsf(fbits(pc) + uint32_t(pc < 0 ? -of : of));
        return fabsf(pc) < origin ? pc + float_scale * nc : pi;
    };
    return {off(p.x, n.x), off(p.y, n.y), off(p.z, n.z)};
}

struct Surface {
    V3 p, ng, ns;
    V2 uv;
    Onb frame{v3(0, 0, 1)};
    MatParams params;
    V3 emission;
    bool front = false;
    int material = -1;
};

inline Surface shade_point(const ph_scene_desc& s, const Ray& ray, const Hit& hit) {
    Surface it;
    const ph_tri_geom& g = s.tris[hit.tri];
    const ph_material& m = s.materials[g.material];
    it.material = g.material;
    float w = 1 - hit.u - hit.v;
    it.p = ray.o + ray.d * hit.t;
    V3 ng = unit(cross(v3p(g.e1), v3p(g.e2)));
    const int* idx = s.indices + 3 * hit.tri;
    V3 ns = ng;
    if (s.normals) {
        V3 n = n_at(s.normals, idx[0]) * w + n_at(s.normals, idx[1]) * hit.u + n_at(s.normals, idx[2]) * hit.v;
        if (len2(n) > 1e-12f) ns = unit(n);
        if (dot(ns, ng) < 0) ns = -ns;
    }
    it.uv = s.uvs ? uv_at(s.uvs, idx[0]) * w + uv_at(s.uvs, idx[1]) * hit.u + uv_at(s.uvs, idx[2]) * hit.v : v2(0, 0);
    it.front = dot(ray.d, ng) < 0;
    if (m.normal_tex >= 0 && s.tangents) {
        V4 t4 = t_at(s.tangents, idx[0]) * w + t_at(s.tangents, idx[1]) * hit.u + t_at(s.tangents, idx[2]) * hit.v;
        V3 t = t4.xyz();
        if (len2(t) > 1e-12f) {
            t = unit(t - ns * dot(ns, t));
            V3 b = cross(ns, t) * (s.tangents[4 * idx[0] + 3] < 0 ? -1.0f : 1.0f);
            V4 tn = tex_bilinear(s, m.normal_tex, it.uv);
            V3 nt = v3(tn.x * 2 - 1, tn.y * 2 - 1, tn.z * 2 - 1);
            nt.x *= m.normal_scale;
            nt.y *= m.normal_scale;
            V3 n = unit(t * nt.x + b * nt.y + ns * nt.z);
            if (dot(n, ng) > 0.01f) ns = n;
        }
    }
    V4 base = v4(m.base_color[0], m.base_color[1], m.base_color[2], m.base_color[3]);
    if (m.base_color_tex >= 0) {
        V4 tc = tex_bilinear(s, m.base_color_tex, it.uv);
        base = v4(base.x * tc.x, base.y * tc.y, base.z * tc.z, base.w * tc.w);
    }
    float metallic = m.metallic, roughness = m.roughness;
    if (m.metallic_roughness_tex >= 0) {
        V4 mr = tex_bilinear(s, m.metallic_roughness_tex, it.uv);
        roughness *= mr.y;
        metallic *= mr.z;
    }
    it.params.base = base.xyz();
    it.params.metallic = metallic;
    it.params.roughness = roughness;
    it.params.transmission = m.transmission;
    it.params.ior = m.ior;
    it.emission = (it.front || m.double_sided) ? tri_emission(s, hit.tri, hit.u, hit.v) : v3s(0);
    if (m.transmission <= 0 && !it.front) {
        ng = -ng;
        ns = -ns;
    }
    V3 wo = -ray.d;
    if (m.transmission <= 0 && dot(wo, ns) <= 0) ns = ng;
    it.ng = ng;
    it.ns = ns;
    it.frame = Onb(ns);
    return it;
}

inline V3 clamp_contribution(V3 c, int depth, float clamp) {
    if (depth > 0 && clamp > 0) {
        float l = lum(c);
        if (l > clamp) c = c * (clamp / l);
    }
    return c;
}

inline V3 radiance(const ph_scene_desc& s, Ray ray, Pcg& rng, const PathOpts& opt, Primary* primary, TraceCounters* cnt) {
    V3 L = v3s(0), beta = v3s(1);
    float prev_pdf = 0;
    bool prev_delta = true;
    bool has_lights = s.p_env + s.p_emissive + s.p_punctual > 0;
    for (int depth = 0;; depth++) {
        Hit hit;
        if (opt.ray_count) ++*opt.ray_count;
        if (!closest(s, ray, hit, cnt)) {
            V3 le = sky_radiance(s, ray.d);
            float w = (prev_delta || s.p_env <= 0) ? 1.0f : mis_power(prev_pdf, sky_pdf(s, ray.d));
            L = L + clamp_contribution(beta * le * w, depth, opt.clamp);
            if (depth == 0 && primary) {
                primary->depth = 0;
                primary->material = -1;
                primary->emission = le;
                primary->albedo = v3s(1);
                primary->normal = -ray.d;
                primary->position = ray.o + ray.d * 1e4f;
            }
            break;
        }
        Surface it = shade_point(s, ray, hit);
        V3 wo_world = -ray.d;
        if (!black(it.emission)) {
            float w = 1.0f;
            if (!prev_delta) w = mis_power(prev_pdf, emitter_pdf(s, hit.tri, ray.d, hit.t));
            L = L + clamp_contribution(beta * it.emission * w, depth, opt.clamp);
        }
        Mix bsdf(it.params);
        if (depth == 0 && primary) {
            primary->position = it.p;
            primary->depth = hit.t;
            primary->normal = it.ns;
            primary->material = it.material;
            primary->albedo = bsdf.albedo();
            primary->emission = it.emission;
        }
        if (depth + 1 >= opt.max_depth) break;

        V3 wo = it.frame.local(wo_world);
        bool nee = opt.nee && has_lights && bsdf.light_samplable();
        if (nee) {
            float ug = rng.uni();
            V2 ul = rng.uni2();
            float uk = rng.uni();
            LightDraw ls = draw_light(s, it.p, ug, ul, uk);
            if (ls.pdf > 0 && !black(ls.radiance)) {
                V3 wi = it.frame.local(ls.wi);
                float pdf_b;
                V3 f = bsdf.eval(wo, wi, &pdf_b);
                if (!black(f)) {
                    Ray shadow;
                    shadow.o = nudge(it.p, dot(ls.wi, it.ng) > 0 ? it.ng : -it.ng);
                    shadow.d = ls.wi;
                    shadow.tmax = ls.dist == kInf ? kInf : ls.dist * (1.0f - 1e-4f);
                    if (opt.ray_count) ++*opt.ray_count;
                    if (!blocked(s, shadow, cnt)) {
                        float w = ls.delta ? 1.0f : mis_power(ls.pdf, pdf_b);
                        L = L + clamp_contribution(beta * f * fabsf(wi.z) * ls.radiance * (w / ls.pdf), depth, opt.clamp);
                    }
                }
            }
        }
        float u_lobe = rng.uni();
        V2 u = rng.uni2();
        float u_sub = rng.uni();
        Scatter bs = bsdf.sample(wo, u_lobe, u, u_sub);
        if (!bs.valid) break;
        beta = beta * (bs.f * (fabsf(bs.wi.z) / bs.pdf));
        if (!finite3(beta) || black(beta)) break;
        prev_pdf = bs.pdf;
        prev_delta = !nee;
        V3 wi_world = it.frame.world(bs.wi);
        float side = dot(wi_world, it.ng) * dot(wo_world, it.ng);
        if (bs.transmitted ? side >= 0 : side <= 0) break;
        ray.o = nudge(it.p, dot(wi_world, it.ng) > 0 ? it.ng : -it.ng);
        ray.d = unit(wi_world);
        ray.tmax = kInf;
        if (depth + 1 >= opt.rr_depth) {
            float q = fmaxf(0.05f, 1.0f - maxc(beta));
            if (rng.uni() < q) break;
            beta = beta * (1.0f / (1.0f - q));
        }
    }
    return L;
}

inline Ray eye_ray(const ph_camera_view& c, float px, float py, int width, int height, V2 lens) {
    float sx = (2.0f * px / float(width) - 1.0f) * c.tan_half_fov * c.aspect;
    float sy = (1.0f - 2.0f * py / float(height)) * c.tan_half_fov;
    V3 fwd = v3p(c.forward), right = v3p(c.right), up = v3p(c.up), pos = v3p(c.position);
    V3 d = unit(fwd + right * sx + up * sy);
    Ray r;
    r.o = pos;
    r.d = d;
    if (c.aperture > 0) {
        V3 focus = pos + d * (c.focus_distance / dot(d, fwd));
        V2 l = disk_concentric(lens) * c.aperture;
        r.o = pos + right * l.x + up * l.y;
        r.d = unit(focus - r.o);
    }
    return r;
}

inline Pcg pixel_rng(uint32_t pixel, uint32_t frame, uint32_t seed) {
    return Pcg(mix32(pixel ^ mix32(seed)) | (uint64_t(mix32(frame + 0x9e3779b9u)) << 32), pixel);
}

inline V3 pixel_sample(const ph_scene_desc& s, const ph_camera_view& cam, int x, int y, int width, int height, uint32_t frame, uint32_t seed,
                       const PathOpts& opt, Primary* primary, TraceCounters* cnt) {
    uint32_t pixel = uint32_t(y) * uint32_t(width) + uint32_t(x);
    Pcg rng = pixel_rng(pixel, frame, seed);
    V2 jitter = rng.uni2();
    V2 lens = rng.uni2();
    Ray ray = eye_ray(cam, float(x) + jitter.x, float(y) + jitter.y, width, height, lens);
    V3 L = radiance(s, ray, rng, opt, primary, cnt);
    if (!finite3(L) || L.x < 0 || L.y < 0 || L.z < 0) L = v3s(0);
    return L;
}

}  // namespace orc
