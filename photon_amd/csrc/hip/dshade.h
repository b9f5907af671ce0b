// Device shading: the GGX / rough-glass mixture BSDF, light sampling with its MIS densities, and the
// surface interaction at a hit. Each function evaluates the reference's expression in the reference's
// order (bsdf.h:45-281, lights.h:22-202, integrator.h:61-148) so that a path draws the same random
// numbers and takes the same branches as on the CPU.
#pragma once

#include "dscene.h"

namespace phd {

// ---- BSDF -------------------------------------------------------------------------------------------
struct MatParams {
    f3 base;
    float metallic, roughness, transmission, ior;
};

struct Scatter {
    f3 wi, f;
    float pdf;
    bool transmitted, valid;
};

PH_DEV float ndf_ggx(f3 h, float alpha) {
    float a2 = alpha * alpha;
    float c2 = h.z * h.z;
    float d = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * d * d);
}
PH_DEV float smith_lambda(f3 w, float alpha) {
    float c2 = w.z * w.z;
    if (c2 >= 1.0f) return 0.0f;
    float t2 = (1.0f - c2) / fmaxf(c2, 1e-12f);
    return 0.5f * (-1.0f + __builtin_sqrtf(1.0f + alpha * alpha * t2));
}
PH_DEV float smith_g1(f3 w, float alpha) { return 1.0f / (1.0f + smith_lambda(w, alpha)); }
PH_DEV float smith_g2(f3 wo, f3 wi, float alpha) { return 1.0f / (1.0f + smith_lambda(wo, alpha) + smith_lambda(wi, alpha)); }

// Visible-normal sampling (Heitz 2018).
PH_DEV f3 visible_normal(f3 wo, float alpha, f2 u) {
    f3 vh = unit(mk3(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    f3 t1 = lensq > 0 ? mk3(-vh.y, vh.x, 0) * (1.0f / __builtin_sqrtf(lensq)) : mk3(1, 0, 0);
    f3 t2 = cross(vh, t1);
    float r = __builtin_sqrtf(u.x);
    float phi = kTwoPi * u.y;
    float sp, cp;
    sincos_cr(phi, &sp, &cp);
    float p1 = r * cp, p2 = r * sp;
    float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * ssqrt(1.0f - p1 * p1) + s * p2;
    f3 nh = t1 * p1 + t2 * p2 + vh * ssqrt(1.0f - p1 * p1 - p2 * p2);
    return unit(mk3(alpha * nh.x, alpha * nh.y, fmaxf(1e-6f, nh.z)));
}
PH_DEV float visible_normal_pdf(f3 wo, f3 h, float alpha) {
    if (wo.z <= 0) return 0.0f;
    return smith_g1(wo, alpha) * fmaxf(0.0f, dot(wo, h)) * ndf_ggx(h, alpha) / wo.z;
}
PH_DEV f3 schlick(f3 f0, float c) {
    float m = sat(1.0f - c);
    float m5 = m * m * m * m * m;
    return f0 + (mk3(1.0f) - f0) * m5;
}
PH_DEV float fresnel_exact(float cos_i, float eta) {
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
PH_DEV float fresnel_mean(float eta) { return (eta - 1.0f) / (4.08567f + 1.00071f * eta); }

struct Mix {
    f3 base;
    float alpha, ior, w_metal, w_opaque, w_glass, f0, diffuse_norm;
};

PH_DEV Mix make_mix(const MatParams& p) {
    Mix b;
    b.base = p.base;
    float r = clampf(p.roughness, 0.0f, 1.0f);
    b.alpha = fmaxf(r * r, 0.002f);
    b.ior = fmaxf(p.ior, 1.0001f);
    float m = sat(p.metallic), t = sat(p.transmission);
    b.w_metal = m;
    b.w_glass = (1 - m) * t;
    b.w_opaque = (1 - m) * (1 - t);
    b.f0 = sq((b.ior - 1) / (b.ior + 1));
    b.diffuse_norm = 1.0f / (1.0f - fresnel_mean(b.ior));
    return b;
}
PH_DEV float spec_choice(const Mix& b, float cos_o) {
    if (b.f0 < 1e-6f) return 0.0f;
    float f = schlick(mk3(b.f0), cos_o).x;
    float d = (1 - f) * fmaxf(lum(b.base), 0.01f);
    return clampf(f / (f + d), 0.1f, 0.9f);
}
PH_DEV bool light_samplable(const Mix& b) { return !(b.alpha <= 0.002f) || b.w_opaque > 0; }
PH_DEV f3 mix_albedo(const Mix& b) { return b.base * (b.w_metal + b.w_opaque) + mk3(b.w_glass); }

PH_DEV f3 dielectric_eval(const Mix& b, f3 wo, f3 wi, float* pdf_out) {
    bool refl = wo.z * wi.z > 0;
    float etap = wo.z > 0 ? b.ior : 1.0f / b.ior;
    f3 wm = refl ? wo + wi : wi * etap + wo;
    if (len2(wm) < 1e-20f || wo.z == 0 || wi.z == 0) {
        *pdf_out = 0;
        return mk3(0);
    }
    wm = unit(wm);
    if (wm.z < 0) wm = -wm;
    if (dot(wm, wi) * wi.z < 0 || dot(wm, wo) * wo.z < 0) {
        *pdf_out = 0;
        return mk3(0);
    }
    float F = fresnel_exact(dot(wo, wm), b.ior);
    float R = F, T = 1 - F;
    f3 wo_up = wo.z > 0 ? wo : -wo;
    float pdf_m = visible_normal_pdf(wo_up, wm, b.alpha);
    float D = ndf_ggx(wm, b.alpha), G = smith_g2(wo, wi, b.alpha);
    if (refl) {
        *pdf_out = pdf_m / (4.0f * fabsf(dot(wo, wm))) * R;
        return mk3(D * G * F / (4.0f * fabsf(wi.z * wo.z)));
    }
    float denom = sq(dot(wi, wm) + dot(wo, wm) / etap);
    float dwm_dwi = fabsf(dot(wi, wm)) / denom;
    *pdf_out = pdf_m * dwm_dwi * T;
    float ft = T * D * G * fabsf(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom)) / (etap * etap);
    return b.base * ft;
}

PH_DEV bool dielectric_sample(const Mix& b, f3 wo, f2 u, float u_lobe, f3& wi) {
    f3 wo_up = wo.z > 0 ? wo : -wo;
    f3 wm = visible_normal(wo_up, b.alpha, u);
    float F = fresnel_exact(dot(wo, wm), b.ior);
    if (u_lobe < F) {
        wi = mirror(wo, wm);
        return wi.z * wo.z > 0;
    }
    float etap = wo.z > 0 ? b.ior : 1.0f / b.ior;
    float cos_i = dot(wo, wm);
    f3 n = wm;
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

// The whole mixture and its exact density. `Lobes` is a compile-time hint of which components can
// have non-zero weight (bit 0 metal, 1 opaque, 2 glass): a component whose weight is zero contributes
// nothing in the reference either, so specialised instances compute identical results.
template <int Lobes = 7>
PH_DEV f3 mix_eval(const Mix& b, f3 wo, f3 wi, float* pdf_out) {
    f3 f = mk3(0);
    float pdf = 0;
    bool same = wo.z * wi.z > 0;
    if ((Lobes & 3) && same && wo.z > 0 && (b.w_metal > 0 || b.w_opaque > 0)) {
        f3 h = unit(wo + wi);
        float D = ndf_ggx(h, b.alpha), G = smith_g2(wo, wi, b.alpha);
        float spec = D * G / (4.0f * wo.z * wi.z);
        float pdf_spec = visible_normal_pdf(wo, h, b.alpha) / (4.0f * dot(wo, h));
        float voh = dot(wo, h);
        if ((Lobes & 1) && b.w_metal > 0) {
            f = f + schlick(b.base, voh) * (spec * b.w_metal);
            pdf += b.w_metal * pdf_spec;
        }
        if ((Lobes & 2) && b.w_opaque > 0) {
            bool matched = b.f0 < 1e-6f;
            float F = matched ? 0.0f : fresnel_exact(voh, b.ior);
            float base_layer = matched ? kInvPi : (1 - fresnel_exact(wo.z, b.ior)) * (1 - fresnel_exact(wi.z, b.ior)) * b.diffuse_norm * kInvPi;
            f = f + (mk3(F * spec) + b.base * base_layer) * b.w_opaque;
            float ps = spec_choice(b, wo.z);
            pdf += b.w_opaque * (ps * pdf_spec + (1 - ps) * hemi_cosine_pdf(wi.z));
        }
    }
    if ((Lobes & 4) && b.w_glass > 0) {
        float pg;
        f = f + dielectric_eval(b, wo, wi, &pg) * b.w_glass;
        pdf += b.w_glass * pg;
    }
    *pdf_out = pdf;
    return f;
}

template <int Lobes = 7>
PH_DEV Scatter mix_sample(const Mix& b, f3 wo, float u_lobe, f2 u, float u_sub) {
    Scatter s;
    s.pdf = 0;
    s.transmitted = false;
    s.valid = false;
    s.wi = mk3(0);
    s.f = mk3(0);
    f3 wi;
    if ((Lobes & 4) && u_lobe < b.w_glass) {
        if (!dielectric_sample(b, wo, u, u_sub, wi)) return s;
    } else {
        if (wo.z <= 0) return s;
        float ul = (u_lobe - b.w_glass) / fmaxf(1 - b.w_glass, 1e-6f);
        bool spec = true;
        if (ul >= b.w_metal / fmaxf(b.w_metal + b.w_opaque, 1e-6f)) spec = u_sub < spec_choice(b, wo.z);
        if (spec) {
            f3 h = visible_normal(wo, b.alpha, u);
            wi = mirror(wo, h);
        } else {
            wi = hemi_cosine(u);
        }
        if (wi.z <= 0) return s;
    }
    s.wi = wi;
    s.f = mix_eval<Lobes>(b, wo, wi, &s.pdf);
    s.transmitted = wi.z * wo.z < 0;
    s.valid = s.pdf > 0 && finite3(s.f);
    return s;
}

// ---- lights -------------------------------------------------------------------------------------------
struct LightDraw {
    f3 wi;
    float dist;
    f3 radiance;
    float pdf;
    bool delta;
};

PH_DEV int alias_pick(const ph_alias_entry* table, int n, float u, float* pmf) {
    float x = u * float(n);
    int i = int(x);
    if (i >= n) i = n - 1;
    float frac = x - float(i);
    const ph_alias_entry e = ldg(table + i);
    int chosen = frac < e.prob ? i : e.alias;
    *pmf = ldg(&table[chosen].pmf);
    return chosen;
}

PH_DEV f3 yaw(f3 d, float angle) {
    float s, c;
    sincos_cr(angle, &s, &c);
    return {c * d.x + s * d.z, d.y, -s * d.x + c * d.z};
}

PH_DEV f3 env_texel(const DScene& s, int x, int y) { return n_at(s.env_texels, y * s.env_w + x); }

PH_DEV f3 sky_radiance(const DScene& s, f3 dir) {
    if (!s.env_texels) return ld3(s.env_color) * s.env_intensity;
    f2 uv = dir_to_latlong(yaw(dir, -s.env_rotation));
    int x = int(uv.x * float(s.env_w));
    int y = int(uv.y * float(s.env_h));
    x = x < 0 ? 0 : (x >= s.env_w ? s.env_w - 1 : x);
    y = y < 0 ? 0 : (y >= s.env_h ? s.env_h - 1 : y);
    return env_texel(s, x, y) * s.env_intensity;
}

PH_DEV int cdf_search(const float* cdf, int n, float u) {
    int lo = 0, hi = n;
    while (lo + 1 < hi) {
        int mid = (lo + hi) / 2;
        if (ldg(cdf + mid) <= u)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// The same index as cdf_search, starting from a guide table (see DScene): with m a power of two, u m is
// exact, so for k = floor(u m) the answer lies in [guide[k], guide[k + 1]] and the bisection runs over
// that range only (a few entries instead of log2(n) dependent loads).
PH_DEV int cdf_search_guided(const float* cdf, int n, const int* guide, int m, float u) {
    const int k = min(int(u * float(m)), m - 1);
    int lo = ldg(guide + k), hi = min(ldg(guide + k + 1) + 1, n);
    while (lo + 1 < hi) {
        int mid = (lo + hi) / 2;
        if (ldg(cdf + mid) <= u)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

PH_DEV float sky_pdf_dir(const DScene& s, f3 dir) {
    if (!s.env_texels) return 1.0f / (4.0f * kPi);
    f3 d = yaw(dir, -s.env_rotation);
    f2 uv = dir_to_latlong(d);
    float sin_theta = ssqrt(1 - d.y * d.y);
    if (sin_theta <= 0 || s.env_integral <= 0) return 0;
    int x = int(uv.x * float(s.env_w)), y = int(uv.y * float(s.env_h));
    x = x < 0 ? 0 : (x >= s.env_w ? s.env_w - 1 : x);
    y = y < 0 ? 0 : (y >= s.env_h ? s.env_h - 1 : y);
    float row_sin = sin_cr((float(y) + 0.5f) / float(s.env_h) * kPi);
    float f = lum(env_texel(s, x, y)) * row_sin;
    float pdf_uv = f / s.env_integral;
    return pdf_uv / (2.0f * kPi * kPi * sin_theta);
}

PH_DEV bool sky_sample(const DScene& s, f2 u, f3& dir, float& pdf) {
    if (!s.env_texels) {
        dir = sphere_uniform(u);
        pdf = 1.0f / (4.0f * kPi);
        return true;
    }
    const int w = s.env_w, h = s.env_h;
    const bool guided = s.env_guide_marginal != nullptr;
    int y = guided ? cdf_search_guided(s.env_marginal, h, s.env_guide_marginal, s.env_guide_mh, u.y) : cdf_search(s.env_marginal, h, u.y);
    float c0 = ldg(s.env_marginal + y), c1 = ldg(s.env_marginal + y + 1);
    float dv = c1 > c0 ? (u.y - c0) / (c1 - c0) : 0.5f;
    const float* row = s.env_conditional + size_t(y) * size_t(w + 1);
    int x = guided ? cdf_search_guided(row, w, s.env_guide_conditional + size_t(y) * size_t(s.env_guide_mw + 1), s.env_guide_mw, u.x)
                   : cdf_search(row, w, u.x);
    float r0 = ldg(row + x), r1 = ldg(row + x + 1);
    float du = r1 > r0 ? (u.x - r0) / (r1 - r0) : 0.5f;
    f2 uv = mk2((float(x) + du) / float(w), (float(y) + dv) / float(h));
    f3 d = latlong_to_dir(uv);
    dir = yaw(d, s.env_rotation);
    float sin_theta = ssqrt(1 - d.y * d.y);
    if (sin_theta <= 1e-6f) return false;
    float row_sin = sin_cr((float(y) + 0.5f) / float(h) * kPi);
    float f = lum(env_texel(s, x, y)) * row_sin;
    pdf = f / s.env_integral / (2.0f * kPi * kPi * sin_theta);
    return pdf > 0;
}

PH_DEV f3 tri_emission(const DScene& s, int tri, const Mat& m, float u, float v) {
    f3 e = m.emission;
    if (m.emissive_tex >= 0 && s.uvs) {
        const int* idx = s.indices + 3 * tri;
        f2 uv = uv_at(s.uvs, index_at(idx, 0)) * (1 - u - v) + uv_at(s.uvs, index_at(idx, 1)) * u + uv_at(s.uvs, index_at(idx, 2)) * v;
        f4 t = tex_bilinear(s, m.emissive_tex, uv);
        e = e * mk3(t.x, t.y, t.z);
    }
    return e;
}

PH_DEV float emitter_pdf(const DScene& s, int tri, f3 dir, float dist) {
    if (s.num_emissive == 0 || !s.tri_emissive) return 0;
    int k = ldg(s.tri_emissive + tri);
    if (k < 0) return 0;
    const ph_emissive_tri e = ldg(s.emissive + k);
    Tri g = load_tri(s, tri);
    f3 n = unit(cross(g.e1, g.e2));
    float cos_l = fabsf(dot(n, dir));
    if (cos_l <= 0) return 0;
    return s.p_emissive * ldg(&s.alias[k].pmf) * dist * dist / (e.area * cos_l);
}

PH_DEV float sky_pdf(const DScene& s, f3 dir) { return s.p_env * sky_pdf_dir(s, dir); }

PH_DEV LightDraw draw_light(const DScene& s, f3 p, float u_group, f2 u, float u_light) {
    LightDraw ls;
    ls.wi = mk3(0);
    ls.dist = 0;
    ls.radiance = mk3(0);
    ls.pdf = 0;
    ls.delta = false;
    if (u_group < s.p_emissive && s.num_emissive > 0) {
        float pmf;
        int k = alias_pick(s.alias, s.num_emissive, u_light, &pmf);
        const ph_emissive_tri e = ldg(s.emissive + k);
        Tri g = load_tri(s, e.tri);
        f2 b = tri_uniform(u);
        f3 q = g.v0 + g.e1 * b.x + g.e2 * b.y;
        f3 d = q - p;
        float dist2 = len2(d);
        if (dist2 <= 0) return ls;
        float dist = __builtin_sqrtf(dist2);
        f3 wi = d / dist;
        f3 n = unit(cross(g.e1, g.e2));
        float cos_l = fabsf(dot(n, wi));
        if (cos_l <= 1e-8f) return ls;
        Mat m = load_material(s, g.material);
        f3 nn = cross(g.e1, g.e2);
        if (!(m.double_sided != 0 || dot(nn, -wi) > 0)) return ls;
        ls.wi = wi;
        ls.dist = dist;
        ls.radiance = tri_emission(s, e.tri, m, b.x, b.y);
        ls.pdf = s.p_emissive * pmf * dist2 / (e.area * cos_l);
        return ls;
    }
    if (u_group < s.p_emissive + s.p_punctual && s.num_punctual > 0) {
        int k = int(u_light * float(s.num_punctual));
        if (k >= s.num_punctual) k = s.num_punctual - 1;
        const ph_punctual_light l = ldg(s.punctual + k);
        float pick = s.p_punctual / float(s.num_punctual);
        ls.delta = true;
        ls.pdf = pick;
        if (l.type == PH_LIGHT_DIRECTIONAL) {
            ls.wi = -ld3(l.direction);
            ls.dist = kInf;
            ls.radiance = ld3(l.intensity);
            return ls;
        }
        f3 d = ld3(l.position) - p;
        float dist2 = len2(d);
        if (dist2 <= 0) {
            ls.pdf = 0;
            ls.delta = false;
            return ls;
        }
        float dist = __builtin_sqrtf(dist2);
        ls.wi = d / dist;
        ls.dist = dist;
        f3 li = ld3(l.intensity) / dist2;
        if (l.type == PH_LIGHT_SPOT) {
            float cd = dot(-ls.wi, ld3(l.direction));
            float t = sat((cd - l.cos_outer) / fmaxf(l.cos_inner - l.cos_outer, 1e-4f));
            li = li * (t * t);
        }
        ls.radiance = li;
        return ls;
    }
    if (s.p_env > 0) {
        f3 dir;
        float pdf;
        if (!sky_sample(s, u, dir, pdf)) return ls;
        ls.wi = dir;
        ls.dist = kInf;
        ls.radiance = sky_radiance(s, dir);
        ls.pdf = s.p_env * pdf;
    }
    return ls;
}

// ---- surface interaction ------------------------------------------------------------------------------
// Wachter-Binder origin offset (integrator.h:61-69).
PH_DEV f3 nudge(f3 p, f3 n) {
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    auto off = [&](float pc, float nc) {
        int of = int(int_scale * nc);
        float pi = __uint_as_float(__float_as_uint(pc) + uint32_t(pc < 0 ? -of : of));
        return fabsf(pc) < origin ? pc + float_scale * nc : pi;
    };
    return {off(p.x, n.x), off(p.y, n.y), off(p.z, n.z)};
}

struct Surface {
    f3 p, ng, ns;
    Onb frame;
    MatParams params;
    f3 emission;
    int material;
};

PH_DEV Surface shade_point(const DScene& s, f3 ro, f3 rd, float t, int tri, float hu, float hv) {
    Surface it;
    Tri g = load_tri(s, tri);
    Mat m = load_material(s, g.material);
    it.material = g.material;
    float w = 1 - hu - hv;
    it.p = ro + rd * t;
    f3 ng = unit(cross(g.e1, g.e2));
    const int* idx = s.indices + 3 * tri;
    const int i0 = index_at(idx, 0), i1 = index_at(idx, 1), i2 = index_at(idx, 2);
    f3 ns = ng;
    if (s.normals) {
        f3 n = n_at(s.normals, i0) * w + n_at(s.normals, i1) * hu + n_at(s.normals, i2) * hv;
        if (len2(n) > 1e-12f) ns = unit(n);
        if (dot(ns, ng) < 0) ns = -ns;
    }
    f2 uv = s.uvs ? uv_at(s.uvs, i0) * w + uv_at(s.uvs, i1) * hu + uv_at(s.uvs, i2) * hv : mk2(0, 0);
    const bool front = dot(rd, ng) < 0;
    if (m.normal_tex >= 0 && s.tangents) {
        f4 t4 = t_at(s.tangents, i0) * w + t_at(s.tangents, i1) * hu + t_at(s.tangents, i2) * hv;
        f3 tt = mk3(t4.x, t4.y, t4.z);
        if (len2(tt) > 1e-12f) {
            tt = unit(tt - ns * dot(ns, tt));
            f3 bt = cross(ns, tt) * (ldg(s.tangents + 4 * i0 + 3) < 0 ? -1.0f : 1.0f);
            f4 tn = tex_bilinear(s, m.normal_tex, uv);
            f3 nt = mk3(tn.x * 2 - 1, tn.y * 2 - 1, tn.z * 2 - 1);
            nt.x *= m.normal_scale;
            nt.y *= m.normal_scale;
            f3 n = unit(tt * nt.x + bt * nt.y + ns * nt.z);
            if (dot(n, ng) > 0.01f) ns = n;
        }
    }
    float4 base = m.base_color;
    if (m.base_color_tex >= 0) {
        f4 tc = tex_bilinear(s, m.base_color_tex, uv);
        base = make_float4(base.x * tc.x, base.y * tc.y, base.z * tc.z, base.w * tc.w);
    }
    float metallic = m.metallic, roughness = m.roughness;
    if (m.mr_tex >= 0) {
        f4 mr = tex_bilinear(s, m.mr_tex, uv);
        roughness *= mr.y;
        metallic *= mr.z;
    }
    it.params.base = mk3(base.x, base.y, base.z);
    it.params.metallic = metallic;
    it.params.roughness = roughness;
    it.params.transmission = m.transmission;
    it.params.ior = m.ior;
    it.emission = (front || m.double_sided) ? tri_emission(s, tri, m, hu, hv) : mk3(0);
    if (m.transmission <= 0 && !front) {
        ng = -ng;
        ns = -ns;
    }
    f3 wo = -rd;
    if (m.transmission <= 0 && dot(wo, ns) <= 0) ns = ng;
    it.ng = ng;
    it.ns = ns;
    it.frame = onb(ns);
    return it;
}

PH_DEV f3 clamp_contribution(f3 c, int depth, float clamp) {
    if (depth > 0 && clamp > 0) {
        float l = lum(c);
        if (l > clamp) c = c * (clamp / l);
    }
    return c;
}

// A camera ray through (px, py) with a thin lens when the aperture is open (integrator.h:266-280).
PH_DEV void eye_ray(const ph_camera_view& c, float px, float py, int width, int height, f2 lens, f3& o, f3& d) {
    float sx = (2.0f * px / float(width) - 1.0f) * c.tan_half_fov * c.aspect;
    float sy = (1.0f - 2.0f * py / float(height)) * c.tan_half_fov;
    f3 fwd = ld3(c.forward), right = ld3(c.right), up = ld3(c.up), pos = ld3(c.position);
    d = unit(fwd + right * sx + up * sy);
    o = pos;
    if (c.aperture > 0) {
        f3 focus = pos + d * (c.focus_distance / dot(d, fwd));
        f2 l = disk_concentric(lens) * c.aperture;
        o = pos + right * l.x + up * l.y;
        d = unit(focus - o);
    }
}

}  // namespace phd
