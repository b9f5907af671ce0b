// Producers of the compiled scene's sampling inputs. Their outputs are compared bit for bit with what
// the reference's host code computes from the same scene, so each keeps that code's float and double
// evaluation order (the code around the arithmetic is this library's own):
//   fill_tangents    <- src/scene/scene.cpp:50-80  (Lengyel's per-vertex tangents)
//   alias_table      <- src/scene/scene.cpp:82-111 (Vose's alias method in double)
//   light_tables     <- src/scene/scene.cpp:113-176 (emitted power, sample split, 5 % floor)
//   backdrop_tables  <- src/scene/scene.cpp:180-210 (luminance * sin(theta) CDFs)
#include <algorithm>
#include <cmath>

#include "scene.hpp"

namespace phh {

void fill_tangents(Mesh& m, const std::vector<ph_material>& materials) {
    if (m.uv.empty() || !m.tan.empty()) return;
    if (std::none_of(materials.begin(), materials.end(), [](const ph_material& x) { return x.normal_tex >= 0; })) return;
    const size_t nv = m.pos.size();
    // Per vertex: the sums of the adjacent triangles' d(position)/du and d(position)/dv.
    std::vector<vec3> du(nv, vec3(0)), dv(nv, vec3(0));
    for (size_t k = 0; k + 2 < m.corner.size(); k += 3) {
        const size_t v[3] = {size_t(m.corner[k]), size_t(m.corner[k + 1]), size_t(m.corner[k + 2])};
        const vec3 e1 = m.pos[v[1]] - m.pos[v[0]], e2 = m.pos[v[2]] - m.pos[v[0]];
        const vec2 t1 = m.uv[v[1]] - m.uv[v[0]], t2 = m.uv[v[2]] - m.uv[v[0]];
        const float det = t1.x * t2.y - t2.x * t1.y;
        if (std::fabs(det) < 1e-20f) continue;  // degenerate parametrisation
        const float inv = 1.0f / det;
        const vec3 along_u = (e1 * t2.y - e2 * t1.y) * inv;
        const vec3 along_v = (e2 * t1.x - e1 * t2.x) * inv;
        for (const size_t c : v) {
            du[c] += along_u;
            dv[c] += along_v;
        }
    }
    m.tan.resize(nv);
    for (size_t i = 0; i < nv; i++) {
        const vec3 n = m.nrm.empty() ? vec3(0, 0, 1) : m.nrm[i];
        vec3 t = du[i] - n * dot(n, du[i]);  // Gram-Schmidt against the normal
        if (length_squared(t) < 1e-20f) t = basis_tangent(normalize(n));
        t = normalize(t);
        const float handed = dot(cross(n, t), dv[i]) < 0 ? -1.0f : 1.0f;
        m.tan[i] = vec4(t, handed);
    }
}

std::vector<ph_alias_entry> alias_table(const std::vector<float>& weights) {
    const size_t n = weights.size();
    std::vector<ph_alias_entry> table(n);
    if (n == 0) return table;
    double total = 0;
    for (const float w : weights) total += std::max(0.0f, w);
    // Scaled probabilities n * p_i; bins below 1 borrow from bins above 1. Both work lists are stacks
    // filled in index order and consumed from the top.
    std::vector<double> share(n);
    std::vector<size_t> low(n), high(n);
    size_t n_low = 0, n_high = 0;
    for (size_t i = 0; i < n; i++) {
        const double p = total > 0 ? std::max(0.0f, weights[i]) / total : 1.0 / double(n);
        table[i] = ph_alias_entry{0.0f, int(i), float(p), 0};
        share[i] = p * double(n);
        if (share[i] < 1.0)
            low[n_low++] = i;
        else
            high[n_high++] = i;
    }
    while (n_low > 0 && n_high > 0) {
        const size_t lo = low[--n_low], hi = high[n_high - 1];
        table[lo].prob = float(share[lo]);
        table[lo].alias = int(hi);
        share[hi] = share[hi] + share[lo] - 1.0;
        if (share[hi] < 1.0) {
            n_high--;
            low[n_low++] = hi;
        }
    }
    // What is left holds its own bin with certainty (rounding leaves some on either list).
    for (size_t k = 0; k < n_low; k++) table[low[k]].prob = 1.0f;
    for (size_t k = 0; k < n_high; k++) table[high[k]].prob = 1.0f;
    return table;
}

namespace {

// Bilinear RGBA lookup of a pooled texture (the reference's sample_texture, src/render/texture.h:9-44),
// used to weigh textured emitters.
class PoolSampler {
public:
    explicit PoolSampler(const Compiled& c) : c_(c) {}

    vec4 at(int tex, vec2 uv) const {
        const ph_texture_info& t = c_.texture_table[size_t(tex)];
        const float x = uv.x * float(t.width) - 0.5f, y = uv.y * float(t.height) - 0.5f;
        const float fx = std::floor(x), fy = std::floor(y);
        if (!(std::fabs(fx) < 1e8f) || !(std::fabs(fy) < 1e8f)) return vec4(0, 0, 0, 1);
        const float wx = x - fx, wy = y - fy;
        const bool tile = t.wrap_repeat != 0;
        const int x0 = wrap(int(fx), t.width, tile), x1 = wrap(int(fx) + 1, t.width, tile);
        const int y0 = wrap(int(fy), t.height, tile), y1 = wrap(int(fy) + 1, t.height, tile);
        const uint32_t* img = c_.texel_pool.data() + t.offset;
        auto px = [&](int xx, int yy) { return decode(img[yy * t.width + xx], t.srgb != 0); };
        const vec4 top = px(x0, y0) * (1 - wx) + px(x1, y0) * wx;
        const vec4 bottom = px(x0, y1) * (1 - wx) + px(x1, y1) * wx;
        return top * (1 - wy) + bottom * wy;
    }

private:
    const Compiled& c_;

    static int wrap(int i, int n, bool tile) {
        if (!tile) return std::clamp(i, 0, n - 1);
        const int r = i % n;
        return r < 0 ? r + n : r;
    }

    vec4 decode(uint32_t p, bool srgb) const {
        const uint32_t r = p & 0xFF, g = (p >> 8) & 0xFF, b = (p >> 16) & 0xFF, a = p >> 24;
        if (srgb) return vec4(c_.srgb_lut[r], c_.srgb_lut[g], c_.srgb_lut[b], float(a) * (1.0f / 255.0f));
        return vec4(float(r), float(g), float(b), float(a)) * (1.0f / 255.0f);
    }
};

// Barycentric points at which a textured emitter's brightness is probed.
constexpr float kProbe[7][2] = {{1.f / 3, 1.f / 3}, {0.1f, 0.1f}, {0.8f, 0.1f}, {0.1f, 0.8f}, {0.45f, 0.1f}, {0.45f, 0.45f}, {0.1f, 0.45f}};

// Mean luminance of an emitter over the probe points (emission factor times its texture).
float textured_luminance(const Scene& s, const PoolSampler& tex, size_t tri, vec3 factor, int image) {
    const int* c = &s.out.corner[3 * tri];
    const vec2 a = s.mesh.uv[size_t(c[0])], b = s.mesh.uv[size_t(c[1])], d = s.mesh.uv[size_t(c[2])];
    float sum = 0;
    for (const auto& q : kProbe) sum += luminance(factor * tex.at(image, a * (1 - q[0] - q[1]) + b * q[0] + d * q[1]).xyz());
    return sum / 7.0f;
}

}  // namespace

LightTables light_tables(const Scene& s) {
    const Compiled& c = s.out;
    LightTables L;
    L.slot_of_tri.assign(c.tri.size(), -1);
    std::vector<float> power;
    const PoolSampler tex(c);
    for (size_t i = 0; i < c.tri.size(); i++) {
        const ph_material& m = s.materials[size_t(c.tri[i].material)];
        const vec3 e(m.emission[0], m.emission[1], m.emission[2]);
        float lum = luminance(e);
        if (lum <= 0) continue;
        if (m.emissive_tex >= 0 && !s.mesh.uv.empty()) {
            lum = textured_luminance(s, tex, i, e, m.emissive_tex);
            if (lum <= 1e-4f) continue;  // dark under its texture: left to BSDF sampling
        }
        const vec3 e1(c.tri[i].e1[0], c.tri[i].e1[1], c.tri[i].e1[2]), e2(c.tri[i].e2[0], c.tri[i].e2[1], c.tri[i].e2[2]);
        const float area = 0.5f * length(cross(e1, e2));
        if (area <= 0) continue;
        L.slot_of_tri[i] = int(L.emitters.size());
        L.emitters.push_back(ph_emissive_tri{int(i), area});
        const float watts = area * lum * kPi * (m.double_sided ? 2.0f : 1.0f);
        power.push_back(watts);
        L.emitted_power += watts;
    }
    L.alias = alias_table(power);

    // Power estimates of the other two kinds, over the scene's bounding sphere.
    const Bounds& box = c.tree.bounds;
    const float r = box.empty() ? 1.0f : 0.5f * length(box.high - box.low);
    float sky = 0;
    if (!s.backdrop.rgb.empty()) {
        double sum = 0;
        for (const vec3& t : s.backdrop.rgb) sum += luminance(t);
        sky = float(sum / double(s.backdrop.rgb.size())) * s.backdrop.gain * kPi * kPi * r * r * 4.0f;
    } else {
        sky = luminance(s.backdrop.flat) * s.backdrop.gain * kPi * kPi * r * r * 4.0f;
    }
    float lamps = 0;
    for (const ph_punctual_light& l : s.lamps) {
        const float y = luminance(vec3(l.intensity[0], l.intensity[1], l.intensity[2]));
        lamps += l.type == PH_LIGHT_DIRECTIONAL ? y * kPi * r * r : y * 4.0f * kPi;
    }
    const float total = sky + L.emitted_power + lamps;
    if (total > 0) {
        // Proportional to power, with a 5 % floor for every kind present (the estimates ignore occlusion),
        // renormalised.
        float p[3] = {sky / total, L.emitted_power / total, lamps / total};
        const float present[3] = {sky, L.emitted_power, lamps};
        for (int k = 0; k < 3; k++)
            if (present[k] > 0) p[k] = std::max(p[k], 0.05f);
        const float norm = p[0] + p[1] + p[2];
        L.p_backdrop = p[0] / norm;
        L.p_emitters = p[1] / norm;
        L.p_punctual = p[2] / norm;
    }
    return L;
}

BackdropTables backdrop_tables(const Backdrop& b) {
    BackdropTables T;
    if (b.rgb.empty()) return T;
    const int w = b.width, h = b.height;
    const size_t stride = size_t(w) + 1;
    T.conditional.resize(size_t(h) * stride);
    T.marginal.resize(size_t(h) + 1);
    std::vector<double> row_total(static_cast<size_t>(h));
    for (int y = 0; y < h; y++) {
        // Row y covers polar angles around theta = (y + 1/2) / h * pi.
        const float sin_theta = std::sin((float(y) + 0.5f) / float(h) * kPi);
        float* cdf = &T.conditional[size_t(y) * stride];
        const vec3* row = &b.rgb[size_t(y) * size_t(w)];
        double run = 0;
        cdf[0] = 0;
        for (int x = 0; x < w; x++) {
            run += double(luminance(row[x])) * sin_theta;
            cdf[x + 1] = float(run);
        }
        row_total[size_t(y)] = run;
        for (int x = 1; x <= w; x++) cdf[x] = run > 0 ? float(cdf[x] / run) : float(x) / float(w);
    }
    double run = 0;
    T.marginal[0] = 0;
    for (int y = 0; y < h; y++) {
        run += row_total[size_t(y)];
        T.marginal[size_t(y) + 1] = float(run);
    }
    for (int y = 1; y <= h; y++) T.marginal[size_t(y)] = run > 0 ? float(T.marginal[size_t(y)] / run) : float(y) / float(h);
    T.mean = float(run / (double(w) * double(h)));
    return T;
}

}  // namespace phh
