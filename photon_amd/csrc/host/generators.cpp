// Procedural scenes.
//
// The reference's built-in scenes are restated with the same geometry so that tests written against
// them (furnace, Cornell, materials, glass) keep their meaning (src/scene/builtin.cpp:12-186). The
// synthetic workload scenes are new:
//   cornell-c1  the reference Cornell geometry with the workload's materials and a 60 degree camera;
//   spheres     N spheres of 1024 triangles (a tetrahedron subdivided 4 times and projected), seeded
//               centres, log-uniform radii and radial vertex noise of 5% of the radius, an optional
//               alpha-mask fraction with a seeded 64x64 binary alpha texture (traversal workload);
//   field       1024 such spheres with random metal/glass/opaque materials, 64 emitters and a ground
//               quad under an HDR sky with a 0.5 degree sun (path tracing workload).
// All randomness comes from the PCG32 generator with explicit (seed, stream) pairs.
#include <algorithm>
#include <cmath>
#include <map>

#include "scene.hpp"

namespace phh {

void add_sphere(Scene& scene, vec3 centre, float radius, int material, int segments) {
    int rings = segments / 2;
    std::vector<vec3> pos, nrm;
    std::vector<vec2> uv;
    std::vector<int> idx;
    for (int r = 0; r <= rings; r++) {
        float theta = kPi * float(r) / float(rings);
        for (int s = 0; s <= segments; s++) {
            float phi = kTwoPi * float(s) / float(segments);
            vec3 n(std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi));
            pos.push_back(centre + n * radius);
            nrm.push_back(n);
            uv.push_back(vec2(float(s) / float(segments), float(r) / float(rings)));
        }
    }
    const int row = segments + 1;
    for (int r = 0; r < rings; r++)
        for (int s = 0; s < segments; s++) {
            int a = r * row + s, b = a + row;
            if (r > 0) idx.insert(idx.end(), {a, a + 1, b});
            if (r < rings - 1) idx.insert(idx.end(), {a + 1, b + 1, b});
        }
    append_mesh(scene.mesh, pos, idx, material, nrm, uv);
}

void add_quad(Scene& scene, vec3 corner, vec3 eu, vec3 ev, int material) {
    vec3 n = normalize(cross(eu, ev));
    std::vector<vec3> pos = {corner, corner + eu, corner + eu + ev, corner + ev};
    std::vector<vec3> nrm(4, n);
    std::vector<vec2> uv = {vec2(0, 0), vec2(1, 0), vec2(1, 1), vec2(0, 1)};
    append_mesh(scene.mesh, pos, {0, 1, 2, 0, 2, 3}, material, nrm, uv);
}

void add_box(Scene& scene, vec3 centre, vec3 h, float rot, int material) {
    const float c = std::cos(rot), s = std::sin(rot);
    auto place = [&](vec3 p) { return centre + vec3(c * p.x + s * p.z, p.y, -s * p.x + c * p.z); };
    auto turn = [&](vec3 d) { return vec3(c * d.x + s * d.z, d.y, -s * d.x + c * d.z); };
    // Six faces with outward winding: origin corner, then the two edges.
    const vec3 faces[6][3] = {
        {vec3(-h.x, -h.y, h.z), vec3(2 * h.x, 0, 0), vec3(0, 2 * h.y, 0)},
        {vec3(h.x, -h.y, -h.z), vec3(-2 * h.x, 0, 0), vec3(0, 2 * h.y, 0)},
        {vec3(h.x, -h.y, h.z), vec3(0, 0, -2 * h.z), vec3(0, 2 * h.y, 0)},
        {vec3(-h.x, -h.y, -h.z), vec3(0, 0, 2 * h.z), vec3(0, 2 * h.y, 0)},
        {vec3(-h.x, h.y, h.z), vec3(2 * h.x, 0, 0), vec3(0, 0, -2 * h.z)},
        {vec3(-h.x, -h.y, -h.z), vec3(2 * h.x, 0, 0), vec3(0, 0, 2 * h.z)},
    };
    for (const auto& f : faces) add_quad(scene, place(f[0]), turn(f[1]), turn(f[2]), material);
}

void make_sky(Backdrop& env, vec3 sun, float sun_intensity, int width) {
    const int height = width / 2;
    env.width = width;
    env.height = height;
    env.rgb.assign(size_t(width) * size_t(height), vec3(0));
    sun = normalize(sun);
    const float sun_cos = std::cos(0.6f * kPi / 180.0f);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            vec3 d = equirect_to_direction(vec2((float(x) + 0.5f) / float(width), (float(y) + 0.5f) / float(height)));
            vec3 col;
            if (d.y >= 0) {
                float t = std::pow(d.y, 0.45f);
                col = lerp(vec3(0.85f, 0.9f, 1.0f), vec3(0.22f, 0.42f, 0.85f), t) * 1.2f;
                float cs = dot(d, sun);
                col += vec3(1.0f, 0.85f, 0.6f) * (0.35f * std::pow(std::max(cs, 0.0f), 32.0f));
                if (cs > sun_cos) col = vec3(1.0f, 0.95f, 0.85f) * sun_intensity;
            } else {
                col = vec3(0.25f, 0.23f, 0.21f) * (0.5f + 0.5f * std::exp(8.0f * d.y));
            }
            env.rgb[size_t(y) * size_t(width) + size_t(x)] = col;
        }
}

static ph_material matte(vec3 c) {
    ph_material m = default_material();
    m.base_color[0] = c.x, m.base_color[1] = c.y, m.base_color[2] = c.z, m.base_color[3] = 1;
    m.metallic = 0;
    m.roughness = 1;
    return m;
}

static void cornell_geometry(Scene& s, int white, int red, int green, int lamp, int short_box, int tall_box) {
    const float S = 5.5f;
    add_quad(s, vec3(0, 0, 0), vec3(0, 0, S), vec3(S, 0, 0), white);   // floor
    add_quad(s, vec3(0, S, 0), vec3(S, 0, 0), vec3(0, 0, S), white);   // ceiling
    add_quad(s, vec3(0, 0, S), vec3(0, S, 0), vec3(S, 0, 0), white);   // back wall
    add_quad(s, vec3(S, 0, 0), vec3(0, 0, S), vec3(0, S, 0), red);     // left wall (x = S)
    add_quad(s, vec3(0, 0, S), vec3(0, 0, -S), vec3(0, S, 0), green);  // right wall (x = 0)
    add_quad(s, vec3(2.13f, S - 0.01f, 2.27f), vec3(1.3f, 0, 0), vec3(0, 0, 1.05f), lamp);
    add_box(s, vec3(1.85f, 0.825f, 1.69f), vec3(0.825f), -0.29f, short_box);
    add_box(s, vec3(3.68f, 1.65f, 3.51f), vec3(0.825f, 1.65f, 0.825f), 0.29f, tall_box);
    s.view.eye = vec3(2.78f, 2.73f, -8.0f);
    s.view.aim = vec3(2.78f, 2.73f, 0.0f);
}

bool make_builtin_scene(const std::string& name, Scene& s) {
    s = Scene();
    s.label = name;
    if (name == "cornell") {
        int white = s.add_material(matte(vec3(0.725f, 0.71f, 0.68f)));
        int red = s.add_material(matte(vec3(0.63f, 0.065f, 0.05f)));
        int green = s.add_material(matte(vec3(0.14f, 0.45f, 0.091f)));
        ph_material light = matte(vec3(0.78f));
        light.emission[0] = 17, light.emission[1] = 12, light.emission[2] = 4;
        int lamp = s.add_material(light);
        cornell_geometry(s, white, red, green, lamp, white, white);
        s.view.fov_deg = 39.3f;
        return true;
    }
    if (name == "cornell-c1") {
        int white = s.add_material(matte(vec3(0.73f)));
        int red = s.add_material(matte(vec3(0.65f, 0.05f, 0.05f)));
        int green = s.add_material(matte(vec3(0.12f, 0.45f, 0.15f)));
        ph_material light = matte(vec3(0.78f));
        light.emission[0] = 17, light.emission[1] = 12, light.emission[2] = 4;
        int lamp = s.add_material(light);
        ph_material gold = default_material();  // GGX conductor
        gold.base_color[0] = 1.0f, gold.base_color[1] = 0.78f, gold.base_color[2] = 0.34f;
        gold.metallic = 1;
        gold.roughness = 0.3f;
        int short_box = s.add_material(gold);
        ph_material glass = default_material();  // rough dielectric
        glass.metallic = 0;
        glass.transmission = 1;
        glass.ior = 1.5f;
        glass.roughness = 0.1f;
        int tall_box = s.add_material(glass);
        cornell_geometry(s, white, red, green, lamp, short_box, tall_box);
        s.view.fov_deg = 60.0f;
        return true;
    }
    if (name == "materials" || name == "glass") {
        ph_material ground = matte(vec3(0.4f, 0.4f, 0.42f));
        ground.roughness = 0.9f;
        int g = s.add_material(ground);
        add_quad(s, vec3(-20, 0, 20), vec3(40, 0, 0), vec3(0, 0, -40), g);
        if (name == "materials") {
            for (int row = 0; row < 3; row++)
                for (int col = 0; col < 6; col++) {
                    ph_material m = default_material();
                    m.roughness = float(col) / 5.0f;
                    if (row == 0) {
                        m.base_color[0] = 0.8f, m.base_color[1] = 0.15f, m.base_color[2] = 0.1f;
                        m.metallic = 0;
                    } else if (row == 1) {
                        m.base_color[0] = 0.95f, m.base_color[1] = 0.75f, m.base_color[2] = 0.35f;
                        m.metallic = 1;
                    } else {
                        m.metallic = 0;
                        m.transmission = 1;
                        m.ior = 1.5f;
                    }
                    add_sphere(s, vec3(-5.0f + 2.0f * float(col), 0.8f, 2.4f * float(row) - 2.4f), 0.8f, s.add_material(m));
                }
            s.view.eye = vec3(0, 6.5f, 11.0f);
            s.view.aim = vec3(0, 0.3f, 0);
            s.view.fov_deg = 40;
        } else {
            ph_material glass = default_material();
            glass.metallic = 0;
            glass.roughness = 0;
            glass.transmission = 1;
            glass.ior = 1.5f;
            add_sphere(s, vec3(-1.2f, 1.0f, 0), 1.0f, s.add_material(glass), 96);
            ph_material frosted = glass;
            frosted.roughness = 0.35f;
            add_sphere(s, vec3(1.2f, 1.0f, 0), 1.0f, s.add_material(frosted), 96);
            add_box(s, vec3(0, 0.5f, -2.5f), vec3(2.5f, 0.5f, 0.3f), 0, s.add_material(matte(vec3(0.9f, 0.3f, 0.2f))));
            s.view.eye = vec3(0, 2.2f, 6.0f);
            s.view.aim = vec3(0, 0.9f, 0);
            s.view.fov_deg = 40;
        }
        make_sky(s.backdrop, vec3(0.5f, 0.7f, 0.4f), 30000.0f, 1024);
        return true;
    }
    if (name == "furnace") {
        ph_material m = matte(vec3(1.0f));
        m.ior = 1.0f;  // index matched: a pure Lambertian
        add_sphere(s, vec3(0), 1.0f, s.add_material(m), 64);
        s.backdrop.flat = vec3(1.0f);
        s.view.eye = vec3(0, 0, 4);
        s.view.aim = vec3(0);
        s.view.fov_deg = 40;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------------------------------
// Workload generators.

namespace {

// Unit sphere mesh of 4 * 4^levels triangles: a tetrahedron whose faces are split in four `levels`
// times, with midpoints shared between neighbouring faces and every vertex projected to the sphere.
struct UnitMesh {
    std::vector<vec3> dirs;
    std::vector<int> idx;
};

UnitMesh subdivided_tetrahedron(int levels) {
    UnitMesh m;
    const double k = 1.0 / std::sqrt(3.0);
    const double base[4][3] = {{k, k, k}, {k, -k, -k}, {-k, k, -k}, {-k, -k, k}};
    std::vector<double> p;  // 3 doubles per vertex
    for (const auto& b : base) p.insert(p.end(), {b[0], b[1], b[2]});
    std::vector<int> tri = {0, 1, 2, 0, 3, 1, 0, 2, 3, 1, 3, 2};
    for (int l = 0; l < levels; l++) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double x = p[3 * size_t(a)] + p[3 * size_t(b)], y = p[3 * size_t(a) + 1] + p[3 * size_t(b) + 1],
                   z = p[3 * size_t(a) + 2] + p[3 * size_t(b) + 2];
            double n = std::sqrt(x * x + y * y + z * z);
            p.insert(p.end(), {x / n, y / n, z / n});
            int id = int(p.size() / 3) - 1;
            mid.emplace(key, id);
            return id;
        };
        std::vector<int> next;
        for (size_t t = 0; t < tri.size(); t += 3) {
            int a = tri[t], b = tri[t + 1], c = tri[t + 2];
            int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            next.insert(next.end(), {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca});
        }
        tri.swap(next);
    }
    for (size_t v = 0; v < p.size() / 3; v++) m.dirs.push_back(vec3(float(p[3 * v]), float(p[3 * v + 1]), float(p[3 * v + 2])));
    m.idx = tri;
    return m;
}

const UnitMesh& unit_sphere_1024() {
    static const UnitMesh mesh = subdivided_tetrahedron(4);
    return mesh;
}

float log_uniform(Pcg32& rng, float lo, float hi) {
    return float(double(lo) * std::exp(double(rng.next()) * std::log(double(hi) / double(lo))));
}

// One noisy sphere of 1024 triangles; returns the index of its first triangle.
int add_noisy_sphere(Scene& s, vec3 centre, float radius, int material, uint64_t seed, uint64_t stream) {
    const UnitMesh& m = unit_sphere_1024();
    Pcg32 rng(seed, stream);
    std::vector<vec3> pos, nrm;
    std::vector<vec2> uv;
    for (vec3 d : m.dirs) {
        float r = radius * (1.0f + 0.05f * (2.0f * rng.next() - 1.0f));
        pos.push_back(centre + d * r);
        nrm.push_back(d);
        double u = 0.5 + std::atan2(double(d.z), double(d.x)) / (2.0 * 3.14159265358979323846);
        double v = std::acos(std::max(-1.0, std::min(1.0, double(d.y)))) / 3.14159265358979323846;
        uv.push_back(vec2(float(u), float(v)));
    }
    int first = s.mesh.triangles();
    append_mesh(s.mesh, pos, m.idx, material, nrm, uv);
    s.sphere_meta.push_back(vec4(centre, radius));
    return first;
}

}  // namespace

// Alpha texture for masked triangles: 64x64, each texel opaque or clear with probability 1/2.
static int add_alpha_texture(Scene& s, uint64_t seed) {
    Image8 t;
    t.width = t.height = 64;
    t.tile = true;
    t.srgb = true;
    Pcg32 rng(seed, 0xA1FAu);
    t.px.resize(64 * 64);
    for (auto& px : t.px) px = 0x00FFFFFFu | (rng.next() < 0.5f ? 0xFF000000u : 0u);
    s.images.push_back(t);
    return int(s.images.size()) - 1;
}

void make_sphere_soup(Scene& s, int count, float box, float alpha_fraction, uint64_t seed) {
    s = Scene();
    s.label = "spheres";
    ph_material opaque = matte(vec3(0.7f));
    int m_opaque = s.add_material(opaque);
    int m_masked = m_opaque;
    if (alpha_fraction > 0) {
        ph_material masked = matte(vec3(0.7f));
        masked.alpha_mode = PH_ALPHA_MASK;
        masked.alpha_cutoff = 0.5f;
        masked.base_color_tex = add_alpha_texture(s, seed);
        m_masked = s.add_material(masked);
    }
    Pcg32 rng(seed, 0x5EEDu);
    for (int i = 0; i < count; i++) {
        vec3 c(rng.next() * box, rng.next() * box, rng.next() * box);
        float r = log_uniform(rng, 0.5f, 5.0f);
        add_noisy_sphere(s, c, r, m_opaque, seed, uint64_t(i) + 1);
    }
    if (alpha_fraction > 0) {
        Pcg32 pick(seed, 0xA1F4u);
        for (int& m : s.mesh.mat) m = pick.next() < alpha_fraction ? m_masked : m_opaque;
    }
    s.view.eye = vec3(0.5f * box, 0.5f * box, -0.8f * box);
    s.view.aim = vec3(0.5f * box, 0.5f * box, 0.5f * box);
    s.view.fov_deg = 60.0f;
}

// The sky of the path tracing workload: luminance 0.2 at the horizon rising linearly in d.y to 1.0 at
// the zenith, 0.2 below the horizon, and a sun disk of 0.5 degrees diameter with luminance 1e4.
void make_gradient_sky(Backdrop& env, vec3 sun, int width) {
    const int height = width / 2;
    env.width = width;
    env.height = height;
    env.rgb.assign(size_t(width) * size_t(height), vec3(0));
    sun = normalize(sun);
    const double sun_cos = std::cos(0.25 * 3.14159265358979323846 / 180.0);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            vec3 d = equirect_to_direction(vec2((float(x) + 0.5f) / float(width), (float(y) + 0.5f) / float(height)));
            float lum = d.y > 0 ? 0.2f + 0.8f * d.y : 0.2f;
            if (double(dot(d, sun)) > sun_cos) lum = 1e4f;
            env.rgb[size_t(y) * size_t(width) + size_t(x)] = vec3(lum);
        }
}

void make_sphere_field(Scene& s, int count, uint64_t seed, int env_width) {
    s = Scene();
    s.label = "field";
    Pcg32 rng(seed, 0xF1E1Du);
    ph_material ground = matte(vec3(0.5f));
    int g = s.add_material(ground);
    add_quad(s, vec3(-150, 0, 150), vec3(300, 0, 0), vec3(0, 0, -300), g);
    for (int i = 0; i < count; i++) {
        ph_material m = default_material();
        m.metallic = rng.next() < 0.3f ? 1.0f : 0.0f;
        m.transmission = rng.next() < 0.2f ? 1.0f : 0.0f;
        m.roughness = 0.02f + 0.98f * rng.next();
        for (int k = 0; k < 3; k++) m.base_color[k] = 0.05f + 0.9f * rng.next();
        m.ior = 1.5f;
        float r = log_uniform(rng, 0.5f, 5.0f);
        vec3 c(-100.0f + 200.0f * rng.next(), 0.0f, -100.0f + 200.0f * rng.next());
        c.y = r + 20.0f * rng.next();
        if (i % (count / 64 > 0 ? count / 64 : 1) == 0) {
            float e = log_uniform(rng, 1.0f, 50.0f);
            m.emission[0] = m.emission[1] = m.emission[2] = e;
        }
        add_noisy_sphere(s, c, r, s.add_material(m), seed, uint64_t(i) + 1);
    }
    make_gradient_sky(s.backdrop, vec3(0.4f, 0.7f, 0.3f), env_width);
    s.view.eye = vec3(0, 45, -170);
    s.view.aim = vec3(0, 5, 0);
    s.view.fov_deg = 60.0f;
}

}  // namespace phh
