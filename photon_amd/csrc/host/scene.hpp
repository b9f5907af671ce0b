// The host scene: geometry, materials, textures and lights as authored, plus the derived device inputs
// (BVH-order triangles, texel pool, emissive alias table, environment CDFs) produced by Scene::build.
// Mirrors the reference's Scene (src/scene/scene.h:41-95) so the same producers feed the device library.
#pragma once

#include <string>
#include <vector>

#include "bvh.hpp"
#include "hmath.hpp"
#include "photon_types.h"

namespace phh {

struct Texture {
    int width = 0, height = 0;
    std::vector<uint32_t> texels;  // RGBA8, row 0 at the top
    bool repeat = true;
    bool srgb = false;
};

struct Environment {
    std::vector<vec3> texels;  // equirectangular linear radiance
    int width = 0, height = 0;
    vec3 color = vec3(0);
    float intensity = 1.0f;
    float rotation = 0.0f;
};

struct Camera {
    vec3 position = vec3(0, 1, 5);
    vec3 target = vec3(0, 1, 0);
    vec3 up = vec3(0, 1, 0);
    float vfov = 45.0f;  // degrees
    float aperture = 0.0f;
    float focus_distance = 0.0f;  // 0: distance to the target
    ph_camera_view view(float aspect) const;
};

ph_material default_material();

class Scene {
public:
    std::vector<vec3> positions;
    std::vector<vec3> normals;
    std::vector<vec2> uvs;
    std::vector<vec4> tangents;
    std::vector<int> indices;
    std::vector<int> tri_material;
    std::vector<ph_material> materials;
    std::vector<Texture> textures;
    std::vector<ph_punctual_light> punctual;
    Environment env;
    Camera camera;
    std::vector<Camera> cameras;  // every camera a scene file defined, in file order
    std::string name;
    std::vector<vec4> spheres;  // generator metadata: centre and radius of each procedural sphere

    int num_tris() const { return int(indices.size() / 3); }
    void add_mesh(const std::vector<vec3>& pos, const std::vector<int>& idx, int material, const std::vector<vec3>& nrm = {},
                  const std::vector<vec2>& uv = {});
    int add_material(const ph_material& m);
    void build(const BvhOptions& options = {});
    void rebuild_lights();
    ph_scene_desc desc() const;

    // Built data.
    BvhResult bvh;
    std::vector<ph_tri_geom> tris;
    std::vector<int> bvh_indices;
    std::vector<ph_emissive_tri> emissive;
    std::vector<ph_alias_entry> emissive_alias;
    std::vector<int> tri_emissive;
    float emissive_power = 0;
    std::vector<uint32_t> texel_pool;
    std::vector<ph_texture_info> texture_infos;
    std::vector<float> env_marginal, env_conditional;
    float env_integral = 0;
    std::vector<float> srgb_lut;
    float p_env = 0, p_emissive = 0, p_punctual = 0;
    bool built = false;

private:
    void compute_tangents();
    void build_lights();
    void build_environment();
};

std::vector<ph_alias_entry> build_alias_table(const std::vector<float>& weights);

// Procedural content (builtin.cpp / generators.cpp).
void add_sphere(Scene& s, vec3 centre, float radius, int material, int segments = 48);
void add_quad(Scene& s, vec3 corner, vec3 edge_u, vec3 edge_v, int material);
void add_box(Scene& s, vec3 centre, vec3 half, float rotation_y, int material);
void make_sky(Environment& env, vec3 sun_direction, float sun_intensity, int width);
bool make_builtin_scene(const std::string& name, Scene& s);
void make_sphere_soup(Scene& s, int count, float box, float alpha_fraction, uint64_t seed);
void make_sphere_field(Scene& s, int count, uint64_t seed, int env_width);
void make_gradient_sky(Environment& env, vec3 sun, int width);

}  // namespace phh
