// The host scene in two parts: what was authored (mesh streams, materials, images, lights, backdrop,
// viewpoints), which generators and the glTF importer fill in, and the device arrays compiled from it
// (the flat arrays ph_scene_desc points at: BVH, triangle records in BVH order, texel pool, light and
// backdrop sampling tables). Compilation is a fixed pipeline of producers (tables.cpp) whose float
// arithmetic follows the reference's host code (src/scene/scene.cpp:13-255) step for step, because the
// triangle IDs, light tables and CDFs it yields are compared bit for bit with the reference's.
#pragma once

#include <string>
#include <vector>

#include "bvh.hpp"
#include "hmath.hpp"
#include "photon_types.h"

namespace phh {

// An RGBA8 image (R in the low byte, first row at the top), as a material texture.
struct Image8 {
    int width = 0, height = 0;
    std::vector<uint32_t> px;
    bool tile = true;   // repeat outside [0, 1); clamp to the edge otherwise
    bool srgb = false;  // colour data stored sRGB-encoded
};

// Light arriving from infinitely far away: an equirectangular radiance image, or one colour when the
// image is empty; scaled by `gain` and turned by `yaw` radians about +y.
struct Backdrop {
    std::vector<vec3> rgb;
    int width = 0, height = 0;
    vec3 flat = vec3(0);
    float gain = 1.0f;
    float yaw = 0.0f;
};

// A look-at camera: vertical field of view in degrees, thin-lens radius, and the focus distance (0
// focuses on the point looked at).
struct Viewpoint {
    vec3 eye = vec3(0, 1, 5);
    vec3 aim = vec3(0, 1, 0);
    vec3 sky = vec3(0, 1, 0);
    float fov_deg = 45.0f;
    float lens = 0.0f;
    float focus = 0.0f;
};

// The device's camera for a viewpoint and an aspect ratio (the reference's Camera::view, scene.cpp:13-25).
ph_camera_view camera_frame(const Viewpoint& v, float aspect);

// Shared vertex streams (each either empty or one entry per vertex), three corners per triangle and one
// material per triangle.
struct Mesh {
    std::vector<vec3> pos;
    std::vector<vec3> nrm;
    std::vector<vec2> uv;
    std::vector<vec4> tan;
    std::vector<int> corner;
    std::vector<int> mat;
    int triangles() const { return int(corner.size() / 3); }
};

// Appends `idx.size() / 3` triangles over the vertices `p` (with optional normals and uvs) in one
// material. A stream one part of the mesh lacks is padded with zeros so every stream stays complete.
void append_mesh(Mesh& m, const std::vector<vec3>& p, const std::vector<int>& idx, int material, const std::vector<vec3>& n = {},
                 const std::vector<vec2>& t = {});

// Emitters and the split of light samples between the three kinds of light.
struct LightTables {
    std::vector<ph_emissive_tri> emitters;
    std::vector<ph_alias_entry> alias;  // over the emitters, by emitted power
    std::vector<int> slot_of_tri;       // BVH-order triangle -> emitter index, -1 if it does not emit
    float emitted_power = 0.0f;
    float p_backdrop = 0.0f, p_emitters = 0.0f, p_punctual = 0.0f;
};

// Importance-sampling tables of the backdrop image: the marginal CDF over rows (height + 1 entries) and
// one conditional CDF per row (width + 1 entries each), weighted by luminance * sin(theta).
struct BackdropTables {
    std::vector<float> marginal, conditional;
    float mean = 0.0f;  // integral of luminance * sin(theta) over the image, per texel
};

struct Compiled {
    BvhTree tree;
    std::vector<ph_tri_geom> tri;  // BVH order
    std::vector<int> corner;       // BVH order, three per triangle
    std::vector<uint32_t> texel_pool;
    std::vector<ph_texture_info> texture_table;
    std::vector<float> srgb_lut;  // 256 decoded sRGB code values
    LightTables lights;
    BackdropTables backdrop;
    bool done = false;
};

class Scene {
public:
    Mesh mesh;
    std::vector<ph_material> materials;
    std::vector<Image8> images;
    std::vector<ph_punctual_light> lamps;
    Backdrop backdrop;
    Viewpoint view;
    std::vector<Viewpoint> file_views;  // every camera a scene file defined, in file order
    std::string label;
    std::vector<vec4> sphere_meta;      // centre and radius of each procedural sphere (generators)

    Compiled out;

    int add_material(const ph_material& m);
    int add_image(Image8 img);
    // Missing tangents, the BVH, the BVH-order triangle records, the texel pool and the light and
    // backdrop tables. A scene without materials gets the glTF default.
    void compile(const BvhSettings& settings = {});
    // After materials or the backdrop of a compiled scene changed: the light and backdrop tables again
    // (the tree stays).
    void relight();
    ph_scene_desc describe() const;
};

ph_material default_material();

// ---- producers (tables.cpp) ----------------------------------------------------------------------

// Per-vertex tangents from the uv parametrisation (Lengyel), for normal mapping: only when the mesh has
// uvs, has no tangents yet and some material uses a normal map.
void fill_tangents(Mesh& m, const std::vector<ph_material>& materials);
// Vose's alias table over non-negative weights, in double precision.
std::vector<ph_alias_entry> alias_table(const std::vector<float>& weights);
LightTables light_tables(const Scene& s);
BackdropTables backdrop_tables(const Backdrop& b);

// ---- procedural content (generators.cpp) ---------------------------------------------------------

void add_sphere(Scene& s, vec3 centre, float radius, int material, int segments = 48);
void add_quad(Scene& s, vec3 corner, vec3 edge_u, vec3 edge_v, int material);
void add_box(Scene& s, vec3 centre, vec3 half, float rotation_y, int material);
void make_sky(Backdrop& b, vec3 sun_direction, float sun_intensity, int width);
bool make_builtin_scene(const std::string& name, Scene& s);
void make_sphere_soup(Scene& s, int count, float box, float alpha_fraction, uint64_t seed);
void make_sphere_field(Scene& s, int count, uint64_t seed, int env_width);
void make_gradient_sky(Backdrop& b, vec3 sun, int width);

}  // namespace phh
