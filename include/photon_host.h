/* Host-side scene producers of the MI355X rendering core, as a C ABI.
 *
 * These build what the device library consumes: the binned-SAH BVH over 64-byte dual-box nodes, the
 * BVH-order triangle array, the texel pool, the emissive alias table and the environment CDFs. Each
 * entry point replaces a piece of the reference's host side:
 *   phs_scene_* construction          <- Scene::add_mesh / add_material (src/scene/scene.h:61-64)
 *   phs_scene_builtin                 <- make_builtin_scene (src/scene/builtin.h:12-13)
 *   phs_scene_build                   <- Scene::build (src/scene/scene.h:66-67, scene.cpp:212-255)
 *   phs_scene_rebuild_lights          <- Scene::rebuild_lights (src/scene/scene.h:68-70)
 *   phs_scene_desc                    <- Scene::view (src/scene/scene.h:72-73)
 *   phs_scene_camera_view             <- Camera::view (src/scene/scene.h:30-39)
 *   phs_scene_order / phs_scene_stats <- Bvh::order, BvhStats (src/bvh/bvh.h:39-52)
 * plus the synthetic workload generators (sphere soups, sphere fields, ray sets) that have no
 * counterpart in the reference.
 * Every function returns 0 on success and a negative code on failure; phs_last_error() then holds a
 * message for the calling thread.
 */
#ifndef PHOTON_HOST_H
#define PHOTON_HOST_H

#include "photon_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phs_scene phs_scene;

typedef struct phs_bvh_stats {
    int32_t nodes, leaves, max_depth;
    float sah_cost;
    double build_ms;
} phs_bvh_stats;

const char* phs_last_error(void);

int phs_scene_create(phs_scene** out);
void phs_scene_destroy(phs_scene* s);

/* "cornell", "cornell-c1", "materials", "glass", "furnace". */
int phs_scene_builtin(phs_scene* s, const char* name);
/* `count` spheres of 1024 triangles with centres uniform in [0, box]^3, radii log-uniform in [0.5, 5]
 * and a seeded fraction of alpha-masked triangles. */
int phs_scene_spheres(phs_scene* s, int count, float box, float alpha_fraction, uint64_t seed);
/* `count` spheres of 1024 triangles over a ground quad, random metal/glass/opaque materials, count/64
 * emitters, and an HDR gradient sky with a sun disk (env_width x env_width/2 texels). */
int phs_scene_field(phs_scene* s, int count, uint64_t seed, int env_width);

int phs_scene_add_mesh(phs_scene* s, const float* positions, int num_vertices, const int32_t* indices, int num_indices, int material,
                       const float* normals, const float* uvs);
int phs_scene_add_material(phs_scene* s, const ph_material* m);
int phs_scene_set_material(phs_scene* s, int index, const ph_material* m);
int phs_scene_get_material(const phs_scene* s, int index, ph_material* out);
int phs_scene_add_texture(phs_scene* s, int width, int height, const uint32_t* texels, int repeat, int srgb);
int phs_scene_add_punctual(phs_scene* s, const ph_punctual_light* l);
int phs_scene_set_env_color(phs_scene* s, const float rgb[3], float intensity);
int phs_scene_set_env_image(phs_scene* s, int width, int height, const float* rgb, float intensity, float rotation);
/* The reference's procedural sky (make_sky, src/scene/builtin.h:15-16). */
int phs_scene_set_sky(phs_scene* s, const float sun[3], float sun_intensity, int width);
int phs_scene_clear_normals(phs_scene* s);
int phs_scene_set_camera(phs_scene* s, const float position[3], const float target[3], const float up[3], float vfov_degrees, float aperture,
                         float focus_distance);
int phs_scene_get_camera(const phs_scene* s, float position[3], float target[3], float up[3], float* vfov_degrees);
int phs_scene_camera_view(const phs_scene* s, float aspect, ph_camera_view* out);

/* Builds BVH and tables. bins/max_leaf/traversal_cost/threads <= 0 take the defaults 32/8/1/hardware. */
int phs_scene_build(phs_scene* s, int bins, int max_leaf, float traversal_cost, int threads);
int phs_scene_rebuild_lights(phs_scene* s);
/* Pointers stay valid until the scene is modified, rebuilt or destroyed. */
int phs_scene_desc(const phs_scene* s, ph_scene_desc* out);
int phs_scene_num_tris(const phs_scene* s);
int phs_scene_num_materials(const phs_scene* s);
const int32_t* phs_scene_order(const phs_scene* s);
int phs_scene_stats(const phs_scene* s, phs_bvh_stats* out);
/* Host bytes of the device-resident scene arrays (for admission by footprint). */
int64_t phs_scene_device_bytes(const phs_scene* s);

/* The authored (pre-build) arrays of a scene: the inputs of Scene::build. Pointers stay valid until
 * the scene is modified or destroyed. */
typedef struct phs_raw_scene {
    const float* positions;
    int32_t num_vertices;
    const float* normals; /* NULL when absent */
    const float* uvs;     /* NULL when absent */
    const int32_t* indices;
    int32_t num_tris;
    const int32_t* tri_material;
    const ph_material* materials;
    int32_t num_materials;
    int32_t num_textures;
    const int32_t* tex_width;
    const int32_t* tex_height;
    const uint32_t* const* tex_texels;
    const int32_t* tex_repeat;
    const int32_t* tex_srgb;
    const ph_punctual_light* punctual;
    int32_t num_punctual;
    const float* env_texels; /* NULL when the environment is a constant colour */
    int32_t env_width, env_height;
    float env_color[3];
    float env_intensity, env_rotation;
    float cam_position[3], cam_target[3], cam_up[3];
    float cam_vfov, cam_aperture, cam_focus;
    const float* tangents; /* 4 floats per vertex (xyz, handedness), NULL when absent */
} phs_raw_scene;

int phs_scene_raw(phs_scene* s, phs_raw_scene* out);

/* ---- scene files and images ---------------------------------------------------------------------- */

/* Appends a glTF 2.0 file (.gltf or .glb) to the scene: meshes of the default scene under the node
 * hierarchy, PBR materials with the transmission / ior / emissive-strength extensions, PNG and JPEG textures,
 * perspective cameras and KHR_lights_punctual lights. The first camera becomes the scene camera.
 * <- load_gltf (src/scene/gltf.h:12, gltf.cpp:72-497). */
int phs_scene_load_gltf(phs_scene* s, const char* path);
/* Writes meshes, materials, camera and punctual lights as a .gltf with an embedded buffer.
 * <- save_gltf (src/scene/gltf.h:16, gltf.cpp:501-640). */
int phs_scene_save_gltf(const phs_scene* s, const char* path);
/* Cameras found in loaded files (Scene::cameras, src/scene/scene.h:55). */
int phs_scene_num_cameras(const phs_scene* s);
int phs_scene_camera_at(const phs_scene* s, int index, float position[3], float target[3], float up[3], float* vfov_degrees);
/* Equirectangular environment from a Radiance .hdr, a PFM or an 8-bit PNG read as sRGB.
 * <- load_environment (src/scene/image.h:18, image.cpp:53-81). */
int phs_scene_load_environment(phs_scene* s, const char* path, float intensity, float rotation);

/* Decodes a PNG or a baseline JPEG to RGBA8 (R in the low byte, row 0 at the top). With rgba NULL only
 * the size is returned. <- decode_image (src/scene/image.h:14, image.cpp:23-37). */
int phs_image_decode(const uint8_t* data, size_t size, int* width, int* height, uint32_t* rgba, size_t capacity);
/* <- save_png (src/scene/image.h:20, image.cpp:84-86). */
int phs_image_save_png(const char* path, int width, int height, const uint32_t* rgba);
/* Linear RGB float rows, top to bottom in memory, bottom to top in the file.
 * <- save_pfm / load_pfm (src/scene/image.h:21-22, image.cpp:88-116). */
int phs_image_save_pfm(const char* path, int width, int height, const float* rgb);
int phs_image_load_pfm(const char* path, int* width, int* height, float* rgb, size_t capacity);
/* Parses JSON text and writes it canonically (members sorted, numbers as %.17g, no white space) into
 * `out` when `capacity` exceeds its length; returns the length, or a negative code on a parse error.
 * <- json::parse (src/scene/json.h:55). */
int64_t phs_json_canonical(const char* text, char* out, size_t capacity);

/* Pinhole rays through pixel centres (camera_ray with jitter 0.5, integrator.h:266-280), row-major. */
int phs_camera_rays(const ph_camera_view* cam, int width, int height, ph_ray* out);
/* Shadow rays of a sphere scene: origin just outside a random sphere, target a uniform point on
 * another random sphere, tmax = distance * (1 - 1e-4). */
int phs_sphere_shadow_rays(const phs_scene* s, int count, uint64_t seed, ph_ray* out);

#ifdef __cplusplus
}
#endif

#endif /* PHOTON_HOST_H */
