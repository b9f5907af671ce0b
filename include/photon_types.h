/* Plain-data layouts shared by the host scene producers, the HIP device library and the CPU oracle.
 *
 * Every struct here is a byte-for-byte restatement of the reference's scene POD types so that a scene
 * built on the host can be handed to the device library (and to the oracle) without conversion:
 *   ph_bvh_node        <- BvhNode        (src/render/scene_view.h:13-22), 64 bytes
 *   ph_tri_geom        <- TriGeom        (src/render/scene_view.h:25-32), 48 bytes
 *   ph_texture_info    <- TextureInfo    (src/render/scene_view.h:38-43), 20 bytes
 *   ph_material        <- Material       (src/render/scene_view.h:46-62), 80 bytes
 *   ph_emissive_tri    <- EmissiveTri    (src/render/scene_view.h:65-68), 8 bytes
 *   ph_alias_entry     <- AliasEntry     (src/render/scene_view.h:71-76), 16 bytes
 *   ph_punctual_light  <- PunctualLight  (src/render/scene_view.h:81-88), 48 bytes
 *   ph_camera_view     <- CameraView     (src/render/scene_view.h:106-115), 64 bytes
 *   ph_scene_desc      <- SceneView      (src/render/scene_view.h:117-150), as pointers + counts
 *   ph_ray / ph_hit    <- Ray / Hit      (src/render/traverse.h:10-20)
 *   ph_path_options    <- PathOptions    (src/render/integrator.h:21-27) minus the host-only ray counter
 *   ph_render_settings <- RenderSettings (src/render/pixel.h:10-16)
 *   ph_svgf_settings   <- SvgfSettings   (src/render/gpu_renderer.h:18-25)
 * Plain C so that ctypes, C++ and HIP all read the same declaration.
 */
#ifndef PHOTON_TYPES_H
#define PHOTON_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Triangle flag: the triangle's material is alpha tested (scene_view.h:34). */
#define PH_TRI_ALPHA_TESTED 1

/* Alpha modes (scene_view.h:36). */
#define PH_ALPHA_OPAQUE 0
#define PH_ALPHA_MASK 1
#define PH_ALPHA_BLEND 2

/* Punctual light types (scene_view.h:78). */
#define PH_LIGHT_POINT 0
#define PH_LIGHT_SPOT 1
#define PH_LIGHT_DIRECTIONAL 2

/* Tone mapping operators (core/color.h:28). */
#define PH_TONEMAP_LINEAR 0
#define PH_TONEMAP_ACES 1
#define PH_TONEMAP_REINHARD 2

/* Display modes (gpu_renderer.h:16). */
#define PH_DISPLAY_COLOR 0
#define PH_DISPLAY_ALBEDO 1
#define PH_DISPLAY_NORMAL 2
#define PH_DISPLAY_DEPTH 3
#define PH_DISPLAY_VARIANCE 4
#define PH_DISPLAY_HEATMAP 5

/* Distance used for "no hit" and for distant lights (core/math.h:25). */
#define PH_INF 1e30f

typedef struct ph_bvh_node {
    float lmin[3];
    int32_t left;   /* leaf: first triangle (BVH order); interior: node index; -1: empty */
    float lmax[3];
    int32_t lcount; /* > 0: leaf with this many triangles */
    float rmin[3];
    int32_t right;
    float rmax[3];
    int32_t rcount;
} ph_bvh_node;

typedef struct ph_tri_geom {
    float v0[3];
    int32_t material;
    float e1[3]; /* v1 - v0 */
    int32_t flags;
    float e2[3]; /* v2 - v0 */
    int32_t pad;
} ph_tri_geom;

typedef struct ph_texture_info {
    int32_t width, height;
    uint32_t offset;     /* first texel in the RGBA8 pool */
    int32_t wrap_repeat; /* 1 repeat, 0 clamp */
    int32_t srgb;        /* texels are sRGB encoded */
} ph_texture_info;

typedef struct ph_material {
    float base_color[4];
    float emission[3];
    float metallic;
    float roughness;
    float ior;
    float transmission;
    float normal_scale;
    float alpha_cutoff;
    int32_t alpha_mode;
    int32_t double_sided;
    int32_t base_color_tex;
    int32_t metallic_roughness_tex; /* G roughness, B metallic */
    int32_t normal_tex;
    int32_t emissive_tex;
    int32_t occlusion_tex;
} ph_material;

typedef struct ph_emissive_tri {
    int32_t tri;
    float area;
} ph_emissive_tri;

typedef struct ph_alias_entry {
    float prob;
    int32_t alias;
    float pmf;
    int32_t pad;
} ph_alias_entry;

typedef struct ph_punctual_light {
    float position[3];
    int32_t type;
    float direction[3];
    float cos_inner;
    float intensity[3];
    float cos_outer;
} ph_punctual_light;

typedef struct ph_camera_view {
    float position[3];
    float tan_half_fov;
    float forward[3];
    float aspect;
    float right[3];
    float aperture;
    float up[3];
    float focus_distance;
} ph_camera_view;

typedef struct ph_ray {
    float o[3];
    float d[3];
    float tmax;
} ph_ray;

typedef struct ph_hit {
    float t;
    int32_t tri; /* BVH order, -1 for a miss */
    float u, v;
} ph_hit;

typedef struct ph_path_options {
    int32_t max_depth; /* path vertices after the camera (default 8) */
    int32_t rr_depth;  /* Russian roulette from this depth (default 3) */
    float clamp;       /* 0: off */
    int32_t nee;       /* next event estimation on/off (default 1) */
} ph_path_options;

typedef struct ph_render_settings {
    int32_t width, height;
    ph_path_options path;
    float exposure;
    int32_t tonemap;
    uint32_t seed;
} ph_render_settings;

typedef struct ph_svgf_settings {
    int32_t atrous_iterations; /* 5 */
    float alpha_color;         /* 0.2 */
    float alpha_moments;       /* 0.2 */
    float phi_color;           /* 4 */
    float phi_normal;          /* 128 */
    float phi_depth;           /* 1 */
} ph_svgf_settings;

/* A built scene as flat arrays (the SceneView of the reference). Pointers refer to host memory when the
 * descriptor is handed to upload functions or the oracle. Arrays that are absent are NULL with count 0. */
typedef struct ph_scene_desc {
    const ph_bvh_node* nodes;
    int32_t num_nodes;
    const ph_tri_geom* tris;
    int32_t num_tris;
    const int32_t* indices; /* 3 per triangle, BVH order */
    const float* normals;   /* 3 floats per vertex, or NULL */
    const float* uvs;       /* 2 floats per vertex, or NULL */
    const float* tangents;  /* 4 floats per vertex, or NULL */
    int32_t num_vertices;

    const ph_material* materials;
    int32_t num_materials;
    const ph_texture_info* textures;
    int32_t num_textures;
    const uint32_t* texels;
    int64_t num_texels;
    const float* srgb_lut; /* 256 entries */

    const ph_emissive_tri* emissive;
    const int32_t* tri_emissive; /* per triangle, -1 if not emissive */
    const ph_alias_entry* emissive_alias;
    int32_t num_emissive;
    float emissive_power;
    const ph_punctual_light* punctual;
    int32_t num_punctual;

    /* Environment (EnvironmentView, scene_view.h:92-104). */
    const float* env_texels;      /* 3 floats per texel, or NULL */
    const float* env_marginal;    /* height + 1 */
    const float* env_conditional; /* height * (width + 1) */
    int32_t env_width, env_height;
    float env_intensity;
    float env_rotation;
    float env_color[3];
    float env_marginal_integral;

    float p_env, p_emissive, p_punctual;
    float bounds_min[3], bounds_max[3];
} ph_scene_desc;

#ifdef __cplusplus
}
#endif

#endif /* PHOTON_TYPES_H */
