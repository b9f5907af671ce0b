/* ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load liboracle.so; the product (photon_amd/) never does.
 *
 * C ABI of the CPU restatement of the reference renderer (oracle_render.hpp) and of its SVGF denoiser
 * (svgf_oracle.cpp). All buffers are host memory.
 */
#ifndef PHOTON_ORACLE_H
#define PHOTON_ORACLE_H

#include "photon_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* CpuRenderer::render (src/render/cpu_renderer.cpp:20-39): adds `spp` samples with frame indices
 * frame0 .. frame0+spp-1 to `sum` (3 floats per pixel, row-major), `threads` row workers
 * (0: hardware concurrency). Optional per-pixel G-buffer of the first sample (albedo, normal+depth as
 * 4 floats each) like render_kernel's k == 0 outputs (src/render/gpu_renderer.cu:60-80). */
int or_render(const ph_scene_desc* s, const ph_camera_view* cam, const ph_render_settings* rs, int spp, uint32_t frame0, int threads, float* sum,
              float* albedo4, float* normal_depth4);

/* Radiance of single samples for a list of (x, y, frame), and the rays each path traced. */
int or_render_samples(const ph_scene_desc* s, const ph_camera_view* cam, const ph_render_settings* rs, int n, const int32_t* xs, const int32_t* ys,
                      const uint32_t* frames, float* rgb, uint32_t* ray_counts);

/* sample_kernel (src/render/gpu_renderer.cu:83-95): one sample per pixel plus the G-buffer, as five
 * 4-float-per-pixel planes (color, albedo, normal+depth, position+material, emission). */
int or_render_gbuffer(const ph_scene_desc* s, const ph_camera_view* cam, const ph_render_settings* rs, uint32_t frame, int threads, float* color4,
                      float* albedo4, float* normal_depth4, float* position4, float* emission4);

/* intersect / occluded (src/render/traverse.h:140-147) over a ray list. Any-hit sets only t and tri.
 * counters (optional, 3 per ray): nodes visited, triangles tested, alpha tests. */
int or_trace(const ph_scene_desc* s, const ph_ray* rays, int n, int anyhit, ph_hit* hits, int64_t* counters, int threads);

/* heatmap_kernel's node-visit count for pixel-centre primary rays (gpu_renderer.cu:98-106). */
int or_heatmap(const ph_scene_desc* s, const ph_camera_view* cam, int width, int height, float* visits);

/* display_kernel for the Color mode of a progressive image (gpu_renderer.cu:129-175). */
int or_display_color(const float* sum, int n, int samples, float exposure, int tonemap, uint32_t* out);

/* display_kernel, every mode (0 colour, 1 albedo, 2 normal, 3 depth, 4 variance, 5 heat map), over
 * the renderer's planes: accum = (sum of radiance, sum of luminance^2), denoised colour, albedo,
 * normal+depth, the denoiser's variance plane (.w) and the node-visit counts (gpu_renderer.cu:108-175). */
int or_display(int n, int mode, float exposure, int tonemap, int samples, float max_depth, float heat_scale, int denoised,
               const float* accum4, const float* denoised4, const float* albedo4, const float* nd4, const float* variance4,
               const float* heat, uint32_t* out);

/* Function-level probes (batch form) for the statistical tests. */
int or_rng_u32(uint64_t seed, uint64_t stream, int n, uint32_t* out);
int or_rng_float(uint64_t seed, uint64_t stream, int n, float* out);
int or_hash_u32(const uint32_t* in, int n, uint32_t* out);
/* kind: 0 concentric disk (2 out), 1 cosine hemisphere (3), 2 uniform sphere (3), 3 uniform triangle (2),
 * 4 direction->equirect (in 3, out 2), 5 equirect->direction (in 2, out 3). */
int or_warp(int kind, int n, const float* in, float* out);
int or_frame(int n, const float* normals, const float* v, float* local, float* world_of_local, float* basis9);
int or_offset_ray(int n, const float* p, const float* nrm, float* out);
/* BSDF with (base rgb, metallic, roughness, transmission, ior) = params[7]. u = 4 floats per sample. */
int or_bsdf_sample(const float* params, int n, const float* wo, const float* u, float* wi, float* f, float* pdf, int32_t* flags);
int or_bsdf_eval(const float* params, int n, const float* wo, const float* wi, float* f, float* pdf);
int or_bsdf_info(const float* params, float* out6); /* alpha, w_metal, w_opaque, w_glass, albedo-luminance, samples_lights */
/* sample_light at points p with u = 4 floats (group, 2D, light); out 12 floats per sample:
 * wi(3) dist radiance(3) pdf delta tri bary(2). */
int or_sample_light(const ph_scene_desc* s, int n, const float* p, const float* u, float* out);
int or_env_eval(const ph_scene_desc* s, int n, const float* dirs, float* radiance3, float* pdf_dir);
int or_env_sample(const ph_scene_desc* s, int n, const float* u, float* dirs, float* pdf, int32_t* ok);
int or_emissive_pdf(const ph_scene_desc* s, int n, const int32_t* tri, const float* dirs, const float* dist, float* pdf);
/* surface_at emission for hits (for the hemisphere-vs-light test). */
int or_hit_emission(const ph_scene_desc* s, int n, const ph_ray* rays, const ph_hit* hits, float* emission3);
int or_tonemap(int n, const float* rgb, float exposure, int op, float* out, uint32_t* packed);
int or_srgb(int n, const float* in, int encode, float* out);

/* SVGF restatement (src/denoise/svgf.cu:37-255) with the reference's five vec4 G-buffer planes. */
typedef struct or_svgf or_svgf;
int or_svgf_create(int width, int height, or_svgf** out);
void or_svgf_destroy(or_svgf* f);
void or_svgf_reset(or_svgf* f);
int or_svgf_run(or_svgf* f, const float* color4, const float* albedo4, const float* normal_depth4, const float* position4, const float* emission4,
                const ph_camera_view* prev_cam, const ph_svgf_settings* st, float* out4);
/* Internal planes for tests: 0 prev_illum, 1 prev_moments, 2 prev_nd, 3 prev_pos, 4 illum, 5 moments,
 * 6 filter0, 7 filter1. */
const float* or_svgf_plane(const or_svgf* f, int which);
/* One a-trous pass alone (step 1, 2, 4, ...) on w*h float4 planes, for tests. */
int or_svgf_atrous(or_svgf* f, const float* normal_depth4, const float* position4, int step, float phi_color, float phi_normal, float phi_depth,
                   const float* src4, float* dst4);

/* The compact G-buffer of the real-time workload: fp16 conversion, octahedral normals, conversion from
 * the float4 planes, and the denoiser on a compact frame traced from `cam`. */
uint32_t or_f32_to_f16(float f);
float or_f16_to_f32(uint32_t h);
uint32_t or_oct_encode(const float* n);
void or_oct_decode(uint32_t code, float* n);
int or_gbuffer_compact(int n, const float* color4, const float* albedo4, const float* nd4, const float* pos4, const float* emission4, float* color3,
                       float* albedo3, float* emission3, float* depth, uint32_t* normal, int32_t* material);
int or_svgf_run_compact(or_svgf* f, const float* color3, const float* albedo3, const float* emission3, const float* depth, const uint32_t* normal,
                        const int32_t* material, const ph_camera_view* cam, const ph_camera_view* prev_cam, const ph_svgf_settings* st, float* out4);

#ifdef __cplusplus
}
#endif

#endif
