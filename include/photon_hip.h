/* The MI355X rendering core: C ABI of the HIP device library (libphoton_hip.so, gfx950).
 *
 * This is the drop-in boundary for the reference's GPU path. The reference's GpuRenderer / Svgf C++ API
 * (src/render/gpu_renderer.h:16-75, src/denoise/svgf.h:24-59) is kept as the header-only wrapper
 * include/photon/gpu_renderer.hpp, whose methods call these entry points one for one:
 *
 *   ph_device_count / ph_device_info  <- cuda_available(GpuInfo*)           gpu_renderer.h:27-34
 *   ph_renderer_create / _destroy     <- GpuRenderer() / ~GpuRenderer()      gpu_renderer.h:38-39
 *   ph_renderer_upload                <- GpuRenderer::upload                 gpu_renderer.h:44
 *   ph_renderer_update_lights         <- GpuRenderer::update_lights          gpu_renderer.h:46
 *   ph_renderer_resize / _reset       <- GpuRenderer::resize / reset         gpu_renderer.h:47-48
 *   ph_renderer_render                <- GpuRenderer::render                 gpu_renderer.h:51
 *   ph_renderer_render_denoised       <- GpuRenderer::render_denoised        gpu_renderer.h:53
 *   ph_renderer_display               <- GpuRenderer::display                gpu_renderer.h:56
 *   ph_renderer_download_image        <- GpuRenderer::download_image         gpu_renderer.h:58
 *   ph_renderer_download_display      <- GpuRenderer::download_display       gpu_renderer.h:59
 *   ph_renderer_width/height/samples/last_frame_ms/display_buffer            gpu_renderer.h:61-65
 *   ph_svgf_*                         <- Svgf::resize / reset / run / variance_buffer  svgf.h:32-59
 * New entry points with no reference counterpart: ray-buffer tracing (traversal workload), engine
 * selection and the autotuner of the wavefront engine, per-kernel statistics, and the serving loop.
 *
 * Conventions: every function returns 0 on success and a negative code on failure, with a message in
 * ph_last_error() for the calling thread (the C++ wrapper turns it into false/error or an exception,
 * as the reference does). Render, display and download calls block until the device work is done,
 * as the reference's do. Pointers named *_dev are device memory; all others are host memory.
 */
#ifndef PHOTON_HIP_H
#define PHOTON_HIP_H

#include "photon_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PH_ENGINE_MEGAKERNEL 0 /* one thread per pixel, whole paths (the reference's structure) */
#define PH_ENGINE_WAVEFRONT 1  /* queue-sorted wavefront with persistent kernels */

typedef struct ph_gpu_info {
    char name[256];
    char arch[64];
    int64_t memory;
    int32_t compute_units;
    int32_t wavefront_size;
    int32_t lds_per_block;
    int32_t clock_khz;
    int32_t memory_clock_khz;
    int32_t memory_bus_width;
} ph_gpu_info;

/* Launch shape of the wavefront engine (chosen by ph_renderer_autotune or set explicitly). */
typedef struct ph_wavefront_config {
    int32_t block;         /* threads per block of the traversal kernels: 64, 128, 256 */
    int32_t stack_depth;   /* LDS stack entries per lane: 8, 16, 24, 32, 64 (rest spills to scratch) */
    int32_t path_slots;    /* paths in flight (0: one per pixel, capped by memory) */
    int32_t blocks_per_cu; /* persistent grid: blocks per compute unit */
    int32_t sort_materials;/* shade in material-class queues */
    int32_t use_graph;     /* replay the bounce loop from a captured HIP graph */
    int32_t graph_iterations; /* bounces per captured graph */
    int32_t refill_lanes;  /* a traversal wave takes new rays once this many of its 64 lanes are free */
    int32_t lds_nodes;     /* top BVH nodes (breadth-first) each traversal block stages in LDS */
    int32_t ray_bins;      /* trace camera rays, then continuing rays by direction octant (1), or all together (0) */
    int32_t xcd_parts;     /* 8: each XCD takes a contiguous eighth of every queue (L2 locality); 1: all waves interleave */
    int32_t sample_split;  /* 1: a path slot runs all of a pixel's samples in order, and the accumulation is bit-identical
                              to the per-pixel kernel's (the reference's order). S > 1: each pixel's samples run as S
                              sub-streams of spp/S samples (S rounded down to a divisor of spp) that any slot may take,
                              into partial sums that are added to the accumulation in sub-stream order; deterministic,
                              within rounding of the in-order sum, and free of the tail of pixels with long paths.
                              With S = spp every partial sum is one sample and the accumulation is the in-order one,
                              bit for bit */
    int32_t finish_percent;/* 0-100. Once every work item of a render is handed out and at most this percentage of the
                              path slots still holds a path, one kernel runs each remaining path to the end of its work
                              item in its own thread (the per-pixel kernel's order, so the image is unchanged) instead of
                              draining them over many sparse bounces. 0: off. Not used when serving */
} ph_wavefront_config;

/* Statistics of the last render call. After render_denoised, ms_shade is the one-sample G-buffer trace
 * and ms_other the SVGF passes. Per-kernel times are measured with HIP events when the bounce loop is
 * launched directly; a graph replay reports only ms_total. */
typedef struct ph_render_stats {
    double ms_total;          /* wall time of the call */
    double ms_extend;         /* summed time of closest-hit traversal launches (HIP events) */
    double ms_shadow;         /* summed time of any-hit traversal launches */
    double ms_shade;          /* summed time of shading launches */
    double ms_other;          /* generation, accumulation, bookkeeping */
    int64_t extend_launches, shadow_launches, shade_launches, iterations;
    int64_t rays_extend, rays_shadow; /* rays traced */
    int64_t samples;
    /* Filled only when counting is enabled: closest-hit (extension) and any-hit (shadow) traversal work. */
    int64_t node_visits, tri_tests, alpha_tests;
    int64_t shadow_node_visits, shadow_tri_tests, shadow_alpha_tests;
    /* Wavefront engine: paths finished one per thread at the end of the render (finish_percent), and the
     * extension and shadow rays they traced (not in rays_extend / rays_shadow, nor in the counts above). */
    int64_t finish_paths, rays_finish;
} ph_render_stats;

/* The five G-buffer planes the reference denoiser consumes (SvgfInput, svgf.h:24-30), as device
 * pointers to 4 floats per pixel. */
typedef struct ph_svgf_input {
    const float* color_dev;
    const float* albedo_dev;
    const float* normal_depth_dev; /* xyz normal, w distance (0 for misses) */
    const float* position_dev;     /* xyz position, w material id (-1 for misses) */
    const float* emission_dev;
} ph_svgf_input;

/* The compact G-buffer of the real-time workload, 48 bytes per pixel against the 80 of ph_svgf_input.
 * No counterpart in the reference, whose denoiser reads only the five float4 planes; a frame in this
 * format is denoised exactly as its decoded planes would be (ph_svgf_run_compact). Device pointers:
 *   color, albedo, emission: 3 floats per pixel (the sample's radiance, the primary albedo, the
 *                            primary surface's emission);
 *   depth:    float per pixel, the primary hit distance (0 where the primary ray missed);
 *   normal:   uint32 per pixel, the primary shading normal in octahedral form, two IEEE half floats
 *             (x in the low 16 bits), rounded to nearest even;
 *   material: int32 per pixel, the primary material id (-1 where the primary ray missed).
 * Positions are not stored: the denoiser rebuilds them at `depth` along each pixel's centre ray from
 * the camera's origin. */
typedef struct ph_compact_gbuffer {
    void* color_dev;
    void* albedo_dev;
    void* emission_dev;
    void* depth_dev;
    void* normal_dev;
    void* material_dev;
} ph_compact_gbuffer;

const char* ph_last_error(void);
int ph_device_count(void);
int ph_device_info(int device, ph_gpu_info* out);

/* Binned-SAH BVH build on the device. No counterpart in the reference, which builds on the CPU
 * (src/bvh/bvh.cpp:51-255); the result is that builder's, bit for bit: the same node array, the same
 * triangle order (BVH position -> input triangle) and the same statistics, so a scene built either way
 * traces identically. Inputs in host memory: positions (3 floats per vertex), indices (3 per triangle,
 * each in [0, num_vertices)). bins (0: 32, at most 64), max_leaf (0: 8) and traversal_cost (0: 1) are
 * BvhBuildOptions (src/bvh/bvh.h:32-37). Outputs in host memory: nodes (room for node_capacity,
 * at least max(1, 2 num_tris - 1)) and order (num_tris). Returns the node count, or -1.
 * At most PH_BVH_BUILD_MAX_TRIS triangles: larger inputs fail with an error before any device work
 * (a 1M-triangle build made a GPU memory fault whose cause is not yet found). */
#define PH_BVH_BUILD_MAX_TRIS 262144
typedef struct ph_bvh_build_stats {
    int32_t nodes, leaves, max_depth;
    float sah_cost;
    float bounds_min[3], bounds_max[3];
    double build_ms; /* wall time of the call, uploads and downloads included */
} ph_bvh_build_stats;
int ph_bvh_build(const float* positions, int num_vertices, const int32_t* indices, int num_tris, int bins, int max_leaf, float traversal_cost,
                 ph_bvh_node* nodes, int node_capacity, int32_t* order, ph_bvh_build_stats* stats);

/* Device memory helpers for callers without a HIP toolchain of their own (bindings, tests). */
int ph_device_alloc(int64_t bytes, void** out_dev);
int ph_device_free(void* dev);
int ph_copy_to_device(void* dst_dev, const void* src, int64_t bytes);
int ph_copy_to_host(void* dst, const void* src_dev, int64_t bytes);
int ph_copy_device(void* dst_dev, const void* src_dev, int64_t bytes);
int ph_device_memset(void* dev, int value, int64_t bytes);
/* Pinned host memory (hipHostMalloc), which kernels can write directly: ph_renderer_display may take
 * such a buffer as its target, so that the tonemapped frame lands in host memory without a separate
 * device-to-host copy (the reference's viewer copies the display buffer to the host each frame, or
 * registers it with GL interop: gpu_renderer.cu:129-175). No counterpart in the reference. */
int ph_host_alloc(int64_t bytes, void** out_host);
int ph_host_free(void* host);
/* What kind of memory a pointer is, as kernels see it: PH_POINTER_DEVICE, PH_POINTER_PINNED_HOST,
 * PH_POINTER_MANAGED, or PH_POINTER_UNUSABLE (NULL, pageable host memory, or unknown). Entry points
 * that take caller targets refuse unusable pointers with this test before any launch. No counterpart in
 * the reference (its display target is documented as device memory only, gpu_renderer.h:56). */
enum { PH_POINTER_UNUSABLE = 0, PH_POINTER_DEVICE = 1, PH_POINTER_PINNED_HOST = 2, PH_POINTER_MANAGED = 3 };
int ph_pointer_kind(const void* p, int* kind);

typedef struct ph_renderer ph_renderer;
int ph_renderer_create(ph_renderer** out);
void ph_renderer_destroy(ph_renderer* r);
int ph_renderer_upload(ph_renderer* r, const ph_scene_desc* scene);
int ph_renderer_update_lights(ph_renderer* r, const ph_scene_desc* scene);
int ph_renderer_resize(ph_renderer* r, int width, int height);
int ph_renderer_reset(ph_renderer* r);
int ph_renderer_render(ph_renderer* r, const ph_camera_view* camera, const ph_render_settings* settings, int spp);
int ph_renderer_render_denoised(ph_renderer* r, const ph_camera_view* camera, const ph_camera_view* previous, const ph_render_settings* settings,
                                const ph_svgf_settings* svgf);
/* target: NULL (the renderer's display buffer), caller device memory, or pinned host memory from ph_host_alloc.
 * Any other target (pageable host memory) fails with a message before a launch, and so does a device
 * allocation smaller than width*height*4 bytes. */
int ph_renderer_display(ph_renderer* r, const ph_render_settings* settings, int mode, uint32_t* target_dev);
int ph_renderer_download_image(ph_renderer* r, float* rgb);
int ph_renderer_download_display(ph_renderer* r, uint32_t* rgba);
int ph_renderer_width(const ph_renderer* r);
int ph_renderer_height(const ph_renderer* r);
int ph_renderer_samples(const ph_renderer* r);
float ph_renderer_last_frame_ms(const ph_renderer* r);
uint32_t* ph_renderer_display_buffer(const ph_renderer* r);
/* One sample per pixel into a caller-owned compact G-buffer: the sample render_denoised traces (the
 * reference's sample_kernel, gpu_renderer.cu:83-95), written in the compact format. Uses and advances
 * the renderer's frame counter like render_denoised; blocks until done (last_frame_ms is its time). */
int ph_renderer_sample_compact(ph_renderer* r, const ph_camera_view* camera, const ph_render_settings* settings, const ph_compact_gbuffer* out);

/* Engine selection, tuning and statistics. */
int ph_renderer_set_engine(ph_renderer* r, int engine);
int ph_renderer_get_engine(const ph_renderer* r);
int ph_renderer_set_wavefront_config(ph_renderer* r, const ph_wavefront_config* cfg);
int ph_renderer_get_wavefront_config(const ph_renderer* r, ph_wavefront_config* cfg);
/* Times candidate launch shapes on a short render of the given view and keeps the fastest. */
int ph_renderer_autotune(ph_renderer* r, const ph_camera_view* camera, const ph_render_settings* settings, int spp, ph_wavefront_config* chosen);
int ph_renderer_set_counting(ph_renderer* r, int enable);
int ph_renderer_stats(const ph_renderer* r, ph_render_stats* out);
/* Device pointers of the frame buffers (4 floats per pixel): 0 accum (rgb sum, w sum of luminance^2),
 * 1 albedo, 2 normal+depth, 3 colour, 4 position+material, 5 emission, 6 denoised; 7 is the heat map
 * (1 float per pixel: BVH nodes visited by the pixel-centre primary ray); 8 is the denoiser's variance
 * plane (.w) that the Variance view of a denoised frame shows (svgf.h:55 variance_buffer). */
const float* ph_renderer_buffer(const ph_renderer* r, int which);

/* Traversal of a ray buffer against the uploaded scene: closest hit (t, tri, u, v) or any hit
 * (t, tri). `counts_dev` (optional, 3 uint32 per ray): nodes visited, triangles tested, alpha tests. */
int ph_trace(ph_renderer* r, const ph_ray* rays_dev, int n, int anyhit, ph_hit* hits_dev, uint32_t* counts_dev);
/* The same as ph_trace with an explicit launch shape (block 64/128/256, stack depth 8..64; a persistent
 * grid stages the top lds_nodes BVH nodes in each block's LDS), timed with HIP events on the library's
 * stream; ms may be NULL. */
int ph_trace_config(ph_renderer* r, const ph_ray* rays_dev, int n, int anyhit, ph_hit* hits_dev, int block, int stack_depth, int persistent,
                    int lds_nodes, float* ms);

/* SVGF on caller G-buffers (the reference's Svgf class). */
typedef struct ph_svgf ph_svgf;
int ph_svgf_create(ph_svgf** out);
void ph_svgf_destroy(ph_svgf* f);
int ph_svgf_resize(ph_svgf* f, int width, int height);
int ph_svgf_reset(ph_svgf* f);
int ph_svgf_run(ph_svgf* f, const ph_svgf_input* in, const ph_camera_view* previous_camera, const ph_svgf_settings* settings, float* output_dev);
/* The denoiser on a compact G-buffer traced from view `camera` (no counterpart in the reference). The
 * previous frame is kept in the format of the call that stored it; switching formats starts a new
 * history. */
int ph_svgf_run_compact(ph_svgf* f, const ph_compact_gbuffer* in, const ph_camera_view* camera, const ph_camera_view* previous_camera,
                        const ph_svgf_settings* settings, float* output_dev);
/* ph_svgf_run without an allocation: no counterpart in the reference. For callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid allocation). It runs
 * the same nine passes as ph_svgf_run on one frame of at most PH_SVGF_STATIC_MAX_PIXELS pixels (larger
 * frames fail with a message). The frame-local planes (illumination, moments, two a-trous planes: 64
 * bytes per pixel) live in one static device array that each GPU the library runs on holds until the
 * process ends (PH_SVGF_STATIC_MAX_PIXELS * 64 bytes = 92 MB). The history that ph_svgf keeps inside
 * the object is the caller's instead: `history_dev` holds four float4 planes of width*height pixels
 * (illumination + history length, moments, normal+depth, position+material), is read when
 * `has_history` is nonzero (0 for the first frame, as after ph_svgf_reset) and is rewritten by the call.
 * Runs on the default stream and returns when the frame is done. Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (from several host threads) race and must be serialised by the
 * caller. */
#define PH_SVGF_STATIC_MAX_PIXELS (1600 * 900)
int ph_svgf_run_static(const ph_svgf_input* in, int width, int height, const ph_camera_view* previous_camera, const ph_svgf_settings* settings,
                       float* history_dev, int has_history, float* output_dev);
const float* ph_svgf_variance_buffer(const ph_svgf* f);
float ph_svgf_last_ms(const ph_svgf* f);

/* Serving: several scenes stay resident; render jobs of mixed resolution, sample count and depth are
 * admitted by HBM footprint while others run, and their paths share the same wavefronts (continuous
 * batching). A job's image depends only on the job and the server's configuration (how small jobs are
 * split, which lane runs them), so it is identical whether the job runs alone or among others; with
 * min_items = 1 it is bit-identical to the renderer's.
 * No counterpart in the reference. */
typedef struct ph_server ph_server;

typedef struct ph_server_config {
    int64_t memory_budget;       /* bytes of job frame buffers admitted at once (0: 16 GiB) */
    int32_t path_slots;          /* paths in flight shared by all jobs (0: 2^21) */
    int32_t iterations_per_step; /* bounces per ph_server_step (0: 4) */
    int32_t max_jobs;            /* jobs admitted at once, at most 1024 (0: 1024) */
    int64_t pass_samples;        /* pixel samples per pass of a job (0: 4 x path_slots) */
    int32_t schedule;            /* PH_SCHEDULE_SRPT, _FIFO, _HRRN, _SHARE or _DEADLINE */
    int32_t min_items;           /* a job with fewer pixels runs each pixel's samples as sub-streams so that
                                    pixels x sub-streams reaches this (0: an eighth of the main lane's slots,
                                    all of the express lane's; -1 or 1: never split) */
    int64_t express_samples;     /* jobs of at most this many samples run in the express lane
                                    (0: 16 x express_slots) */
    int32_t express_slots;       /* path slots of the express lane, a second engine on its own stream that
                                    runs concurrently with the main lane (0: path_slots / 2; -1: no lane) */
    int32_t express_iterations;  /* bounces per run of the express lane (0: iterations_per_step) */
    double deadline_stretch;     /* PH_SCHEDULE_DEADLINE: a job's deadline is its submission time plus this
                                    many times its estimated service time (0: 4) */
} ph_server_config;

#define PH_SCHEDULE_SRPT 0 /* unstarted work of the job with the least remaining work first */
#define PH_SCHEDULE_FIFO 1 /* passes in the order they were queued (round robin over running jobs) */
#define PH_SCHEDULE_HRRN 2 /* highest response ratio next: (waiting time + estimated remaining service) / estimated
                              remaining service, re-evaluated every step; evens out slowdown across job sizes */
#define PH_SCHEDULE_SHARE 3 /* processor sharing: the running jobs' unstarted work items are handed out in
                               round-robin slices, so every running job gets an equal share of the slots that
                               free up (a job that cannot use its share leaves it to the others) */
#define PH_SCHEDULE_DEADLINE 4 /* earliest deadline first, with deadline = submission time + deadline_stretch x
                                  the job's nominal service time (its samples at 1e6 samples per ms): a job
                                  waits at most about that multiple of its own size behind newer work */

typedef struct ph_job_desc {
    int32_t scene; /* id returned by ph_server_add_scene */
    int32_t width, height, spp;
    ph_camera_view camera;
    ph_path_options path;
    uint32_t seed, frame0;
} ph_job_desc;

#define PH_JOB_QUEUED 0
#define PH_JOB_RUNNING 1
#define PH_JOB_DONE 2
#define PH_JOB_RELEASED 3

typedef struct ph_job_status {
    int32_t state;
    int32_t pixels;
    int64_t samples_done, samples;       /* pixel samples finished (as of the last step) / requested */
    double submit_ms, admit_ms, done_ms; /* on the server's clock (ph_server_now_ms) */
} ph_job_status;

int ph_server_create(const ph_server_config* cfg, ph_server** out);
void ph_server_destroy(ph_server* s);
int ph_server_set_wavefront_config(ph_server* s, const ph_wavefront_config* cfg);
/* Uploads a scene and keeps it resident; returns its id. */
int ph_server_add_scene(ph_server* s, const ph_scene_desc* scene);
/* Queues a job (admitted by a later step); returns its id. */
int ph_server_submit(ph_server* s, const ph_job_desc* job);
/* Admits queued jobs that fit, advances all paths in flight by iterations_per_step bounces (blocking),
 * queues the next pass of every job whose pass finished, and returns the number of jobs finished since
 * the server started. A job runs as passes of at most pass_samples pixel samples; unstarted work is
 * handed to path slots in the order of the schedule. While the main lane runs, the express lane runs
 * as many times as it has work, so its jobs can finish within one step. */
int ph_server_step(ph_server* s);
int ph_server_status(const ph_server* s, int job, ph_job_status* out);
/* Average radiance of a finished job (3 floats per pixel); frees its frame buffer. */
int ph_server_result(ph_server* s, int job, float* rgb);
/* Frees a finished job's frame buffer without reading it. */
int ph_server_release(ph_server* s, int job);
/* Jobs queued or running. */
int ph_server_pending(const ph_server* s);
double ph_server_now_ms(const ph_server* s);

#ifdef __cplusplus
}
#endif

#endif /* PHOTON_HIP_H */
