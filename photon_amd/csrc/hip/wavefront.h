// The wavefront path tracing engine: paths in flight as structure-of-arrays state in HBM, advanced
// one bounce at a time by persistent traversal and shading kernels that consume compacted queues.
// It renders one job to completion (the progressive renderer) or serves many jobs at once
// (continuous batching: jobs are admitted between bounces and share the same wavefronts).
#pragma once

#include <memory>
#include <vector>

#include "internal.h"

namespace phk {

// One render request: `spp` samples per pixel with frame indices frame0.., added to `accum` (rgb sum,
// w sum of luminance^2); the G-buffer of the first sample goes to albedo / nd when they are set.
struct WavefrontJob {
    phd::DScene scene;
    ph_camera_view camera;
    int width, height, spp;
    uint32_t frame0, seed;
    phd::PathOpts opt;
    float4* accum;
    float4* albedo;
    float4* nd;
    const int* mat_class;  // per-material class table in device memory (nullptr: computed per render)
    // Serving: the samples of each pixel run as `split` sub-streams of spp / split samples, each added
    // to its own partial image (accum holds split planes of width x height); 1 for the renderer.
    int split = 1;
};

// Per-material class of the shading queues: 0 metal only, 1 opaque only, 2 glass only, 3 mixed.
std::vector<int> material_classes(const phd::DScene& scene, hipStream_t st);

class WavefrontEngine {
public:
    static constexpr int kMaxJobs = 1024;

    WavefrontEngine();
    ~WavefrontEngine();
    WavefrontEngine(const WavefrontEngine&) = delete;
    WavefrontEngine& operator=(const WavefrontEngine&) = delete;

    void set_config(const ph_wavefront_config& cfg);
    ph_wavefront_config config() const;
    // Renders one job to completion (blocking).
    void render(hipStream_t st, const WavefrontJob& job, bool counting, ph_render_stats& stats);
    ph_wavefront_config autotune(hipStream_t st, const WavefrontJob& job);

    // Serving. serve_begin makes `slots` idle paths; serve_admit places a job in a free table entry
    // (returns it, or -1 when the table is full); serve_enqueue queues one pass of an admitted job (every
    // pixel of every sub-stream, samples k0..k1-1 of the sub-stream) and returns false when the pass
    // table is full; serve_run lays out
    // the unstarted items of the queued passes by priority (smaller first, then queue order) and
    // advances every path in flight by `iterations` bounces (blocking). serve_pixels_done counts the
    // finished items (pixel sub-streams) of an entry's passes; a pass must not be queued before the
    // previous pass of the same job has finished. serve_release frees the entry.
    void serve_begin(hipStream_t st, int slots);
    int serve_admit(hipStream_t st, const WavefrontJob& job);
    bool serve_enqueue(int entry, int k0, int k1, double priority);
    // New priority for the queued, not yet started work of an entry (the next serve_run lays it out again).
    void serve_reprioritize(int entry, double priority);
    // Round-robin layout of the running passes' unstarted items, `quantum` items of each per round
    // (processor sharing); 0 lays each pass out in one piece, in priority order.
    void serve_share(int quantum);
    void serve_run(hipStream_t st, int iterations);
    int serve_pixels_done(int entry) const;
    void serve_release(int entry);
    int serve_in_flight() const;  // paths currently traced or shaded
    int64_t serve_work_remaining() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace phk
